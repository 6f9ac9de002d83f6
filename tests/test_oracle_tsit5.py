"""Pins the oracle's Tsit5 table (oracle.TSIT5_*, Tsitouras 2011 as OrdinaryDiffEq's Tsit5 steps it
with a fixed dt) by what the mathematics fixes for a fifth-order explicit Runge-Kutta method:
- the row sums sum_j a_ij = c_i of every stage;
- the 17 order conditions of the rooted trees through order 5 (Butcher), to round-off of the
  published 16-digit coefficients, and a sixth-order one that fails (the method is of order 5);
- on y' = lambda y, the stability function's Taylor coefficients b^T A^(k-1) 1 = 1 / k! for k <= 5;
- the oracle's loop on a linear ODE system against the matrix exponential converges at order 5.
A wrong digit in any coefficient breaks a row sum or an order condition by far more than 1e-13."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O


def tableau():
    A = np.zeros((7, 7))
    for i, row in enumerate(O.TSIT5_A):
        A[i + 1, : len(row)] = row
    A[6, :6] = O.TSIT5_B  # the seventh stage's inputs are the solution weights (FSAL)
    b = np.zeros(7)
    b[:6] = O.TSIT5_B
    return A, b, np.array(O.TSIT5_C)


def test_row_sums_are_the_nodes():
    A, b, c = tableau()
    assert np.max(np.abs(A.sum(axis=1) - c)) <= 1e-14
    assert abs(b.sum() - 1.0) <= 1e-15


def test_order_conditions_through_order_5():
    A, b, c = tableau()
    Ac, Ac2, Ac3 = A @ c, A @ c**2, A @ c**3
    conds = [
        (b.sum(), 1.0),
        (b @ c, 1 / 2),
        (b @ c**2, 1 / 3), (b @ Ac, 1 / 6),
        (b @ c**3, 1 / 4), (b @ (c * Ac), 1 / 8), (b @ Ac2, 1 / 12), (b @ (A @ Ac), 1 / 24),
        (b @ c**4, 1 / 5), (b @ (c**2 * Ac), 1 / 10), (b @ (c * Ac2), 1 / 15), (b @ (c * (A @ Ac)), 1 / 30),
        (b @ (Ac * Ac), 1 / 20), (b @ Ac3, 1 / 20), (b @ (A @ (c * Ac)), 1 / 40), (b @ (A @ Ac2), 1 / 60),
        (b @ (A @ (A @ Ac)), 1 / 120),
    ]
    assert len(conds) == 17
    for n, (got, want) in enumerate(conds):
        assert abs(got - want) <= 1e-13, (n, got, want)
    # not of order 6: the bushy tree of order 6 fails
    assert abs(b @ c**5 - 1 / 6) > 1e-4


def test_stability_function_taylor_coefficients():
    A, b, _ = tableau()
    one = np.ones(7)
    v = one
    for k in range(1, 6):
        assert abs(b @ v - 1 / math.factorial(k)) <= 1e-13, k
        v = A @ v
    assert abs(b @ v - 1 / math.factorial(6)) > 1e-6  # b^T A^5 1 is free: order 5, not 6


@pytest.mark.parametrize("seed", [0, 1])
def test_loop_converges_at_order_5_on_a_linear_system(seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((6, 6))
    M = M - M.T - 0.3 * np.eye(6)  # oscillatory, slightly damped
    u0 = rng.standard_normal(6)
    exact = expm(M * 1.0) @ u0
    errs = []
    for n in (20, 40, 80):
        got = O.tsit5(lambda v: M @ v, u0, 1.0 / n, n)
        errs.append(np.linalg.norm(got - exact))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(orders) >= 4.8, (orders, errs)
