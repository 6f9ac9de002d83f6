"""filter_exp (src/Polynomial/poly_filter.jl:11-47): the oracle's restatement pinned by what the
exponential filter of Hesthaven & Warburton (KitBase filter_exp1d's published algorithm; KitBase
itself is absent, so parity with its code stays unpinned) fixes -- sigma = 1 up to the cutoff
mode, sigma = eps at the highest mode, sigma = eps^(((d - Nc)/(N - Nc))^s) in between, V's columns
as the eigenvectors of F -- and the reference's own test call (test/runtests.jl:29-33, a 1D
degree-5 Vandermonde matrix with N = 2, which takes the triangle branch); then libfrhip's host
fr_filter_exp against the oracle (CPU only)."""
import ctypes as C

import numpy as np
import pytest

import frhip
from frhip import _lib
from oracle import oracle as O

EPS = 2.220446049250313e-16


@pytest.mark.parametrize("N", [2, 3, 4, 6])
@pytest.mark.parametrize("s", [2, 4, 10])
@pytest.mark.parametrize("Nc", [0, 1])
def test_oracle_filter_exp1d_pins(N, s, Nc):
    d = O.filter_exp1d(N, s, Nc)
    assert np.all(d[: Nc + 1] == 1.0)
    assert abs(d[N] - EPS) <= 1e-12 * EPS
    for i in range(Nc, N + 1):
        assert abs(d[i] - EPS ** (((i - Nc) / (N - Nc)) ** s)) <= 1e-13 * d[i]
    assert np.all(np.diff(d) <= 0.0)
    # a denominator of N instead of N - Nc would leave the highest mode at eps^((N-Nc)/N)^s
    if Nc:
        assert abs(EPS ** (((N - Nc) / N) ** s) - d[N]) > 1e-6 * EPS


@pytest.mark.parametrize("N", [2, 3, 5])
def test_oracle_filter_exp_line_eigen(N):
    ops = O.Ops(N)
    V = ops.V1
    F = O.filter_exp(N, 8, V)
    d = O.filter_exp1d(N, 8)
    assert np.allclose(F @ V, V * d[None, :], atol=1e-13)
    # the constant is kept, the highest Legendre mode removed (to eps)
    assert np.allclose(F @ np.ones(N + 1), np.ones(N + 1), atol=1e-13)
    top = V[:, N]
    assert np.max(np.abs(F @ top)) <= 1e-13


def test_reference_test_call_triangle_branch():
    # FR.filter_exp(2, 10, Array(ps.V)) with ps = FRPSpace1D(0, 1, 20, 5): nv = 6 = (2+1)(2+2)/2
    V = O.Ops(5).V1
    F = O.filter_exp(2, 10, V)
    d2 = O.filter_exp2d(2, 10)
    # modes (i, j): (0,0) (0,1) (0,2) (1,0) (1,1) (2,0) -> total degrees 0 1 2 1 2 2
    deg = np.array([0, 1, 2, 1, 2, 2])
    assert np.allclose(d2, EPS ** ((deg / 2.0) ** 10), rtol=1e-13, atol=0)
    assert np.allclose(F @ V, V * d2[None, :], atol=1e-12)
    # a quad basis ((N+1)^2 rows) has no filter in the reference
    assert O.filter_exp(2, 10, np.eye(9)) is None


@pytest.mark.parametrize("N,s,Nc,nv", [(2, 10, 0, 6), (3, 4, 0, 4), (4, 6, 1, 5), (3, 8, 1, 10), (5, 2, 0, 6)])
@pytest.mark.parametrize("with_inv", [False, True])
def test_library_filter_exp_matches_oracle(N, s, Nc, nv, with_inv):
    rng = np.random.default_rng(N * 10 + nv)
    V = O.Ops(nv - 1).V1 if nv <= 8 else np.eye(nv) + 0.1 * rng.standard_normal((nv, nv))
    iV = np.linalg.inv(V) if with_inv else None
    F = frhip.filter_exp(N, s, V, Nc, iV)
    ref = O.filter_exp(N, s, V, Nc, iV if with_inv else None)
    assert np.max(np.abs(F - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))


def test_library_filter_exp_errors():
    V = np.asfortranarray(np.eye(9))
    F = np.zeros((9, 9), order="F")
    L = frhip.lib()
    from frhip import ops

    ops._bind()
    dp = _lib.dptr
    assert L.fr_filter_exp(2, 10, 0, dp(V), 9, None, dp(F)) == 5  # FR_E_UNSUPPORTED: quad basis
    V3 = np.asfortranarray(np.eye(3))
    assert L.fr_filter_exp(2, 10, 2, dp(V3), 3, None, dp(F)) == 1  # Nc must be below N
    assert b"fr_filter_exp" in L.fr_last_error()
