"""frhip.positive_limiter (the reference's per-element positive_limiter, src/dissipation.jl:37-215,
on libfrhip's native host routines) against the oracle's restatement of the same element routine
(pinned by the limiter invariants in tests/test_oracle_residual.py), on seeded elements with
negative density and negative pressure at points, for the scalar, 1D Euler and 2D Euler methods;
the runtests.jl:24-25 calls; an element with a non-positive mean raises and is left unchanged."""
import numpy as np
import pytest

import frhip
from oracle import oracle as O


def euler1d_element(rng, nsp, g):
    rho = 1.0 + 0.3 * rng.random(nsp)
    u = 0.5 * rng.standard_normal(nsp)
    p = 0.5 + 0.5 * rng.random(nsp)
    w = np.stack([rho, rho * u, p / (g - 1.0) + 0.5 * rho * u * u], -1)
    k = rng.integers(nsp)
    w[k, 0] = -0.2  # negative density at a point
    w[(k + 1) % nsp, 2] *= 0.05  # low energy: negative pressure at a point
    return np.asfortranarray(w)


def euler2d_element(rng, nsp, g):
    rho = 1.0 + 0.3 * rng.random((nsp, nsp))
    u, v = 0.4 * rng.standard_normal((2, nsp, nsp))
    p = 0.5 + 0.5 * rng.random((nsp, nsp))
    w = np.stack([rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)], -1)
    k, l = rng.integers(nsp, size=2)
    w[k, l, 0] = -0.3
    w[(k + 1) % nsp, l, 3] *= 0.02
    return np.asfortranarray(w)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_scalar_method(deg):
    ops = O.Ops(deg)
    rng = np.random.default_rng(deg)
    for _ in range(20):
        u = 1.0 + rng.random(deg + 1)
        u[rng.integers(deg + 1)] = -0.3 * rng.random()
        for weights in (ops.w / 2.0, 1.0 / (deg + 1)):
            rc, ref = O.positive_limiter1d_scalar(u, weights, ops.ll, ops.lr)
            got = u.copy()
            frhip.positive_limiter(got, weights, ops.ll, ops.lr)
            assert rc == 0 and np.allclose(got, ref, rtol=1e-14, atol=1e-15)
            assert not np.array_equal(got, u)  # the negative point was lifted


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 7])
def test_euler1d_method(deg):
    g = 1.4
    ops = O.Ops(deg)
    rng = np.random.default_rng(10 + deg)
    for _ in range(20):
        ue = euler1d_element(rng, deg + 1, g)
        rc, ref = O.positive_limiter1d_euler(ue, g, ops.w / 2.0, ops.ll, ops.lr)
        got = ue.copy(order="F")
        frhip.positive_limiter(got, g, ops.w / 2.0, ops.ll, ops.lr)
        assert rc == 0
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert not np.array_equal(got, ue)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_euler2d_method(deg):
    g = 1.4
    ops = O.Ops(deg)
    wp4 = np.asfortranarray(np.outer(ops.w, ops.w) / 4.0)
    rng = np.random.default_rng(20 + deg)
    for _ in range(20):
        ue = euler2d_element(rng, deg + 1, g)
        rc, ref = O.positive_limiter2d(ue, g, wp4, ops.ll, ops.lr)
        got = ue.copy(order="F")
        frhip.positive_limiter(got, g, wp4, ops.ll, ops.lr)
        assert rc == 0
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        # every point and face trace is admissible afterwards: rho > 0 and p >= eps up to round-off
        p = (g - 1.0) * (got[..., 3] - 0.5 * (got[..., 1] ** 2 + got[..., 2] ** 2) / got[..., 0])
        assert got[..., 0].min() > 0.0 and p.min() > -1e-12


def test_admissible_elements_are_returned_unchanged_and_runtests_calls():
    ops = O.Ops(5)
    u = np.ones(6)
    frhip.positive_limiter(u, 1.0 / 6.0, ops.ll, ops.lr)  # runtests.jl:24
    assert np.array_equal(u, np.ones(6))
    ue = np.ones((6, 3), order="F")
    frhip.positive_limiter(ue, 5.0 / 3.0, np.full(6, 1.0 / 6.0), ops.ll, ops.lr)  # runtests.jl:25
    assert np.array_equal(ue, np.ones((6, 3)))
    # a smooth element well inside the admissible set, at its points and its face traces: the
    # reference's density update t1 (u - um) + um runs with t1 = 1, which only rounds
    o3 = O.Ops(3)
    k, l = np.meshgrid(o3.r, o3.r, indexing="ij")
    w2 = np.asfortranarray(np.stack([1.2 + 0.05 * k, 0.1 + 0.02 * l, -0.05 + 0.01 * k * l, 3.0 + 0.1 * l], -1))
    keep = w2.copy(order="F")
    frhip.positive_limiter(w2, 1.4, np.outer(o3.w, o3.w) / 4.0, o3.ll, o3.lr)
    assert np.max(np.abs(w2 - keep)) <= 1e-15 * np.max(np.abs(keep))
    assert np.array_equal(w2[..., 1:], keep[..., 1:])


def test_non_positive_mean_raises_and_keeps_the_element():
    ops = O.Ops(3)
    ue = np.asfortranarray(np.tile([-1.0, 0.0, 2.0], (4, 1)))
    keep = ue.copy(order="F")
    with pytest.raises(frhip.FRError):
        frhip.positive_limiter(ue, 1.4, ops.w / 2.0, ops.ll, ops.lr)
    assert np.array_equal(ue, keep)
    u = -np.ones(4)
    with pytest.raises(frhip.FRError):
        frhip.positive_limiter(u, ops.w / 2.0, ops.ll, ops.lr)
    assert np.array_equal(u, -np.ones(4))


def test_modal_filter_l2():
    """modal_filter (KitBase modal_filter! :l2 as the scripts call it) against the oracle's
    restatement, and what the form fixes: mode 0 kept, lambda = 0 the identity, every higher mode
    damped by 1 / (1 + lambda (k+1)^2 k^2), the factor falling with k."""
    rng = np.random.default_rng(5)
    for n in (3, 9, 16, 36):
        u = rng.standard_normal(n)
        for lam in (0.0, 5e-4, 0.3):
            got = u.copy()
            frhip.modal_filter(got, lam)
            assert np.allclose(got, O.modal_filter_l2(u, lam), rtol=1e-15, atol=0)
            assert got[0] == u[0]
            if lam == 0.0:
                assert np.array_equal(got, u)
        k = np.arange(n)
        ratio = np.ones(n)
        frhip.modal_filter(ratio, 0.3)
        assert np.allclose(ratio, 1.0 / (1.0 + 0.3 * ((k + 1) * k) ** 2), rtol=1e-15, atol=0)
        assert np.all(np.diff(ratio) < 0)
    with pytest.raises(ValueError):
        frhip.modal_filter(np.ones(4), 1e-3, filter="exp")
