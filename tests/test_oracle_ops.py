"""Pins the oracle's operator data (src/Polynomial, src/Transform, src/struct.jl) against
library routines (numpy.polynomial), closed forms and the reference's own Vandermonde identity
(example/vandermonde_lagrange.jl:14-25).  Also checks libfrhip's native host builders (the
product's operator data) against the pinned oracle."""
import numpy as np
import pytest
from numpy.polynomial import legendre as L

import frhip
from oracle import oracle as O


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_legendre_vs_numpy(n):
    x, w = O.gauss_legendre(n)
    xn, wn = L.leggauss(n)
    assert np.allclose(x, xn, rtol=0, atol=2e-16 * n)
    assert np.allclose(w, wn, rtol=0, atol=1e-15)
    # exact for x^(2n-1) and x^(2n-2)
    assert abs(np.sum(w * x ** (2 * n - 2)) - 2.0 / (2 * n - 1)) < 1e-14


def test_gauss_closed_form_p3():
    x, _ = O.gauss_legendre(4)
    a = np.sqrt(3 / 7 - 2 / 7 * np.sqrt(6 / 5))
    b = np.sqrt(3 / 7 + 2 / 7 * np.sqrt(6 / 5))
    assert np.allclose(x, [-b, -a, a, b], atol=1e-16)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_lagrange_invariants(deg):
    ops = O.Ops(deg)
    r = ops.r
    assert abs(ops.ll.sum() - 1) < 1e-14 and abs(ops.lr.sum() - 1) < 1e-14
    # cardinal property and reproduction of polynomials of degree <= p at the end points
    for q in range(deg + 1):
        assert abs(np.dot(r ** q, ops.ll) - (-1.0) ** q) < 1e-13
        assert abs(np.dot(r ** q, ops.lr) - 1.0) < 1e-13
        # dl exact on x^q: d/dx x^q at the points
        d = ops.dl @ (r ** q)
        exact = q * r ** (q - 1) if q > 0 else np.zeros_like(r)
        assert np.allclose(d, exact, atol=1e-12)
    assert np.allclose(ops.dl.sum(axis=1), 0.0, atol=1e-13)


def test_p3_known_values():
    ops = O.Ops(3)
    assert np.allclose(ops.ll, [1.52678813, -0.81363245, 0.40076152, -0.11391720], atol=5e-9)
    assert np.allclose(ops.dhr, [-0.32748486, 0.61452810, -1.24762477, 4.38915297], atol=5e-9)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_radau_closed_form(deg):
    ops = O.Ops(deg)
    r, w = ops.r, ops.w
    cp = np.zeros(deg + 2)
    cp[deg] = 1
    cp1 = np.zeros(deg + 2)
    cp1[deg + 1] = 1
    dgr = 0.5 * (L.legval(r, L.legder(cp)) + L.legval(r, L.legder(cp1)))
    dgl = (-1) ** deg * 0.5 * (L.legval(r, L.legder(cp)) - L.legval(r, L.legder(cp1)))
    assert np.allclose(ops.dhr, dgr, atol=1e-12) and np.allclose(ops.dhl, dgl, atol=1e-12)
    # g_R(1) - g_R(-1) = 1 and g_L(1) - g_L(-1) = -1 through Gauss quadrature of the derivatives
    assert abs(np.dot(w, ops.dhr) - 1.0) < 1e-13 and abs(np.dot(w, ops.dhl) + 1.0) < 1e-13
    # mirror symmetry dhl(x_k) = -dhr(x_{p-k})
    assert np.allclose(ops.dhl, -ops.dhr[::-1], atol=1e-13)


@pytest.mark.parametrize("kind", ["sd", "huynh"])
def test_sd_huynh_closed_form(kind):
    deg = 3
    r, _ = O.gauss_legendre(deg + 1)
    dgl, dgr = O.dcorrection(kind, deg, r)

    def dP(n):
        c = np.zeros(n + 1)
        c[n] = 1
        return L.legval(r, L.legder(c))

    if kind == "sd":
        y = (deg * dP(deg - 1) + (deg + 1) * dP(deg + 1)) / (2 * deg + 1)
    else:
        y = ((deg + 1) * dP(deg - 1) + deg * dP(deg + 1)) / (2 * deg + 1)
    assert np.allclose(dgr, 0.5 * (dP(deg) + y), atol=1e-12)
    assert np.allclose(dgl, (-1) ** deg * 0.5 * (dP(deg) - y), atol=1e-12)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_vandermonde_orthonormal_and_reference_identity(deg):
    ops = O.Ops(deg)
    n = deg + 1
    # orthonormal Legendre basis: V^T diag(w) V = I (1D) and the tensor version (quad)
    assert np.allclose(ops.V1.T @ np.diag(ops.w) @ ops.V1, np.eye(n), atol=1e-13)
    W = np.diag(ops.wp.ravel(order="F"))
    assert np.allclose(ops.V.T @ W @ ops.V, np.eye(n * n), atol=1e-13)
    assert np.allclose(ops.iV @ ops.V, np.eye(n * n), atol=1e-12)
    # example/vandermonde_lagrange.jl:14-25: V' \ psi_f = ll / lr and V' \ Vr = dl
    psi = O.vandermonde_1d(deg, np.array([-1.0, 1.0]))
    lf = np.linalg.solve(ops.V1.T, psi.T).T
    assert np.allclose(lf[0], ops.ll, atol=1e-13) and np.allclose(lf[1], ops.lr, atol=1e-13)
    Vr = O.dvandermonde_1d(deg, ops.r)
    dl = np.linalg.solve(ops.V1.T, Vr.T).T
    assert np.allclose(dl, ops.dl, atol=1e-12)
    # quad mode ordering: mode a*(N+1)+b is P_a(r) P_b(s), node k + n*l at (r_k, r_l)
    k, l, a, b = 1, deg, deg, 0
    pa = O.jacobi_p(ops.r[k], 0, 0, a)
    pb = O.jacobi_p(ops.r[l], 0, 0, b)
    assert abs(ops.V[k + n * l, a * n + b] - pa * pb) < 1e-15


def test_jacobi_vs_normalised_legendre():
    x = np.linspace(-1, 1, 11)
    for N in range(6):
        c = np.zeros(N + 1)
        c[N] = 1
        ref = L.legval(x, c) * np.sqrt((2 * N + 1) / 2)
        got = np.array([O.jacobi_p(xi, 0, 0, N) for xi in x])
        assert np.allclose(got, ref, atol=1e-13)


def test_rs_jacobi_rectangle_and_parallelogram():
    vert = np.array([[0.0, 0.0], [2.0, 0.0], [2.0, 3.0], [0.0, 3.0]])
    J = O.rs_jacobi_quad(0.3, -0.2, vert)
    assert np.allclose(J, [[1.0, 0.0], [0.0, 1.5]])
    # parallelogram of test/runtests.jl:19: constant Jacobian
    v2 = np.array([[0, 0], [np.sqrt(3), -1], [np.sqrt(3) + 1, np.sqrt(3) - 1], [1, np.sqrt(3)]], dtype=float)
    J1 = O.rs_jacobi_quad(-0.7, 0.1, v2)
    J2 = O.rs_jacobi_quad(0.5, 0.9, v2)
    assert np.allclose(J1, J2) and abs(np.linalg.det(J1) - 1.0) < 1e-14


def test_shock_detector_known_answers():
    # test/runtests.jl:12-14 with deg = 3 (S0 = -3 log10 3, kappa = 4)
    assert O.shock_detector(-np.inf, 3) is False
    assert O.shock_detector(np.log10(0.1), 3) is True
    assert O.shock_detector(np.log10(1e5), 3) is True
    # sigma at Se = -1 equals 0.415706 (closed form)
    S0, k = -3 * np.log10(3), 4.0
    sigma = 0.5 * (1 - np.sin(0.5 * np.pi * (-1 - S0) / k))
    assert abs(sigma - 0.415706) < 1e-6


def test_modal_filter_l2_coefficients():
    u = np.ones(9)
    out = O.modal_filter_l2(u, 5e-4)
    n = np.arange(9)
    assert np.allclose(out, 1.0 / (1.0 + 5e-4 * (n + 1) ** 2 * n ** 2))


# ---------------------------------------------------------------- product host operators
@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_library_host_operators_match_oracle(deg):
    ops = O.Ops(deg)
    ps = frhip.FRPSpace2D(0.0, 1.0, 4, 0.0, 2.0, 3, deg, 1, 1)
    assert np.allclose(ps.xpl, ops.r, atol=1e-16)
    assert np.allclose(ps.ll, ops.ll, atol=1e-15) and np.allclose(ps.lr, ops.lr, atol=1e-15)
    assert np.allclose(ps.dl, ops.dl, atol=1e-13)
    assert np.allclose(ps.dhl, ops.dhl, atol=1e-13) and np.allclose(ps.dhr, ops.dhr, atol=1e-13)
    assert np.allclose(ps.V, ops.V, atol=1e-14) and np.allclose(ps.iV, ops.iV, atol=1e-12)
    assert np.allclose(ps.iJ_elem, np.diag([2 / 0.25, 2 / (2 / 3)]))
    ps1 = frhip.FRPSpace1D(0.0, 1.0, 10, deg, 0, "huynh")
    dgl, dgr = O.dcorrection("huynh", deg, ops.r)
    assert np.allclose(ps1.dhl, dgl) and np.allclose(ps1.dhr, dgr)
    assert frhip.shock_detector(np.log10(0.1), 3) is True and frhip.shock_detector(-np.inf, 3) is False


def test_frpspace2d_geometry():
    ps = frhip.FRPSpace2D(0.0, 2.0, 4, 0.0, 1.0, 2, 2, 1, 1)
    assert ps.x.shape == (6, 4) and ps.vertices.shape == (6, 4, 4, 2)
    # xpg of the bilinear map equals centre + r dx/2 on rectangles
    i, j, k, l = 2, 1, 0, 2
    assert np.isclose(ps.xpg[i, j, k, l, 0], ps.x[i, j] + ps.xpl[k] * 0.25)
    assert np.isclose(ps.xpg[i, j, k, l, 1], ps.y[i, j] + ps.xpl[l] * 0.25)
    assert ps.dx[0, 0] == 0.5 and ps.nx == 4


def test_gauss_hermite_vs_numpy():
    from frhip import bgk

    for n in (4, 8, 24):
        x, w = bgk.gauss_hermite(n)
        xn, wn = np.polynomial.hermite.hermgauss(n)
        assert np.allclose(x, xn, atol=1e-13) and np.allclose(w, wn, rtol=1e-12, atol=1e-300)
    u, wu = bgk.velocity_grid(24)
    # moments of a Maxwellian with lambda = 1: exact for the Gauss-Hermite grid
    M = np.exp(-u ** 2) / np.sqrt(np.pi)
    assert abs(np.sum(wu * M) - 1.0) < 1e-13 and abs(np.sum(wu * u * u * M) - 0.5) < 1e-13


def test_bgk_oracle_equilibrium_and_conservation():
    from frhip import bgk

    u, wu = bgk.velocity_grid(8)
    f0 = bgk.maxwellian_state(4, 3, 2, u, u, amp=0.0)
    assert np.max(np.abs(O.bgk2d_rhs(f0, 4, 3, 2, u, wu, u, wu, 1e-2))) < 1e-12
    f = bgk.maxwellian_state(5, 4, 2, u, u, amp=0.1)
    f *= 1.0 + 0.01 * np.sin(np.arange(f.size))
    df = O.bgk2d_rhs(f, 5, 4, 2, u, wu, u, wu, 1e-2)
    # mass: quadrature over points and velocities of df vanishes on the periodic grid
    wq = np.polynomial.legendre.leggauss(3)[1]
    wpt = np.einsum("k,l->kl", wq, wq).ravel(order="F")
    d = df.reshape((5 * 4, 9, 8, 8), order="F")
    tot = np.einsum("cpab,p,a,b->", d, wpt, wu, wu)
    assert abs(tot) < 1e-10 * np.abs(d).sum()
