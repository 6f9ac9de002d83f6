"""The calls of the reference's own test file (test/runtests.jl), in order, on frhip's host
functions, with the values the mathematics fixes for them.  test_triangle.jl and TriFRPSpace
(triangle meshes read with PyCall) are out of scope; GPU counterparts of the solver calls are in
the tests/test_gpu_*.py files."""
import numpy as np

import frhip


def test_errors_runtests_6_10():
    rng = np.random.default_rng(0)
    u, u0 = rng.random(3), rng.random(3)
    d = np.abs(u - u0) * 0.1
    assert np.isclose(frhip.L1_error(u, u0, 0.1), d.sum(), rtol=1e-15)
    assert np.isclose(frhip.L2_error(u, u0, 0.1), np.sqrt((d * d).sum()), rtol=1e-15)
    assert np.isclose(frhip.Linf_error(u, u0, 0.1), d.max(), rtol=1e-15)


def test_shock_detector_runtests_12_14():
    assert frhip.shock_detector(-np.inf, 3) is False
    assert frhip.shock_detector(np.log10(0.1), 3) is True
    assert frhip.shock_detector(np.log10(1e5), 3) is True


def test_rs_jacobi_forms_runtests_16_20():
    ps2 = frhip.FRPSpace2D(0.0, 1.0, 20, 0.0, 1.0, 20, 5, 1, 1)
    r = ps2.xpl
    # a square of side 2 turned by -30 degrees: the bilinear map is a rotation, J constant
    v = np.array([[0.0, 0.0], [np.sqrt(3), -1.0], [np.sqrt(3) + 1, np.sqrt(3) - 1], [1.0, np.sqrt(3)]])
    J = frhip.rs_jacobi(r, v)
    assert J.shape == (6, 6, 2, 2)
    rot = np.array([[np.sqrt(3) / 2, 0.5], [-0.5, np.sqrt(3) / 2]])
    assert np.max(np.abs(J - rot)) <= 1e-15
    # per cell of a random vertex array: the tensor form at every (r[i], r[j]) of every cell
    vr = np.random.default_rng(1).random((3, 3, 4, 2))
    Jr = frhip.rs_jacobi(r, vr)
    assert Jr.shape == (3, 3, 6, 6, 2, 2)
    for (a, b, i, j) in [(0, 0, 0, 0), (2, 1, 3, 5), (1, 2, 5, 1)]:
        assert np.array_equal(Jr[a, b, i, j], frhip.rs_jacobi(r[i], r[j], vr[a, b]))
    # the pairwise vector form and the reference's linear-index matrix form
    pair = frhip.rs_jacobi(r, r[::-1], v)
    assert pair.shape == (6, 2, 2) and np.array_equal(pair[2], frhip.rs_jacobi(r[2], r[3], v))
    ri = np.stack([r, np.ones(6), r[::-1], np.zeros(6)])
    si = np.stack([np.zeros(6), r, np.ones(6), r[::-1]])
    lin = frhip.rs_jacobi(ri, si, vr[1, 1])
    assert lin.shape == (4, 6, 2, 2)
    for f, (rr, ss) in enumerate(((r[0], 0.0), (1.0, r[0]), (r[-1], 1.0), (0.0, r[-1]))):
        assert np.array_equal(lin[f, 4], frhip.rs_jacobi(rr, ss, vr[1, 1]))


def test_positive_limiter_calls_runtests_24_25():
    """positive_limiter(ones(6), 1/6, ll, lr) and the Euler form on ones(6, 3): a constant state is
    admissible, so both return it unchanged (the oracle restates the limiter; the library's device
    limiter is compared with it in the GPU tests)."""
    from oracle import oracle as O

    ops = O.Ops(5)
    rc, u = O.positive_limiter1d_scalar(np.ones(6), 1.0 / 6.0, ops.ll, ops.lr)
    assert rc == 0 and np.array_equal(u, np.ones(6))
    rc, ue = O.positive_limiter1d_euler(np.ones((6, 3), order="F"), 5.0 / 3.0, np.full(6, 1.0 / 6.0), ops.ll, ops.lr)
    assert rc == 0 and np.array_equal(ue, np.ones((6, 3)))


def test_basis_norm_and_filter_exp_runtests_27_31():
    deg = 5
    nrm = frhip.basis_norm(deg)
    x = np.linspace(-1.0, 1.0, 100)
    dx = x[1] - x[0]
    # P_0 = 1 / sqrt(2) on the orthonormal basis: 100 dx / sqrt(2)
    assert np.isclose(nrm[0], 100 * dx / np.sqrt(2.0), rtol=1e-14)
    # P_n from numpy's Legendre module, scaled by sqrt((2n + 1) / 2)
    for n in range(deg + 1):
        Pn = np.polynomial.legendre.legval(x, np.eye(deg + 1)[n]) * np.sqrt((2 * n + 1) / 2)
        assert np.isclose(nrm[n], dx * np.abs(Pn).sum(), rtol=1e-13)
    ps = frhip.FRPSpace1D(0.0, 1.0, 20, deg)
    F = frhip.filter_exp(2, 10, ps.V)  # 6 x 6 V: the triangle branch of N = 2 (nv = 6)
    assert F.shape == (6, 6) and np.all(np.isfinite(F))


def test_interp_face_and_standard_lagrange_runtests_39_43():
    ps = frhip.FRPSpace1D(0.0, 1.0, 20, 5)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((5, 6))
    ll, lr, dl = frhip.standard_lagrange(ps.xpl)
    assert np.allclose(ll, ps.ll, atol=1e-15) and np.allclose(lr, ps.lr, atol=1e-15)
    assert np.allclose(dl @ ps.xpl, np.ones(6), atol=1e-12)  # derivative of x
    # interp_face! is a device routine (tests/test_gpu_ops.py); its host meaning:
    ff = np.stack([f @ ll, f @ lr], axis=1)
    assert ff.shape == (5, 2)


def test_global_sp_and_r_x():
    """global_sp / r_x (geo_points.jl:8-41, geo_transform.jl:7): the line form as bgk_wave.jl:29
    calls it gives FRPSpace1D's xpg; the quadrilateral form on a rectangular vertex grid gives the
    interior of FRPSpace2D's xpg; r_x maps -1, 0, 1 to the ends and the midpoint."""
    assert frhip.r_x(-1.0, 2.0, 5.0) == 2.0 and frhip.r_x(1.0, 2.0, 5.0) == 5.0 and frhip.r_x(0.0, 2.0, 5.0) == 3.5
    ps = frhip.FRPSpace1D(0.0, 1.0, 20, 2)
    xface = np.linspace(0.0, 1.0, 21)
    assert np.allclose(frhip.global_sp(xface, ps.xpl), ps.xpg, rtol=0, atol=1e-15)
    ps2 = frhip.FRPSpace2D(0.0, 2.0, 5, -1.0, 1.0, 4, 3, 1, 1)
    xi, yi = np.meshgrid(np.linspace(0.0, 2.0, 6), np.linspace(-1.0, 1.0, 5), indexing="ij")
    xp, yp = frhip.global_sp(xi, yi, ps2.xpl)
    assert xp.shape == (5, 4, 4, 4)
    assert np.allclose(xp, ps2.xpg[1:-1, 1:-1, :, :, 0], rtol=0, atol=1e-14)
    assert np.allclose(yp, ps2.xpg[1:-1, 1:-1, :, :, 1], rtol=0, atol=1e-14)


def test_dlegendre_closed_forms():
    """∂legendre (poly_legendre.jl:13-22) against numpy's Legendre derivative at interior points
    and the end values P_p'(+-1) = (+-1)^(p+1) p (p + 1) / 2."""
    x = np.linspace(-1.0, 1.0, 41)
    for p in range(0, 10):
        ref = np.polynomial.legendre.legval(x, np.polynomial.legendre.legder(np.eye(p + 1)[p]))
        got = frhip.dlegendre(p, x)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13), p
        assert frhip.dlegendre(p, 1.0) == p * (p + 1) / 2
        assert frhip.dlegendre(p, -1.0) == (-1.0) ** (p + 1) * p * (p + 1) / 2


def test_frpspace1d_keyword_constructor():
    """FRPSpace1D(; cfg...) of example/advection_lowlevel.jl:49-51 (extra keys ignored, as the
    reference's keyword method, struct.jl:87-88, does)."""
    cfg = dict(x0=-1, x1=1, nx=100, deg=2, cfl=0.05, t=0.0, a=1.0)
    a = frhip.FRPSpace1D(**cfg)
    b = frhip.FRPSpace1D(-1, 1, 100, 2)
    assert np.array_equal(a.xpg, b.xpg) and np.array_equal(a.dl, b.dl) and a.nx == 100


def test_frpspace2d_keyword_constructor():
    cfg = dict(x0=0.0, x1=1.0, nx=20, y0=0.0, y1=1.0, ny=10, deg=3, ngx=1, ngy=1, cfl=0.1, gamma=1.4)
    a = frhip.FRPSpace2D(**cfg)
    b = frhip.FRPSpace2D(0.0, 1.0, 20, 0.0, 1.0, 10, 3, 1, 1)
    assert np.array_equal(a.xpg, b.xpg) and np.array_equal(a.iJ, b.iJ)


def test_problem_constructors_take_both_forms():
    """FREulerProblem / FRAdvectionProblem accept frhip's (ps, param, bc[, flux]) and the reference's
    (u, tspan, ps, param, bc) (eq_euler.jl:1, eq_advection.jl:1); argument errors are raised before
    any device call."""
    import pytest

    ps = frhip.FRPSpace1D(0.0, 1.0, 10, 2)
    with pytest.raises(TypeError):
        frhip.FREulerProblem(ps, 1.4, "dirichlet", "hll", speed=1)
    with pytest.raises(ValueError):
        frhip.FREulerProblem(np.zeros((10, 3, 3)), (0.0, 1.0))  # no space
