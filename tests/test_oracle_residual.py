"""Pins the oracle's physics and residuals: Riemann-flux consistency / rotation / upwinding,
2D residual vs an independent numpy restatement and the committed golden vectors, free-stream
preservation, discrete conservation, DG equivalence of the 1D advection residual, the RK
stability polynomials, and the positivity limiter's invariants."""
import os

import numpy as np
import pytest

import np_reference as NR
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vortex_state(nx, ny, deg, L=10.0, seed=3, amp=0.01):
    import frhip
    from frhip import ics

    ps = frhip.FRPSpace2D(0.0, L, nx, 0.0, L, ny, deg, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4)
    return ps, (ics.perturbed(u, seed=seed, amplitude=amp) if amp else u)


# ---------------------------------------------------------------- fluxes
@pytest.mark.parametrize("kind", ["hll", "rusanov", "roe"])
def test_riemann_consistency_and_rotation(kind):
    rng = np.random.default_rng(0)
    for _ in range(20):
        prim = [rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.2, 2)]
        w = O.prim_conserve2(prim, 1.4)
        F, G = O.euler_flux2(w, 1.4)
        assert np.allclose(O.riemann2(kind, w, w, 1.4), F, rtol=1e-13, atol=1e-14)
        # a y-face flux is the x-face flux of the rotated states, rotated back (KitBase frames)
        wl = O.prim_conserve2([prim[0], prim[2], -prim[1], prim[3]], 1.4)
        f = O.riemann2(kind, wl, wl, 1.4)
        assert np.allclose([f[0], -f[2], f[1], f[3]], G, rtol=1e-13, atol=1e-14)


def test_hll_supersonic_upwind():
    wL = O.prim_conserve2([1.0, 5.0, 0.3, 2.0], 1.4)  # supersonic to the right
    wR = O.prim_conserve2([0.5, 4.0, -0.2, 1.0], 1.4)
    F, _ = O.euler_flux2(wL, 1.4)
    assert np.array_equal(O.riemann2("hll", wL, wR, 1.4), F)


def test_prim_conserve_roundtrip():
    prim = np.array([1.3, 0.4, -0.2, 0.7])
    w = O.prim_conserve2(prim, 1.4)
    assert np.allclose(O.conserve_prim2(w, 1.4), prim, rtol=1e-15)
    # lambda = rho / (2 p)
    p = 0.4 * (w[3] - 0.5 * (w[1] ** 2 + w[2] ** 2) / w[0])
    assert np.isclose(prim[3], w[0] / (2 * p))


# ---------------------------------------------------------------- 2D residual
@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("flux", ["hll", "rusanov"])
@pytest.mark.parametrize("bc", ["ghost", "periodic"])
def test_residual_vs_numpy_restatement(deg, flux, bc):
    nx, ny = 5, 4
    _, u = vortex_state(nx, ny, deg, seed=deg)
    orc = O.Fr2d(nx, ny, deg, 10 / nx, 10 / ny, 1.4, flux, bc)
    du = orc.rhs(u)
    ref = NR.fr2d_rhs(u, NR.operators(deg), 1.4, 1 / (10 / nx / 2), 1 / (10 / ny / 2), flux, bc == "periodic")
    assert np.max(np.abs(du - ref)) <= 5e-14 * np.max(np.abs(ref))


def test_residual_golden_vectors():
    files = sorted(f for f in os.listdir(GOLDEN) if f.startswith("fr2d_") and f.endswith(".npz"))
    assert files, "golden fixtures missing (tests/golden/make_golden.py)"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        nx, ny, deg = (int(v) for v in z["dims"])
        flux = str(z["flux"])
        bc = str(z["bc"])
        u = np.asfortranarray(z["u"])
        orc = O.Fr2d(nx, ny, deg, float(z["dx"]), float(z["dy"]), 1.4, flux, bc)
        du = orc.rhs(u)
        assert np.max(np.abs(du - z["du"])) <= 5e-14 * np.max(np.abs(z["du"])), f


@pytest.mark.parametrize("flux", ["hll", "rusanov", "roe"])
def test_free_stream_preservation(flux):
    nx, ny = 6, 5
    u = np.zeros((nx + 2, ny + 2, 4, 4, 4), order="F")
    u[..., 0], u[..., 1], u[..., 2], u[..., 3] = 1.3, 0.4, -0.7, 3.0
    orc = O.Fr2d(nx, ny, 3, 1 / nx, 1 / ny, 1.4, flux, "periodic")
    assert np.max(np.abs(orc.rhs(u))) < 1e-12


def test_discrete_conservation_periodic():
    ps, u = vortex_state(9, 7, 3)
    orc = O.Fr2d(9, 7, 3, 10 / 9, 10 / 7, 1.4, "rusanov", "periodic")
    du = orc.rhs(u)
    tot = np.einsum("ijklm,kl->m", du[1:-1, 1:-1], ps.wp)
    assert np.all(np.abs(tot) < 1e-12 * np.abs(du).sum())


def test_planted_sign_error_is_caught():
    """A sign flip in the correction term (a plausible restatement mistake) breaks both the
    numpy cross-check and conservation: the checks above are not vacuous."""
    nx, ny, deg = 5, 4, 3
    ps, u = vortex_state(nx, ny, deg)
    ops = NR.operators(deg)
    bad = dict(ops)
    bad["dhr"] = -ops["dhr"]
    good = NR.fr2d_rhs(u, ops, 1.4, 2 / (10 / nx), 2 / (10 / ny), "rusanov", True)
    wrong = NR.fr2d_rhs(u, bad, 1.4, 2 / (10 / nx), 2 / (10 / ny), "rusanov", True)
    orc = O.Fr2d(nx, ny, deg, 10 / nx, 10 / ny, 1.4, "rusanov", "periodic")
    ref = orc.rhs(u)
    assert np.max(np.abs(good - ref)) < 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(wrong - ref)) > 1e-3 * np.max(np.abs(ref))
    tot = np.einsum("ijklm,kl->m", wrong[1:-1, 1:-1], ps.wp)
    assert np.max(np.abs(tot)) > 1e-6


# ---------------------------------------------------------------- 1D residuals
def test_advection_upwind_equals_nodal_dg():
    """Radau correction + upwind flux is the nodal DG operator (Huynh 2007): compare with a
    brute-force DG assembly (mass/stiffness by exact quadrature) on a periodic mesh."""
    deg, ncell = 3, 6
    ops = O.Ops(deg)
    dx = 1.0 / ncell
    J = np.full(ncell, dx / 2)
    rng = np.random.default_rng(1)
    u = np.asfortranarray(rng.standard_normal((ncell, deg + 1)))
    du = O.fr1d_advection_rhs(u, J, ops, a=1.0, bc="period")
    # weak-form DG: M du = S^T u - (u_R^upw l_R - u_L^upw l_L), scaled by 1/J, a = 1
    r, w = ops.r, ops.w
    M = np.diag(w)
    S = M @ ops.dl  # S[m, k] = int l_m l_k'
    ref = np.zeros_like(u)
    for i in range(ncell):
        ui = u[i]
        up_left = u[i - 1] @ ops.lr  # upwind value on the left face
        up_right = ui @ ops.lr
        rhs = S.T @ ui - (up_right * ops.lr - up_left * ops.ll)
        ref[i] = np.linalg.solve(M, rhs) / J[i]
    # the reference regularises the upwind speed with +1e-8 / +1e-6, so equality holds to ~1e-6
    assert np.max(np.abs(du - ref)) < 1e-5 * np.max(np.abs(ref))


def test_c1_conditioning():
    """C1 (advection, 100 cells, p = 2, RK4, dt = 0.1 dx, t = 1): a relative perturbation of
    1e-15 in u0 changes the oracle's own final state by about 1e-7, because the reference's
    regularised flux divides face jumps of 1e-7 .. 5e-6 by (jump + 1e-8).  This bounds what any
    two correct FP64 implementations can agree to on this run."""
    import frhip

    ncell, deg = 100, 2
    ps = frhip.FRPSpace1D(0.0, 1.0, ncell, deg)
    ops = O.Ops(deg)
    u0 = np.asfortranarray(np.sin(2 * np.pi * ps.xpg))
    dt = 0.1 / ncell

    def run(pert):
        L = lambda v: O.fr1d_advection_rhs(v, ps.J, ops, 1.0, "period")  # noqa: E731
        uc = u0 * (1 + pert)
        for _ in range(1000):
            k1 = L(uc)
            y = uc + (0.5 * dt) * k1
            acc = uc + (dt / 6.0) * k1
            k2 = L(y)
            y = uc + (0.5 * dt) * k2
            acc = acc + (dt / 3.0) * k2
            k3 = L(y)
            y = uc + dt * k3
            acc = acc + (dt / 3.0) * k3
            uc = acc + (dt / 6.0) * L(y)
        return uc

    a, b = run(0.0), run(1e-15)
    d = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-9 < d < 1e-6
    exact = np.sin(2 * np.pi * (ps.xpg - 1.0))
    assert np.sum(np.abs(a - exact)) / (ncell * (deg + 1)) < 1e-4


def test_euler1d_free_stream_and_conservation():
    deg, ncell = 2, 8
    ops = O.Ops(deg)
    J = np.full(ncell, 0.5 / ncell)
    u = np.zeros((ncell, deg + 1, 3), order="F")
    u[..., 0], u[..., 1], u[..., 2] = 1.0, 0.3, 2.5
    assert np.max(np.abs(O.fr1d_euler_rhs(u, J, ops, 1.4, "hll", "period"))) < 1e-12
    rng = np.random.default_rng(2)
    u[..., 0] += 0.05 * rng.standard_normal((ncell, deg + 1))
    du = O.fr1d_euler_rhs(u, J, ops, 1.4, "rusanov", "period")
    tot = np.einsum("ipk,p->k", du * J[:, None, None], ops.w)
    assert np.all(np.abs(tot) < 1e-12)
    dd = O.fr1d_euler_rhs(u, J, ops, 1.4, "rusanov", "dirichlet")
    assert np.all(dd[0] == 0) and np.all(dd[-1] == 0)


# ---------------------------------------------------------------- steppers
def test_ssprk3_and_rk4_stability_polynomials():
    """On the linear advection operator the one-step maps are the Taylor polynomials of the
    RK stability functions: SSP-RK3 -> 1 + z + z^2/2 + z^3/6, RK4 adds z^4/24 (tested through
    the 2D stage routine with a scalar-like linearised check on a tiny Euler state)."""
    nx, ny, deg = 4, 4, 2
    ps, u = vortex_state(nx, ny, deg, amp=0.0)
    orc = O.Fr2d(nx, ny, deg, 10 / nx, 10 / ny, 1.4, "rusanov", "periodic")
    dt = 1e-3
    # SSP-RK3 step equals the Shu-Osher stage composition
    u3 = u.copy(order="F")
    orc.ssprk3(u3, dt, 1)
    a = orc.stage(u, u, 0.0, 1.0, dt)
    b = orc.stage(u, a, 0.75, 0.25, 0.25 * dt)
    c = orc.stage(u, b, 1 / 3, 2 / 3, 2 / 3 * dt)
    orc.fill_periodic(c)
    assert np.array_equal(u3, c)
    # order check: error vs an RK4 reference with dt/4 shrinks ~8x when dt halves (3rd order)
    def err(h):
        x = u.copy(order="F")
        orc.ssprk3(x, h, int(round(0.02 / h)))
        y = u.copy(order="F")
        orc.rk4(y, h / 4, int(round(0.08 / h)))
        return np.linalg.norm(x - y)

    e1, e2 = err(4e-3), err(2e-3)
    assert 5.0 < e1 / e2 < 12.0


def test_positivity_limiter_invariants():
    ops = O.Ops(3)
    wts = (ops.wp / 4.0).ravel(order="F")
    rng = np.random.default_rng(4)
    ue = np.zeros((4, 4, 4), order="F")
    prim = np.stack([1 + 0.3 * rng.standard_normal((4, 4)), 0.5 * rng.standard_normal((4, 4)),
                     0.5 * rng.standard_normal((4, 4)), np.full((4, 4), 0.5)], axis=-1)
    for k in range(4):
        for l in range(4):
            ue[k, l] = O.prim_conserve2(np.abs(prim[k, l]) + [0, 0, 0, 0], 1.4)
    ue[1, 2, 0] = -0.3  # negative density at one point
    ue[2, 1, 3] *= 0.01  # nearly no internal energy at another
    mean0 = np.einsum("klm,kl->m", ue, ops.wp / 4)
    rc, out = O.positive_limiter2d(ue, 1.4, wts, ops.ll, ops.lr)
    assert rc == 0
    mean1 = np.einsum("klm,kl->m", out, ops.wp / 4)
    assert np.allclose(mean0, mean1, rtol=1e-13)  # cell means preserved
    eps = min(1e-13, mean0[0])
    assert out[..., 0].min() >= eps - 1e-15
    p = 0.4 * (out[..., 3] - 0.5 * (out[..., 1] ** 2 + out[..., 2] ** 2) / out[..., 0])
    assert p.min() > -1e-12


def test_positivity_limiter_1d_euler():
    """dissipation.jl:61-123: cell means kept; density >= eps at solution and flux points after the
    density squeeze; pressure >= eps after the pressure squeeze, whose factor is the root of
    tj_equation (p(t) = eps) at the worst point; a positive element is returned untouched."""
    deg = 4
    ops = O.Ops(deg)
    w = ops.w / 2.0
    rng = np.random.default_rng(8)
    nsp = deg + 1
    prim = np.stack([1 + 0.2 * rng.random(nsp), 0.3 * rng.standard_normal(nsp), np.full(nsp, 0.6)], axis=-1)
    good = np.asfortranarray(np.array([O.prim_conserve1(q, 1.4) for q in prim]))
    rc, same = O.positive_limiter1d_euler(good, 1.4, w, ops.ll, ops.lr)
    assert rc == 0 and np.array_equal(same, good)
    bad = good.copy(order="F")
    bad[1, 0] = -0.2          # negative density at one point
    bad[3, 2] *= 0.02         # pressure driven negative at another
    mean0 = w @ bad
    rc, out = O.positive_limiter1d_euler(bad, 1.4, w, ops.ll, ops.lr)
    assert rc == 0
    assert np.allclose(w @ out, mean0, rtol=1e-13, atol=1e-15)
    rho = np.concatenate([out[:, 0], [ops.ll @ out[:, 0], ops.lr @ out[:, 0]]])
    eps = min(1e-13, mean0[0])
    assert rho.min() >= eps - 1e-14
    p = 0.4 * (out[:, 2] - 0.5 * out[:, 1] ** 2 / out[:, 0])
    assert p.min() >= -1e-12
    # the squeeze is the largest admissible one: the worst point sits on p = eps (up to round-off)
    assert abs(p.min()) < 1e-10


# ----------------------------------------------------------------- parallelogram cells
def _rot(u, c, s):
    """Momentum components rotated by (c, s): (rho, c m_x - s m_y, s m_x + c m_y, E)."""
    v = u.copy(order="F")
    v[..., 1] = c * u[..., 1] - s * u[..., 2]
    v[..., 2] = s * u[..., 1] + c * u[..., 2]
    return v


def _vortex_state(nx, ny, deg, seed):
    from frhip import ics
    import frhip

    ps = frhip.FRPSpace2D(0.0, 1.0, nx, 0.0, 1.0, ny, deg, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4, strength=3.0, center=(0.5, 0.5))
    return ics.perturbed(u, seed=seed, amplitude=0.01)


def test_parallelogram_form_reduces_to_rectangles():
    """The interface terms of dev/parallelogram.jl with iJ = diag(2/dx, 2/dy), n1 = (1, 0),
    n2 = (0, 1) are those of example/shock-vortex.jl, bitwise."""
    nx, ny, deg = 6, 5, 2
    dx, dy = 1.0 / nx, 1.0 / ny
    u = _vortex_state(nx, ny, deg, 3)
    for flux in ("hll", "rusanov", "roe"):
        a = O.Fr2d(nx, ny, deg, dx, dy, flux=flux, bc="periodic").rhs(u)
        b = O.Fr2d(nx, ny, deg, dx, dy, flux=flux, bc="periodic", iJ=[[2 / dx, 0.0], [0.0, 2 / dy]],
                   n1=(1.0, 0.0), n2=(0.0, 1.0)).rhs(u)
        assert np.array_equal(a, b)


def test_rotated_grid_is_the_rotated_problem():
    """Euler is rotation invariant: the same logical grid turned by theta (iJ = D^-1 R^T, face
    normals R e1, R e2) with the momenta turned by R gives the turned residual.  A dropped
    off-diagonal iJ term, a wrong frame sign or a transposed iJ breaks this."""
    nx, ny, deg = 6, 5, 3
    dx, dy = 1.0 / nx, 0.7 / ny
    u = _vortex_state(nx, ny, deg, 4)
    for theta in (0.3, -1.1, 2.5):
        c, s = np.cos(theta), np.sin(theta)
        iJ = np.array([[2 * c / dx, 2 * s / dx], [-2 * s / dy, 2 * c / dy]])
        for flux in ("hll", "rusanov", "roe"):
            du = O.Fr2d(nx, ny, deg, dx, dy, flux=flux, bc="periodic").rhs(u)
            dur = O.Fr2d(nx, ny, deg, dx, dy, flux=flux, bc="periodic", iJ=iJ, n1=(c, s), n2=(-s, c)).rhs(_rot(u, c, s))
            ref = _rot(du, c, s)
            assert np.max(np.abs(dur - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_parallelogram_free_stream():
    """A uniform state on sheared periodic cells (J = [a/2, b/2], a = (dx, 0), b = (t, dy)) has
    zero residual to round-off."""
    nx, ny, deg = 5, 4, 3
    dx, dy, t = 0.2, 0.25, 0.13
    iJ = np.linalg.inv(np.array([[dx / 2, t / 2], [0.0, dy / 2]]))
    u = np.zeros((nx + 2, ny + 2, deg + 1, deg + 1, 4), order="F")
    u[...] = O.prim_conserve2([1.2, 0.7, -0.4, 0.9], 1.4)
    for flux in ("hll", "rusanov", "roe"):
        du = O.Fr2d(nx, ny, deg, dx, dy, flux=flux, bc="periodic", iJ=iJ,
                    n1=iJ[0] / np.linalg.norm(iJ[0]), n2=iJ[1] / np.linalg.norm(iJ[1])).rhs(u)
        assert np.max(np.abs(du)) <= 1e-12 * np.max(np.abs(iJ)) * np.max(np.abs(u))
