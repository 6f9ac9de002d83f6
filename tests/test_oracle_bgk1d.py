"""Pins the oracle of the 1D discrete-velocity BGK model (example/bgk_wave.jl:69-129, mol!):
- free transport (tau = inf) is, for every velocity, the weak-form nodal DG operator with the
  exact upwind flux, assembled here by quadrature (Radau correction = nodal DG, Huynh 2007), for
  velocities of both signs;
- transport conserves the discrete mass, momentum and energy on the periodic mesh;
- the Maxwellian of a state carries the state's own moments (sum_j w_j phi(u_j) M_j against the
  moments of f, phi = 1, u, u^2 / 2, on a fine velocity grid): this fixes the normalisation
  sqrt(lambda / pi) and conserve_prim at gamma = 3 (p = rho / (2 lambda), E = p / 2 + rho U^2 / 2);
- a spatially uniform Maxwellian is a steady state;
- relaxation alone (a spatially uniform, non-Maxwellian f) is (M - f) / tau with M from those
  moments.
"""
import numpy as np
import pytest

from oracle import oracle as O


def vspace(u0, u1, nu):
    du = (u1 - u0) / nu
    return u0 + (np.arange(nu) + 0.5) * du, np.full(nu, du)


def dg_upwind(u, a, ops, J):
    """Weak-form DG of u_t + a u_x = 0 with the exact upwind flux, periodic: M du = a S^T u -
    (F_R l_R - F_L l_L), S = M D (exact for these polynomial degrees)."""
    ncell = u.shape[0]
    M = np.diag(ops.w)
    S = M @ ops.dl
    out = np.zeros_like(u)
    for i in range(ncell):
        if a > 0:
            FL = a * (u[i - 1] @ ops.lr)
            FR = a * (u[i] @ ops.lr)
        else:
            FL = a * (u[i] @ ops.ll)
            FR = a * (u[(i + 1) % ncell] @ ops.ll)
        rhs = a * (S.T @ u[i]) - (FR * ops.lr - FL * ops.ll)
        out[i] = np.linalg.solve(M, rhs) / J
    return out


@pytest.mark.parametrize("deg", [1, 2, 3, 5])
def test_free_transport_is_upwind_nodal_dg(deg):
    nx, dx = 9, 0.125
    uv, wv = vspace(-3.0, 3.0, 8)
    orc = O.Bgk1d(nx, deg, uv, wv, dx, tau=np.inf)
    rng = np.random.default_rng(deg)
    f = np.asfortranarray(rng.uniform(0.1, 1.0, (nx, 8, deg + 1)))
    du = orc.rhs(f)
    for j, a in enumerate(uv):
        ref = dg_upwind(f[:, j, :], a, orc.ops, 0.5 * dx)
        assert np.max(np.abs(du[:, j, :] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_free_transport_conserves_moments():
    nx, deg, dx = 12, 2, 1.0 / 12
    uv, wv = vspace(-5.0, 5.0, 20)
    orc = O.Bgk1d(nx, deg, uv, wv, dx, tau=np.inf)
    rng = np.random.default_rng(4)
    f = np.asfortranarray(rng.uniform(0.1, 1.0, (nx, 20, deg + 1)))
    du = orc.rhs(f)
    # sum over cells and points with the Gauss weights, over velocities with the grid weights
    wq = orc.ops.w
    for phi in (np.ones_like(uv), uv, 0.5 * uv * uv):
        tot = np.einsum("ijk,k,j->", du, wq, wv * phi)
        scale = np.einsum("ijk,k,j->", np.abs(du), wq, wv * np.abs(phi))
        assert abs(tot) <= 1e-13 * scale


def test_maxwellian_carries_the_state_moments():
    """With tau = 1 and a spatially uniform state, du = M - f; its moments vanish when M is the
    Maxwellian of f's own moments (fine grid on [-12, 12], so the discrete moments of M are those
    of the continuous Maxwellian to round-off)."""
    nx, deg = 3, 2
    uv, wv = vspace(-12.0, 12.0, 400)
    orc = O.Bgk1d(nx, deg, uv, wv, 0.1, tau=1.0)
    g = 0.7 * np.exp(-1.3 * (uv - 0.4) ** 2) + 0.4 * np.exp(-0.8 * (uv + 1.1) ** 2)  # not a Maxwellian
    f = np.asfortranarray(np.broadcast_to(g[None, :, None], (nx, uv.size, deg + 1)))
    du = orc.rhs(f)
    assert np.max(np.abs(du)) > 1e-2  # far from equilibrium
    for phi in (np.ones_like(uv), uv, 0.5 * uv * uv):
        m_du = np.einsum("ijk,j->ik", du, wv * phi)
        m_f = np.einsum("ijk,j->ik", f, wv * np.abs(phi))
        assert np.max(np.abs(m_du)) <= 1e-12 * np.max(m_f)
    # and (M - f) itself: the closed-form Maxwellian of those moments
    rho = np.sum(wv * g)
    U = np.sum(wv * uv * g) / rho
    E = 0.5 * np.sum(wv * uv * uv * g)
    lam = 0.5 * rho / (3.0 - 1.0) / (E - 0.5 * rho * U * U)
    M = rho * np.sqrt(lam / np.pi) * np.exp(-lam * (uv - U) ** 2)
    assert np.max(np.abs(du[1, :, 1] - (M - g))) <= 1e-13 * np.max(np.abs(M))


def test_uniform_maxwellian_is_steady():
    nx, deg = 10, 3
    uv, wv = vspace(-10.0, 10.0, 200)
    orc = O.Bgk1d(nx, deg, uv, wv, 0.1, tau=1e-2)
    M = 1.3 * np.sqrt(0.9 / np.pi) * np.exp(-0.9 * (uv - 0.6) ** 2)
    f = np.asfortranarray(np.broadcast_to(M[None, :, None], (nx, uv.size, deg + 1)))
    du = orc.rhs(f)
    assert np.max(np.abs(du)) <= 1e-10 * np.max(M) / 1e-2


def test_bgk_wave_setup_runs_and_keeps_mass():
    """The script's own case (20 cells, degree 2, 100 velocities on [-5, 5], tau = 1e-2) through
    20 Midpoint steps of dt = cfl dx / u1: finite, total mass kept to 1e-7.  Transport conserves
    mass exactly; relaxation only up to the velocity grid's truncation of the Maxwellian (U = 1,
    lambda ~ 1: exp(-16) at u = 5), which moves the mass by ~2.5e-8 here."""
    import frhip
    from frhip import bgk

    ps = frhip.FRPSpace1D(0.0, 1.0, 20, 2)
    vs = bgk.VSpace1D(-5.0, 5.0, 100)
    f0 = bgk.bgk_wave_initial(ps, vs)
    orc = O.Bgk1d(20, 2, vs.u, vs.weights, 1.0 / 20)
    dt = 0.1 * (1.0 / 20) / 5.0
    f = orc.step(f0, dt, 20, "midpoint")
    assert np.all(np.isfinite(f))
    mass = lambda a: np.einsum("ijk,k,j->", a, orc.ops.w, vs.weights)  # noqa: E731
    assert abs(mass(f) - mass(f0)) <= 1e-7 * mass(f0)
