"""Initial states of the workloads, written in the reference's state layout
u(i, j, k, l, m) with axes (0:nx+1, 0:ny+1, 1:nsp, 1:nsp, 1:4) (example/shock-vortex.jl:160).

Primitive convention of KitBase: prim = [rho, U, V, lambda] with lambda = rho / (2 p).
"""
from __future__ import annotations

import numpy as np


def prim_conserve(prim, gamma):
    """KitBase prim_conserve for 2D states; prim[..., 0:4] -> w[..., 0:4]."""
    prim = np.asarray(prim, dtype=np.float64)
    rho, U, V, lam = prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 3]
    w = np.empty_like(prim)
    w[..., 0] = rho
    w[..., 1] = rho * U
    w[..., 2] = rho * V
    w[..., 3] = 0.5 * rho / lam / (gamma - 1.0) + 0.5 * rho * (U * U + V * V)
    return w


def conserve_prim(w, gamma):
    """KitBase conserve_prim for 2D states."""
    w = np.asarray(w, dtype=np.float64)
    prim = np.empty_like(w)
    prim[..., 0] = w[..., 0]
    prim[..., 1] = w[..., 1] / w[..., 0]
    prim[..., 2] = w[..., 2] / w[..., 0]
    prim[..., 3] = 0.5 * w[..., 0] / (gamma - 1.0) / (w[..., 3] - 0.5 * (w[..., 1] ** 2 + w[..., 2] ** 2) / w[..., 0])
    return prim


def _to_state(w):
    """(nx+2, ny+2, nsp, nsp, 4) Fortran-ordered float64 state."""
    return np.asfortranarray(w, dtype=np.float64)


def isentropic_vortex(ps, gamma=1.4, strength=5.0, center=None, rho_inf=1.0, u_inf=1.0, v_inf=1.0, p_inf=1.0,
                      t=0.0):
    """Isentropic vortex of Shu (strength eps, unit radius) on uniform flow, at every solution
    point including the ghost layer:
        du = -eps/(2 pi) e^{(1-r^2)/2} (y - yc),  dv = eps/(2 pi) e^{(1-r^2)/2} (x - xc),
        T = T_inf - (gamma-1) eps^2 / (8 gamma pi^2) e^{1-r^2},  rho = rho_inf (T/T_inf)^{1/(gamma-1)},
        p = rho T   (T = p/rho).
    t > 0 gives the exact solution at time t on the periodic domain of ps: the initial vortex
    carried by (u_inf, v_inf) t, every point measured from the periodic image of the centre
    nearest to it (minimum-image distance)."""
    if center is None:
        center = (0.5 * (ps.x0 + ps.x1), 0.5 * (ps.y0 + ps.y1))
    x = ps.xpg[..., 0]
    y = ps.xpg[..., 1]
    xc, yc = center
    if t:
        Lx, Ly = ps.x1 - ps.x0, ps.y1 - ps.y0
        # offsets from the carried centre wrapped into [-L/2, L/2)
        x = xc + np.mod(x - (xc + u_inf * t) + 0.5 * Lx, Lx) - 0.5 * Lx
        y = yc + np.mod(y - (yc + v_inf * t) + 0.5 * Ly, Ly) - 0.5 * Ly
    r2 = (x - xc) ** 2 + (y - yc) ** 2
    T_inf = p_inf / rho_inf
    e = strength / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = T_inf - (gamma - 1.0) * strength ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = rho_inf * (T / T_inf) ** (1.0 / (gamma - 1.0))
    p = rho * T
    prim = np.stack([rho, u_inf - e * (y - yc), v_inf + e * (x - xc), rho / (2.0 * p)], axis=-1)
    return _to_state(prim_conserve(prim, gamma))


def rankine_hugoniot(mach, gamma):
    """Upstream / downstream states of a stationary normal shock (upstream rho = 1, p = 1,
    u = M c).  Returns (rho1, u1, p1), (rho2, u2, p2)."""
    c1 = np.sqrt(gamma)
    u1 = mach * c1
    m2 = mach * mach
    rr = (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    pr = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0)
    return (1.0, u1, 1.0), (rr, u1 / rr, pr)


def shock_vortex_script(ps, gamma=5.0 / 3.0, mach=1.12):
    """The initial condition of example/shock-vortex.jl:160-190 as the script writes it: cells whose
    centre lies at x <= x1 / 4 take the post-shock state moving at u1 - u2, the others the quiescent
    pre-shock state with the vortex (kappa 0.2, mu 0.204, rc 0.05, centre (0.8, 0.5), within 8 rc)
    added per point, its density and temperature by the script's own expressions.  The shock states
    come from the Rankine-Hugoniot relations of `rankine_hugoniot` (pre-shock rho = 1, p = 1):
    KitBase's ib_rh, which the script calls, is not part of the reference tree, so its
    normalisation is not pinned."""
    (r1, u1, p1), (r2, u2, p2) = rankine_hugoniot(mach, gamma)
    lam1, lam2 = r1 / (2.0 * p1), r2 / (2.0 * p2)
    x = ps.xpg[..., 0]
    y = ps.xpg[..., 1]
    xc = np.broadcast_to(ps.base.x[:, :, None, None], x.shape)
    kap, mu, rc, x0, y0 = 0.2, 0.204, 0.05, 0.8, 0.5
    r = np.sqrt((x - x0) ** 2 + (y - y0) ** 2)
    eta = r / rc
    du = kap * eta * np.exp(mu * (1 - eta ** 2)) * (y - y0) / r
    dv = -kap * eta * np.exp(mu * (1 - eta ** 2)) * (x - x0) / r
    dT = -(gamma - 1) * kap ** 2 / (8 * mu * gamma) * np.exp(2 * mu * (1 - eta ** 2))
    T0 = 1.0 / lam1
    rho_v = r1 ** (gamma - 1) * (T0 + dT) / T0 ** (1 / (gamma - 1))
    vort = r <= rc * 8
    up = xc <= ps.x1 * 0.25
    rho = np.where(up, r2, np.where(vort, rho_v, r1))
    U = np.where(up, u1 - u2, np.where(vort, du, 0.0))
    V = np.where(up, 0.0, np.where(vort, dv, 0.0))
    lam = np.where(up, lam2, np.where(vort, 1.0 / (1.0 / lam1 + dT), lam1))
    prim = np.stack([rho, U, V, lam], axis=-1)
    return _to_state(prim_conserve(prim, gamma))


def shock_vortex(ps, gamma=1.4, shock_mach=1.1, shock_x=0.5, vortex_mach=0.9, vortex_r=0.075,
                 vortex_x=0.25, vortex_y=0.5):
    """Stationary normal shock at x = shock_x with Rankine-Hugoniot states and a Lamb-Oseen
    vortex (peak swirl Mach vortex_mach, core radius vortex_r) upstream at (vortex_x, vortex_y):
    v_theta(r) = A (1 - exp(-r^2/rc^2)) / r with A fixing the peak; pressure from the radial
    balance dp/dr = rho v_theta^2 / r integrated isentropically on the upstream state."""
    (r1, u1, p1), (r2, u2, p2) = rankine_hugoniot(shock_mach, gamma)
    x = ps.xpg[..., 0]
    y = ps.xpg[..., 1]
    c1 = np.sqrt(gamma * p1 / r1)
    # Lamb-Oseen profile; peak of (1 - e^{-s})/sqrt(s) sits at s ~ 1.2564 (r ~ 1.1209 rc)
    s_pk = 1.2564312086261696
    g_pk = (1.0 - np.exp(-s_pk)) / np.sqrt(s_pk)
    A = vortex_mach * c1 * vortex_r / g_pk
    dx_ = x - vortex_x
    dy_ = y - vortex_y
    rr = np.sqrt(dx_ ** 2 + dy_ ** 2)
    rs = np.maximum(rr, 1e-14)
    vt = A * (1.0 - np.exp(-(rr / vortex_r) ** 2)) / rs
    # isentropic radial equilibrium: T(r) = T_inf - (gamma-1)/gamma * int_r^inf v^2/r' dr'
    # (T = p / rho); integral evaluated by quadrature on a fine radial grid
    rg = np.linspace(0.0, 40.0 * vortex_r, 40001)
    rg_s = np.maximum(rg, 1e-14)
    vg = A * (1.0 - np.exp(-(rg / vortex_r) ** 2)) / rg_s
    integrand = vg ** 2 / rg_s
    integrand[0] = 0.0
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(rg))])
    tail = cum[-1] - cum
    T_inf = p1 / r1
    Tg = T_inf - (gamma - 1.0) / gamma * tail
    T = np.interp(rr, rg, Tg, right=T_inf)
    rho_v = r1 * (T / T_inf) ** (1.0 / (gamma - 1.0))
    p_v = rho_v * T
    up = np.where(x <= shock_x, 1.0, 0.0)
    rho = np.where(up > 0, rho_v, r2)
    p = np.where(up > 0, p_v, p2)
    U = np.where(up > 0, u1 - vt * dy_ / rs, u2)
    V = np.where(up > 0, vt * dx_ / rs, 0.0)
    prim = np.stack([rho, U, V, rho / (2.0 * p)], axis=-1)
    return _to_state(prim_conserve(prim, gamma))


def perturbed(u, seed: int, amplitude: float = 0.01):
    """Multiplicative per-value perturbation u * (1 + amplitude * U(-1, 1)) with a counter-based
    splitmix64 stream (value n uses counter n in column-major order), keeping rho, p > 0 for
    small amplitudes.  Used to make residual parity inputs non-smooth."""
    n = u.size
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    uni = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    f = 1.0 + amplitude * (2.0 * uni - 1.0)
    out = np.asfortranarray(u.ravel(order="F") * f).reshape(u.shape, order="F")
    return np.asfortranarray(out)


def max_wave_speed(u, gamma):
    """max over interior solution points of max(|U| + c, |V| + c), c = sqrt(gamma p / rho)."""
    w = np.asarray(u)[1:-1, 1:-1]
    rho = w[..., 0]
    U = w[..., 1] / rho
    V = w[..., 2] / rho
    p = (gamma - 1.0) * (w[..., 3] - 0.5 * rho * (U * U + V * V))
    c = np.sqrt(gamma * p / rho)
    return float(np.max(np.maximum(np.abs(U) + c, np.abs(V) + c)))


def cfl_steps(u, gamma, dx, cfl, t_end):
    """Fixed step for t_end: dt = cfl dx / lambda_max rounded down so that t_end = nsteps dt."""
    lam = max_wave_speed(u, gamma)
    n = int(np.ceil(t_end * lam / (cfl * dx)))
    return t_end / n, n


def prim_conserve1(prim, gamma):
    """KitBase prim_conserve for 1D states, prim = [rho, U, lambda]."""
    prim = np.asarray(prim, dtype=np.float64)
    rho, U, lam = prim[..., 0], prim[..., 1], prim[..., 2]
    w = np.empty_like(prim)
    w[..., 0] = rho
    w[..., 1] = rho * U
    w[..., 2] = 0.5 * rho / lam / (gamma - 1.0) + 0.5 * rho * U * U
    return w


def riemann_ensemble(nprob=512, seed=2024, gamma=1.4, ncell_range=(64, 2048), degs=(2, 3, 4)):
    """Seeded ensemble of 1D Riemann problems on [0, 1]: member b has ncell[b] uniform cells of
    degree deg[b]; piecewise-constant states rho, p in [0.1, 1], u in [-0.5, 0.5] left / right of
    an interface x0 in [0.3, 0.7].  As in the reference's Riemann setup
    (example/euler_highlevel.jl:22-31, `ps.x[i] <= 0.5`), every solution point of a cell takes the
    state of the cell's centre, so the jump lies on the cell face nearest x0 and each cell starts
    constant.  Returns (ncell, deg, dx, [u_b]) with u_b(i, p, k)."""
    rng = np.random.default_rng(seed)
    ncell = rng.integers(ncell_range[0], ncell_range[1] + 1, size=nprob).astype(np.int32)
    deg = rng.choice(np.asarray(degs), size=nprob).astype(np.int32)
    states = rng.uniform(size=(nprob, 2, 3))
    x0 = rng.uniform(0.3, 0.7, size=nprob)
    dx = 1.0 / ncell.astype(np.float64)
    members = []
    for b in range(nprob):
        n, d = int(ncell[b]), int(deg[b])
        xc = (np.arange(n) + 0.5) * dx[b]
        h = np.repeat((xc > x0[b]).astype(np.float64)[:, None], d + 1, axis=1)
        mix = lambda k: (1.0 - h) * states[b, 0, k] + h * states[b, 1, k]  # noqa: E731
        rho = 0.1 + 0.9 * mix(0)
        vel = -0.5 + mix(1)
        pr = 0.1 + 0.9 * mix(2)
        prim = np.stack([rho, vel, rho / (2.0 * pr)], axis=-1)
        members.append(np.asfortranarray(prim_conserve1(prim, gamma)))
    return ncell, deg, dx, members
