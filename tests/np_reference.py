"""Independent numpy restatement of the 2D Euler FR residual (TEST INFRASTRUCTURE ONLY).

Written array-at-a-time (einsum over the element index) rather than loop-by-loop like the C
oracle, from the same reference source (example/shock-vortex.jl:26-118), so that an indexing,
sign or operand-order mistake in one of the two does not reappear in the other.  Operators are
computed here from numpy's own Legendre module, not from the oracle or the library.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as npleg


def operators(deg: int):
    n = deg + 1
    r, w = npleg.leggauss(n)
    # Lagrange basis through closed-form barycentric weights
    ll = np.array([np.prod([(-1.0 - r[j]) / (r[k] - r[j]) for j in range(n) if j != k]) for k in range(n)])
    lr = np.array([np.prod([(1.0 - r[j]) / (r[k] - r[j]) for j in range(n) if j != k]) for k in range(n)])
    # dl[m, k] = l_k'(r_m): differentiate the monomial interpolant, V^{-T} route
    Vm = np.vander(r, n, increasing=True)  # Vm[m, q] = r_m^q
    C = np.linalg.inv(Vm)  # column k = monomial coefficients of l_k
    dl = np.zeros((n, n))
    for k in range(n):
        dc = np.polynomial.polynomial.polyder(C[:, k])
        dl[:, k] = np.polynomial.polynomial.polyval(r, dc) if dc.size else 0.0
    # Radau right/left correction g_R = (P_p + P_{p+1})/2, g_L = (-1)^p (P_p - P_{p+1})/2
    cp = np.zeros(n + 1)
    cp[deg] = 1.0
    cp1 = np.zeros(n + 1)
    cp1[deg + 1] = 1.0
    dgr = 0.5 * (npleg.legval(r, npleg.legder(cp)) + npleg.legval(r, npleg.legder(cp1)))
    dgl = (-1.0) ** deg * 0.5 * (npleg.legval(r, npleg.legder(cp)) - npleg.legval(r, npleg.legder(cp1)))
    return dict(r=r, w=w, ll=ll, lr=lr, dl=dl, dhl=dgl, dhr=dgr)


def _pressure(w, g):
    return (g - 1.0) * (w[..., 3] - 0.5 * (w[..., 1] ** 2 + w[..., 2] ** 2) / w[..., 0])


def _fluxes(w, g):
    p = _pressure(w, g)
    u = w[..., 1] / w[..., 0]
    v = w[..., 2] / w[..., 0]
    F = np.stack([w[..., 1], w[..., 1] * u + p, w[..., 2] * u, (w[..., 3] + p) * u], axis=-1)
    G = np.stack([w[..., 2], w[..., 1] * v, w[..., 2] * v + p, (w[..., 3] + p) * v], axis=-1)
    return F, G


def riemann_x(kind, wL, wR, g):
    FL, _ = _fluxes(wL, g)
    FR, _ = _fluxes(wR, g)
    pL, pR = _pressure(wL, g), _pressure(wR, g)
    uL, uR = wL[..., 1] / wL[..., 0], wR[..., 1] / wR[..., 0]
    cL, cR = np.sqrt(g * pL / wL[..., 0]), np.sqrt(g * pR / wR[..., 0])
    if kind == "hll":
        lmin = np.minimum(uL - cL, uR - cR)[..., None]
        lmax = np.maximum(uL + cL, uR + cR)[..., None]
        with np.errstate(divide="ignore", invalid="ignore"):
            h = (lmax * FL - lmin * FR + lmax * lmin * (wR - wL)) / (lmax - lmin)
        return np.where(lmin >= 0, FL, np.where(lmax <= 0, FR, h))
    if kind == "rusanov":
        s = np.maximum(np.abs(uL) + cL, np.abs(uR) + cR)[..., None]
        return 0.5 * (FL + FR) - 0.5 * s * (wR - wL)
    raise ValueError(kind)


def riemann_y(kind, wB, wT, g):
    rot = lambda w: np.stack([w[..., 0], w[..., 2], -w[..., 1], w[..., 3]], axis=-1)  # noqa: E731
    f = riemann_x(kind, rot(wB), rot(wT), g)
    return np.stack([f[..., 0], -f[..., 2], f[..., 1], f[..., 3]], axis=-1)


def fr2d_rhs(u, ops, g, iJ11, iJ22, flux="hll", periodic=False):
    """u: (nx+2, ny+2, n, n, 4) reference layout (ghost layer read as data unless periodic)."""
    u = np.array(u, dtype=np.float64)
    nx, ny = u.shape[0] - 2, u.shape[1] - 2
    if periodic:
        u[0, 1:-1] = u[nx, 1:-1]
        u[nx + 1, 1:-1] = u[1, 1:-1]
        u[:, 0] = u[:, ny]
        u[:, ny + 1] = u[:, 1]
    ll, lr, dl, dhl, dhr = ops["ll"], ops["lr"], ops["dl"], ops["dhl"], ops["dhr"]
    F, G = _fluxes(u, g)  # (..., k, l, m)
    f1 = iJ11 * F
    f2 = iJ22 * G
    # traces: x-faces along k (index 2), y-faces along l (index 3)
    uL = np.einsum("ijklm,k->ijlm", u, ll)
    uR = np.einsum("ijklm,k->ijlm", u, lr)
    uB = np.einsum("ijklm,l->ijkm", u, ll)
    uT = np.einsum("ijklm,l->ijkm", u, lr)
    f1L = np.einsum("ijklm,k->ijlm", f1, ll)
    f1R = np.einsum("ijklm,k->ijlm", f1, lr)
    f2B = np.einsum("ijklm,l->ijkm", f2, ll)
    f2T = np.einsum("ijklm,l->ijkm", f2, lr)
    # common fluxes: Fx[a] between cells a-1 and a (a = 1..nx+1), rows j = 1..ny
    Fx = riemann_x(flux, uR[0:nx + 1, 1:ny + 1], uL[1:nx + 2, 1:ny + 1], g)  # (nx+1, ny, l, m)
    Gy = riemann_y(flux, uT[1:nx + 1, 0:ny + 1], uB[1:nx + 1, 1:ny + 2], g)  # (nx, ny+1, k, m)
    I = slice(1, nx + 1)
    J = slice(1, ny + 1)
    d1 = np.einsum("kq,ijqlm->ijklm", dl, f1[I, J])
    d2 = np.einsum("lq,ijkqm->ijklm", dl, f2[I, J])
    cl = Fx[0:nx] * iJ11 - f1L[I, J]  # (nx, ny, l, m)
    cr = Fx[1:nx + 1] * iJ11 - f1R[I, J]
    cb = Gy[:, 0:ny] * iJ22 - f2B[I, J]  # (nx, ny, k, m)
    ct = Gy[:, 1:ny + 1] * iJ22 - f2T[I, J]
    corr = (dhl[None, None, :, None, None] * cl[:, :, None, :, :]
            + dhr[None, None, :, None, None] * cr[:, :, None, :, :]
            + dhl[None, None, None, :, None] * cb[:, :, :, None, :]
            + dhr[None, None, None, :, None] * ct[:, :, :, None, :])
    du = np.zeros_like(u)
    du[I, J] = -(d1 + d2 + corr)
    return np.asfortranarray(du)


def fr2d_rhs_metric(u, ops, g, iJ, flux="hll", periodic=False):
    """The same residual with a 2x2 inverse Jacobian per solution point, iJ of shape
    (nx+2, ny+2, n, n, 2, 2) (example/shock-vortex.jl:48-55 and :102-115 read iJ[i, j][k, l])."""
    u = np.array(u, dtype=np.float64)
    nx, ny = u.shape[0] - 2, u.shape[1] - 2
    if periodic:
        u[0, 1:-1] = u[nx, 1:-1]
        u[nx + 1, 1:-1] = u[1, 1:-1]
        u[:, 0] = u[:, ny]
        u[:, ny + 1] = u[:, 1]
    ll, lr, dl, dhl, dhr = ops["ll"], ops["lr"], ops["dl"], ops["dhl"], ops["dhr"]
    F, G = _fluxes(u, g)
    a = iJ[..., None, :, :]  # (..., k, l, 1, 2, 2) against the variables
    f1 = a[..., 0, 0] * F + a[..., 0, 1] * G
    f2 = a[..., 1, 0] * F + a[..., 1, 1] * G
    uL = np.einsum("ijklm,k->ijlm", u, ll)
    uR = np.einsum("ijklm,k->ijlm", u, lr)
    uB = np.einsum("ijklm,l->ijkm", u, ll)
    uT = np.einsum("ijklm,l->ijkm", u, lr)
    f1L = np.einsum("ijklm,k->ijlm", f1, ll)
    f1R = np.einsum("ijklm,k->ijlm", f1, lr)
    f2B = np.einsum("ijklm,l->ijkm", f2, ll)
    f2T = np.einsum("ijklm,l->ijkm", f2, lr)
    Fx = riemann_x(flux, uR[0:nx + 1, 1:ny + 1], uL[1:nx + 2, 1:ny + 1], g)
    Gy = riemann_y(flux, uT[1:nx + 1, 0:ny + 1], uB[1:nx + 1, 1:ny + 2], g)
    I = slice(1, nx + 1)
    J = slice(1, ny + 1)
    d1 = np.einsum("kq,ijqlm->ijklm", dl, f1[I, J])
    d2 = np.einsum("lq,ijkqm->ijklm", dl, f2[I, J])
    j11 = iJ[I, J, :, :, 0, 0][..., None]  # (nx, ny, k, l, 1)
    j22 = iJ[I, J, :, :, 1, 1][..., None]
    cl = Fx[0:nx][:, :, None] * j11 - f1L[I, J][:, :, None]  # (nx, ny, k, l, m)
    cr = Fx[1:nx + 1][:, :, None] * j11 - f1R[I, J][:, :, None]
    cb = Gy[:, 0:ny][:, :, :, None] * j22 - f2B[I, J][:, :, :, None]
    ct = Gy[:, 1:ny + 1][:, :, :, None] * j22 - f2T[I, J][:, :, :, None]
    corr = dhl[None, None, :, None, None] * cl + dhr[None, None, :, None, None] * cr \
        + dhl[None, None, None, :, None] * cb + dhr[None, None, None, :, None] * ct
    du = np.zeros_like(u)
    du[I, J] = -(d1 + d2 + corr)
    return np.asfortranarray(du)


def exact_riemann(wl, wr, g, xi):
    """Exact solution of the 1D Euler Riemann problem (Toro, Riemann Solvers and Numerical Methods
    for Fluid Dynamics, ch. 4) at the similarity coordinates xi = (x - x0) / t.  wl, wr =
    (rho, u, p) of the two states; returns rho, u, p arrays shaped like xi.  The star pressure is
    the root of f_L(p) + f_R(p) + u_R - u_L (shock branch for p above the side's pressure,
    rarefaction branch below), found by bisection to round-off."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    gm, gp = g - 1.0, g + 1.0

    def fk(p, rk, pk, ck):
        if p > pk:
            a, b = 2.0 / (gp * rk), gm / gp * pk
            return (p - pk) * np.sqrt(a / (p + b))
        return 2.0 * ck / gm * ((p / pk) ** (gm / (2 * g)) - 1.0)

    f = lambda p: fk(p, rl, pl, cl) + fk(p, rr, pr, cr) + ur - ul  # noqa: E731
    lo, hi = 1e-14, max(pl, pr)
    while f(hi) < 0:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, cr) - fk(ps, rl, pl, cl))
    xi = np.asarray(xi, dtype=np.float64)
    rho, u, p = np.empty_like(xi), np.empty_like(xi), np.empty_like(xi)
    for n, s in np.ndenumerate(xi):
        if s <= us:
            if ps > pl:
                sl = ul - cl * np.sqrt(gp / (2 * g) * ps / pl + gm / (2 * g))
                w = (rl, ul, pl) if s <= sl else (rl * (ps / pl + gm / gp) / (gm / gp * ps / pl + 1.0), us, ps)
            else:
                csl = cl * (ps / pl) ** (gm / (2 * g))
                if s <= ul - cl:
                    w = (rl, ul, pl)
                elif s > us - csl:
                    w = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / gp * (cl + gm / 2 * (ul - s))
                    w = (rl * (c / cl) ** (2.0 / gm), 2.0 / gp * (cl + gm / 2 * ul + s), pl * (c / cl) ** (2 * g / gm))
        else:
            if ps > pr:
                sr = ur + cr * np.sqrt(gp / (2 * g) * ps / pr + gm / (2 * g))
                w = (rr, ur, pr) if s >= sr else (rr * (ps / pr + gm / gp) / (gm / gp * ps / pr + 1.0), us, ps)
            else:
                csr = cr * (ps / pr) ** (gm / (2 * g))
                if s >= ur + cr:
                    w = (rr, ur, pr)
                elif s <= us + csr:
                    w = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / gp * (cr - gm / 2 * (ur - s))
                    w = (rr * (c / cr) ** (2.0 / gm), 2.0 / gp * (-cr + gm / 2 * ur + s), pr * (c / cr) ** (2 * g / gm))
        rho[n], u[n], p[n] = w
    return rho, u, p, ps, us
