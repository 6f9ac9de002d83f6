"""Host mirror of the reference's operator API and FR spaces, computed by libfrhip's native host
routines (include/fr.h, "operator data").

Names follow src/FluxReconstruction.jl:19-51 (Julia symbols with unicode prefixes are spelled
out: ``∂lagrange`` -> ``dlagrange``, ``∂radau`` -> ``dradau`` ...).  Arrays are numpy, indices
0-based, multi-dimensional arrays keep the reference's column-major (Fortran) layout.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, dptr


class Line:  # element-shape tags (src/data.jl:4-11)
    pass


class Quad:
    pass


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


# ------------------------------------------------------------------ polynomials (src/Polynomial)
def legendre_point(p: int) -> np.ndarray:
    """Gauss-Legendre points of degree p (src/Polynomial/poly_legendre.jl:6)."""
    r = np.zeros(p + 1)
    w = np.zeros(p + 1)
    check(_lib.lib().fr_legendre_point(p, dptr(r), dptr(w)))
    return r


def gausslegendre(n: int):
    r = np.zeros(n)
    w = np.zeros(n)
    check(_lib.lib().fr_legendre_point(n - 1, dptr(r), dptr(w)))
    return r, w


def lagrange_point(sp, x):
    """Lagrange basis of solution points sp at x (src/Polynomial/poly_lagrange.jl:6-31)."""
    sp = _f(sp)
    if np.ndim(x) == 0:
        out = np.zeros(sp.size)
        check(_lib.lib().fr_lagrange_point(dptr(sp), sp.size, float(x), dptr(out)))
        return out
    return np.stack([lagrange_point(sp, xi) for xi in np.asarray(x)], axis=0)


def dlagrange(sp) -> np.ndarray:
    """lpdm[m, k] = l_k'(sp_m) (src/Polynomial/poly_lagrange.jl:38-59)."""
    sp = _f(sp)
    out = np.zeros(sp.size * sp.size)
    check(_lib.lib().fr_dlagrange(dptr(sp), sp.size, dptr(out)))
    return out.reshape((sp.size, sp.size), order="F")


def standard_lagrange(x):
    """(ll, lr, lpdm) (src/Polynomial/poly_lagrange.jl:99-105)."""
    return lagrange_point(x, -1.0), lagrange_point(x, 1.0), dlagrange(x)


def dlegendre(p: int, x):
    """∂legendre(p, x) (src/Polynomial/poly_legendre.jl:13-22): P_p'(x) at a point or an array."""
    scalar = np.ndim(x) == 0
    xs = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)).ravel())
    out = np.zeros_like(xs)
    check(_lib.lib().fr_dlegendre(int(p), dptr(xs), xs.size, dptr(out)))
    return float(out[0]) if scalar else out.reshape(np.shape(x))


def _dcorr(kind: str, p: int, x):
    x = _f(np.atleast_1d(x))
    dgl = np.zeros_like(x)
    dgr = np.zeros_like(x)
    check(_lib.lib().fr_dcorrection(_lib.CORR[kind], p, dptr(x), x.size, dptr(dgl), dptr(dgr)))
    return dgl, dgr


def dradau(p: int, x):
    """Radau correction derivatives (src/Polynomial/poly_legendre.jl:29-37)."""
    return _dcorr("radau", p, x)


def dsd(p: int, x):
    """Spectral-difference correction derivatives (src/Polynomial/poly_legendre.jl:44-54)."""
    return _dcorr("sd", p, x)


def dhuynh(p: int, x):
    """Huynh g2 correction derivatives (src/Polynomial/poly_legendre.jl:61-71)."""
    return _dcorr("huynh", p, x)


def JacobiP(x, alpha, beta, N):
    """Orthonormal Jacobi polynomial (src/Polynomial/poly_jacobi.jl:7-83)."""
    xs = _f(np.atleast_1d(x))
    out = np.zeros_like(xs)
    check(_lib.lib().fr_jacobi_p(dptr(xs), xs.size, float(alpha), float(beta), int(N), dptr(out)))
    return out if np.ndim(x) else float(out[0])


def dJacobiP(x, alpha, beta, N):
    """src/Polynomial/poly_jacobi.jl:88-107."""
    xs = _f(np.atleast_1d(x))
    out = np.zeros_like(xs)
    check(_lib.lib().fr_djacobi_p(dptr(xs), xs.size, float(alpha), float(beta), int(N), dptr(out)))
    return out if np.ndim(x) else float(out[0])


def vandermonde_matrix(*args):
    """vandermonde_matrix(N, r), (Line, N, r) or (Quad, N, r, s) (src/Transform/transform.jl:18-68)."""
    if args[0] is Line:
        args = args[1:]
    if args[0] is Quad:
        _, N, r, s = args
        r = _f(r)
        s = _f(s)
        V = np.zeros(r.size * (N + 1) ** 2)
        check(_lib.lib().fr_vandermonde_quad(N, dptr(r), dptr(s), r.size, dptr(V)))
        return V.reshape((r.size, (N + 1) ** 2), order="F")
    N, r = args
    r = _f(r)
    V = np.zeros(r.size * (N + 1))
    check(_lib.lib().fr_vandermonde_1d(N, dptr(r), r.size, dptr(V)))
    return V.reshape((r.size, N + 1), order="F")


def dvandermonde_matrix(*args):
    """r-derivative of the 1D Vandermonde matrix (src/Transform/transform.jl:76-89)."""
    if args[0] is Line:
        args = args[1:]
    N, r = args
    r = _f(r)
    V = np.zeros(r.size * (N + 1))
    check(_lib.lib().fr_dvandermonde_1d(N, dptr(r), r.size, dptr(V)))
    return V.reshape((r.size, N + 1), order="F")


def inv(A) -> np.ndarray:
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    out = np.zeros(n * n)
    check(_lib.lib().fr_inv(n, dptr(_f(A)), dptr(out)))
    return out.reshape((n, n), order="F")


def _rs_jacobi_point(r: float, s: float, vertices) -> np.ndarray:
    J = np.zeros(4)
    check(_lib.lib().fr_rs_jacobi_quad(float(r), float(s), dptr(_f(vertices)), dptr(J)))
    return J.reshape((2, 2), order="F")


def rs_jacobi(*args) -> np.ndarray:
    """Bilinear quadrilateral Jacobians [xr xs; yr ys] (src/Geometry/geo_jacobi.jl:77-108), in the
    reference's method forms, with a 4 x 2 vertex matrix (one cell) or a (ncx, ncy, 4, 2) vertex
    array (every cell, leading axes (ncx, ncy)); results end in the 2 x 2 matrix:
    - rs_jacobi(r, s, vertices), scalars r, s: one matrix (:77-88);
    - rs_jacobi(r, s, vertices), vectors r, s: one matrix per pair (r[i], s[i]) (:90-91);
    - rs_jacobi(r, s, vertices), matrices r (A x B) and s (C x D): the reference's comprehension
      over i in axes(r, 1), j in axes(s, 2) at (r[i], s[i]) with Julia's linear (column-major)
      index i, an A x D array (:93-94);
    - rs_jacobi(r, vertices), a vector r: the tensor form J[i, j] at (r[i], r[j]) (:102-103)."""
    if len(args) == 2:
        r, vertices = np.asarray(args[0], dtype=np.float64), np.asarray(args[1], dtype=np.float64)
        if vertices.shape == (4, 2):
            J, _ = rs_jacobi_mesh(vertices[None, None], r)
            return J[0, 0]
        J, _ = rs_jacobi_mesh(vertices, r)
        return J
    if len(args) != 3:
        raise TypeError("rs_jacobi(r, s, vertices) or rs_jacobi(r, vertices)")
    r, s, vertices = args
    vertices = np.asarray(vertices, dtype=np.float64)
    if vertices.ndim == 4:
        return np.stack([np.stack([rs_jacobi(r, s, vertices[i, j]) for j in range(vertices.shape[1])])
                         for i in range(vertices.shape[0])])
    if np.ndim(r) == 0:
        return _rs_jacobi_point(r, s, vertices)
    r, s = np.asarray(r, dtype=np.float64), np.asarray(s, dtype=np.float64)
    if r.ndim == 1:
        return np.stack([_rs_jacobi_point(r[i], s[i], vertices) for i in range(r.size)])
    rl, sl = r.ravel(order="F"), s.ravel(order="F")
    out = np.zeros((r.shape[0], s.shape[1], 2, 2))
    for i in range(r.shape[0]):
        out[i, :] = _rs_jacobi_point(rl[i], sl[i], vertices)[None]
    return out


def basis_norm(deg: int) -> np.ndarray:
    """Discrete L1 norms of the orthonormal Legendre basis P_0 .. P_deg on 100 uniform points of
    [-1, 1], dx sum |P_i| (src/Polynomial/poly_filter.jl:54-66)."""
    x = np.linspace(-1.0, 1.0, 100)
    dx = x[1] - x[0]
    return np.array([dx * np.sum(np.abs(JacobiP(x, 0.0, 0.0, i))) for i in range(deg + 1)])


def shock_detector(Se: float, deg: int, S0: float | None = None, kappa: float = 4.0) -> bool:
    """Persson-Peraire switch (src/dissipation.jl:13-23)."""
    if S0 is None:
        S0 = -3.0 * np.log10(deg)
    return bool(_lib.lib().fr_shock_detector(float(Se), int(deg), float(S0), float(kappa)))


def modal_filter(u, lam: float, filter: str = "l2"):
    """KitBase's modal_filter!(u, lambda; filter = :l2) on a modal vector, in place (the call of
    example/shock-vortex.jl:316 and example/euler_highlevel.jl:45 on iV * u of an element): mode
    k (0-based) is divided by 1 + lambda (k+1)^2 k^2.  Only the :l2 filter is built."""
    if str(filter).lstrip(":") != "l2":
        raise ValueError("modal_filter: only filter = l2 is built")
    if not isinstance(u, np.ndarray) or u.dtype != np.float64 or u.ndim != 1:
        raise ValueError("u must be a float64 numpy vector (filtered in place)")
    a = np.ascontiguousarray(u)
    check(_lib.lib().fr_modal_filter_l2(dptr(a), a.size, float(lam)))
    if a is not u:
        u[...] = a
    return None


def positive_limiter(u, *args):
    """positive_limiter of src/dissipation.jl on one element, in place (the reference's three
    methods, chosen by u's dimension like Julia's dispatch), computed by libfrhip's native host
    routine; raises FRError (FR_E_LIMITER) where the reference's @assert fails, leaving u as it was.

    positive_limiter(u, weights, ll, lr[, t0]): scalar element u(nsp) (:37-58; weights an array
    or one number, as the reference accepts);
    positive_limiter(u, gamma, weights, ll, lr[, t0]): 1D Euler u(nsp, 3) (:61-123) or 2D Euler
    u(nsp, nsp, 4) (:125-206), weights summing to one.  u must be a float64 numpy array; 2D and
    matrix elements are taken in column-major order."""
    if not isinstance(u, np.ndarray) or u.dtype != np.float64:
        raise ValueError("u must be a float64 numpy array (limited in place)")
    L = _lib.lib()
    if u.ndim == 1:
        weights, ll, lr = args[:3]
        n = u.size
        scalar = np.ndim(weights) == 0
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, dtype=np.float64), (n,)))
        wsum = float(weights) if scalar else float(np.sum(w))
        a = np.ascontiguousarray(u)
        check(L.fr_positive_limiter_scalar(dptr(a), n, dptr(w), wsum, dptr(_f(ll)), dptr(_f(lr))))
    elif u.ndim in (2, 3):
        g, weights, ll, lr = args[:4]
        t0 = float(args[4]) if len(args) > 4 else 1.0
        nsp = u.shape[0]
        if (u.ndim == 2 and u.shape != (nsp, 3)) or (u.ndim == 3 and u.shape != (nsp, nsp, 4)):
            raise ValueError("u must be (nsp, 3) (1D Euler) or (nsp, nsp, 4) (2D Euler)")
        a = np.asfortranarray(u)
        w = np.asfortranarray(weights, dtype=np.float64)
        fn = L.fr_positive_limiter_euler1d if u.ndim == 2 else L.fr_positive_limiter_euler2d
        check(fn(dptr(a), nsp, float(g), dptr(w), dptr(_f(ll)), dptr(_f(lr)), t0))
    else:
        raise ValueError("u must be an element of 1, 2 or 3 dimensions")
    if a is not u:
        u[...] = a
    return None


# ------------------------------------------------------------------ metrics (src/tools.jl:43-53)
def L1_error(u, ue, dx):
    return float(np.sum(np.abs(np.asarray(u) - np.asarray(ue)) * dx))


def L2_error(u, ue, dx):
    return float(np.sqrt(np.sum((np.abs(np.asarray(u) - np.asarray(ue)) * dx) ** 2)))


def Linf_error(u, ue, dx):
    return float(np.max(np.abs(np.asarray(u) - np.asarray(ue)) * dx))


# ------------------------------------------------------------------ meshes (KitBase PSpace1D/2D)
class PSpace1D:
    """Uniform 1D mesh with ng ghost cells; x[i] for i = 1-ng .. nx+ng stored at i-1+ng."""

    def __init__(self, x0, x1, nx, ng=0):
        self.x0, self.x1, self.nx, self.ng = float(x0), float(x1), int(nx), int(ng)
        d = (self.x1 - self.x0) / nx
        idx = np.arange(1 - ng, nx + ng + 1)
        self.x = self.x0 + (idx - 0.5) * d
        self.dx = np.full(idx.size, d)


class PSpace2D:
    """Uniform 2D mesh with (ngx, ngy) ghost layers, KitBase conventions: centres x[i, j],
    y[i, j] and vertices[i, j, v, d] (v = 1 bottom-left, 2 bottom-right, 3 top-right,
    4 top-left; src/Geometry/geo_jacobi.jl:50-57)."""

    def __init__(self, x0, x1, nx, y0, y1, ny, ngx=0, ngy=0):
        self.x0, self.x1, self.nx = float(x0), float(x1), int(nx)
        self.y0, self.y1, self.ny = float(y0), float(y1), int(ny)
        self.ngx, self.ngy = int(ngx), int(ngy)
        self.dx0 = (self.x1 - self.x0) / nx
        self.dy0 = (self.y1 - self.y0) / ny
        ii = np.arange(1 - ngx, nx + ngx + 1)
        jj = np.arange(1 - ngy, ny + ngy + 1)
        X, Y = np.meshgrid(self.x0 + (ii - 0.5) * self.dx0, self.y0 + (jj - 0.5) * self.dy0, indexing="ij")
        self.x = np.asfortranarray(X)
        self.y = np.asfortranarray(Y)
        self.dx = np.full_like(self.x, self.dx0)
        self.dy = np.full_like(self.y, self.dy0)
        v = np.zeros(X.shape + (4, 2), order="F")
        hx, hy = self.dx0 / 2, self.dy0 / 2
        v[..., 0, 0], v[..., 0, 1] = X - hx, Y - hy
        v[..., 1, 0], v[..., 1, 1] = X + hx, Y - hy
        v[..., 2, 0], v[..., 2, 1] = X + hx, Y + hy
        v[..., 3, 0], v[..., 3, 1] = X - hx, Y + hy
        self.vertices = v


class CSpace2D:
    """Polar (curvilinear) 2D mesh, KitBase's CSpace2D(r0, r1, nr, θ0, θ1, nθ, ngr, ngθ) as the
    reference's curvilinear scripts build it (dev/cylinder.jl:23, dev/curve_euler.jl:24): cells of
    constant radial width dr and angle dθ with ngr / ngθ ghost layers; centres r[i, j], θ[i, j],
    x = r cos θ, y = r sin θ; dr, dθ and the arc length darc = r dθ per cell (the scripts read
    ps.nr, ps.nθ, ps.dθ[1, j], ps.x, ps.y); vertices[i, j, v, d] in the order of PSpace2D with r
    in place of x and θ in place of y (v = 1 (r-, θ-), 2 (r+, θ-), 3 (r+, θ+), 4 (r-, θ+)).  KitBase
    is not part of the reference tree, so these conventions follow its use in the scripts and are
    not pinned against KitBase itself.  nx, ny, ngx, ngy alias nr, nθ, ngr, ngθ for the solvers."""

    def __init__(self, r0, r1, nr, θ0, θ1, nθ, ngr=0, ngθ=0):
        self.r0, self.r1, self.nr = float(r0), float(r1), int(nr)
        self.θ0, self.θ1, self.nθ = float(θ0), float(θ1), int(nθ)
        self.ngr, self.ngθ = int(ngr), int(ngθ)
        dr, dθ = (self.r1 - self.r0) / nr, (self.θ1 - self.θ0) / nθ
        ii = np.arange(1 - ngr, nr + ngr + 1)
        jj = np.arange(1 - ngθ, nθ + ngθ + 1)
        R, T = np.meshgrid(self.r0 + (ii - 0.5) * dr, self.θ0 + (jj - 0.5) * dθ, indexing="ij")
        self.r = np.asfortranarray(R)
        self.θ = np.asfortranarray(T)
        self.x = np.asfortranarray(R * np.cos(T))
        self.y = np.asfortranarray(R * np.sin(T))
        self.dr = np.full_like(self.r, dr)
        self.dθ = np.full_like(self.θ, dθ)
        self.darc = np.asfortranarray(R * dθ)
        v = np.zeros(R.shape + (4, 2), order="F")
        for q, (sr, st) in enumerate(((-1, -1), (1, -1), (1, 1), (-1, 1))):
            rv, tv = R + sr * dr / 2, T + st * dθ / 2
            v[..., q, 0], v[..., q, 1] = rv * np.cos(tv), rv * np.sin(tv)
        self.vertices = v

    nx = property(lambda self: self.nr)
    ny = property(lambda self: self.nθ)
    ngx = property(lambda self: self.ngr)
    ngy = property(lambda self: self.ngθ)


def r_x(r, vl, vr):
    """The point r of the standard element [-1, 1] on the real element [vl, vr]
    (src/Geometry/geo_transform.jl:7)."""
    return ((1.0 - r) / 2.0) * vl + ((1.0 + r) / 2.0) * vr


def global_sp(*args):
    """Global coordinates of the solution points (src/Geometry/geo_points.jl:8-41).

    global_sp(xi, r): line elements between the n + 1 faces xi, shape (n, len(r)).
    global_sp(xi, yi, r): quadrilaterals of the vertex arrays xi, yi of shape (nx + 1, ny + 1);
    returns (xp, yp) of shape (nx, ny, len(r), len(r)): xp[i, j, k, l] = r_x(r[k], xi[i, j],
    xi[i+1, j]) and yp[i, j, k, l] = r_x(r[l], yi[i, j], yi[i, j+1]).  The reference's loop runs j
    over all ny + 1 vertex columns, whose last one reads yi[i, ny + 2] past the end of the array;
    here j covers the ny cells."""
    if len(args) == 2:
        xi = np.asarray(args[0], dtype=np.float64)
        r = np.asarray(args[1], dtype=np.float64)
        return np.asfortranarray(r_x(r[None, :], xi[:-1, None], xi[1:, None]))
    if len(args) == 3:
        xi = np.asarray(args[0], dtype=np.float64)
        yi = np.asarray(args[1], dtype=np.float64)
        r = np.asarray(args[2], dtype=np.float64)
        if xi.ndim != 2 or xi.shape != yi.shape:
            raise ValueError("xi and yi must be vertex arrays of one shape (nx + 1, ny + 1)")
        xp = r_x(r[None, None, :, None], xi[:-1, :-1, None, None], xi[1:, :-1, None, None])
        yp = r_x(r[None, None, None, :], yi[:-1, :-1, None, None], yi[:-1, 1:, None, None])
        n = r.size
        shape = (xi.shape[0] - 1, xi.shape[1] - 1, n, n)
        return np.asfortranarray(np.broadcast_to(xp, shape)), np.asfortranarray(np.broadcast_to(yp, shape))
    raise TypeError("global_sp(xi, r) or global_sp(xi, yi, r)")


def rs_jacobi_mesh(vertices, r):
    """J[i, j][k, l] = rs_jacobi(r[k], r[l], vertices[i, j]) of every cell and solution point and
    its inverse (src/struct.jl:134-142, geo_jacobi.jl:90-98), as (ncx, ncy, n, n, 2, 2) arrays."""
    v = np.asfortranarray(vertices, dtype=np.float64)
    if v.ndim != 4 or v.shape[2:] != (4, 2):
        raise ValueError("vertices must have shape (ncx, ncy, 4, 2)")
    r = np.ascontiguousarray(r, dtype=np.float64)
    ncx, ncy, n = v.shape[0], v.shape[1], r.size
    J = np.zeros(ncx * ncy * n * n * 4)
    iJ = np.zeros_like(J)
    check(_lib.lib().fr_rs_jacobi_mesh(ncx, ncy, dptr(_f(v)), n, dptr(r), dptr(J), dptr(iJ)))
    shape = (ncx, ncy, n, n, 2, 2)
    return J.reshape(shape, order="F"), iJ.reshape(shape, order="F")


class _Forward:
    """Forward KitBase mesh fields through the FR space (src/tools.jl:5-37)."""

    _base_fields = ("x0", "x1", "nx", "y0", "y1", "ny", "ng", "ngx", "ngy", "x", "y", "dx", "dy", "dx0", "dy0",
                    "vertices", "r0", "r1", "nr", "θ0", "θ1", "nθ", "ngr", "ngθ", "r", "θ", "dr", "dθ", "darc")

    def __getattr__(self, name):
        if name in self._base_fields and "base" in self.__dict__:
            return getattr(self.__dict__["base"], name)
        raise AttributeError(name)


class FRPSpace1D(_Forward):
    """1D FR space (src/struct.jl:13-88): J, xpl, xpg, wp, dl, ll, lr, dll, dlr, dhl, dhr, V, iV.
    Positional or keyword arguments; other keywords are ignored, as the reference's keyword
    constructor FRPSpace1D(; x0, x1, nx, deg, ng, correction, kwargs...) (struct.jl:87-88) ignores
    them (example/advection_lowlevel.jl:51 passes its whole configuration)."""

    def __init__(self, x0, x1, nx, deg, ng=0, correction="radau", **kwargs):
        self.base = PSpace1D(x0, x1, nx, ng)
        self.deg = int(deg)
        self.np = deg + 1
        self.correction = correction
        self.J = self.base.dx / 2.0
        r, w = gausslegendre(deg + 1)
        self.xpl = r
        self.wp = w
        xi = np.append(self.base.x - 0.5 * self.base.dx, self.base.x[-1] + 0.5 * self.base.dx[-1])
        self.xpg = np.asfortranarray(((1.0 - r[None, :]) / 2.0) * xi[:-1, None] + ((1.0 + r[None, :]) / 2.0) * xi[1:, None])
        self.ll, self.lr, self.dl = standard_lagrange(r)
        self.V = vandermonde_matrix(deg, r)
        self.iV = inv(self.V)
        dVf = dvandermonde_matrix(deg, np.array([-1.0, 1.0]))
        self.dll = np.linalg.solve(self.V.T, dVf[0, :])
        self.dlr = np.linalg.solve(self.V.T, dVf[1, :])
        self.dhl, self.dhr = _dcorr(correction, deg, r)


class FRPSpace2D(_Forward):
    """2D FR space (src/struct.jl:99-245): FRPSpace2D(x0, x1, nx, y0, y1, ny, deg, ngx, ngy) on
    uniform rectangles, or FRPSpace2D(base, deg) over any mesh with a vertices[i, j, v, d] array
    (PSpace2D, CSpace2D, ...).

    Fields: deg, np, xpl, xpg(i, j, k, l, d), wp(k, l), dl, ll, lr, dll, dlr, dhl, dhr, V, iV
    (tensor Vandermonde, modes i-outer / j-inner) and the metrics.  On uniform rectangles
    (`uniform`) J / iJ are one 2x2 matrix per element (diag(dx/2, dy/2) and its inverse, exposed
    as (nx+2g, ny+2g, 2, 2) views; J_elem / iJ_elem).  Over any other mesh they are per solution
    point as in the reference, J[i, j, k, l] = rs_jacobi(r[k], r[l], vertices[i, j]) with iJ its
    inverse, arrays of shape (nx+2g, ny+2g, n, n, 2, 2), and Ji the Jacobians the reference forms
    at its interface points (struct.jl:144-158).  Keyword construction ignores other keywords, as
    the reference's FRPSpace2D(; x0, x1, nx, y0, y1, ny, deg, ngx, ngy, kwargs...) does
    (struct.jl:244-245)."""

    def __init__(self, x0, x1=None, nx=None, y0=None, y1=None, ny=None, deg=None, ngx=0, ngy=0, **kwargs):
        if hasattr(x0, "vertices"):  # FRPSpace2D(base, deg)
            base, deg = x0, x1
        else:
            base = PSpace2D(x0, x1, nx, y0, y1, ny, ngx, ngy)
        self.base = base
        self.deg = int(deg)
        self.np = (deg + 1) ** 2
        n = deg + 1
        r, w = gausslegendre(n)
        self.xpl = r
        self.wp = np.asfortranarray(np.outer(w, w))
        shape = base.x.shape
        self.uniform = isinstance(base, PSpace2D)
        if self.uniform:
            Jc = rs_jacobi(0.0, 0.0, base.vertices[0, 0])
            self.J_elem = Jc
            self.iJ_elem = inv(Jc)
            self.J = np.broadcast_to(Jc, shape + (2, 2))
            self.iJ = np.broadcast_to(self.iJ_elem, shape + (2, 2))
        else:
            self.J_elem = self.iJ_elem = None
            self.J, self.iJ = rs_jacobi_mesh(base.vertices, r)
        v = base.vertices
        rk = r[:, None]
        rl = r[None, :]
        xpg = np.zeros(shape + (n, n, 2), order="F")
        for d in range(2):
            xpg[..., d] = ((rk - 1.0) * (rl - 1.0) / 4)[None, None] * v[:, :, 0, d, None, None] \
                + ((rk + 1.0) * (1.0 - rl) / 4)[None, None] * v[:, :, 1, d, None, None] \
                + ((rk + 1.0) * (rl + 1.0) / 4)[None, None] * v[:, :, 2, d, None, None] \
                + ((1.0 - rk) * (rl + 1.0) / 4)[None, None] * v[:, :, 3, d, None, None]
        self.xpg = xpg
        self.ll, self.lr, self.dl = standard_lagrange(r)
        V1 = vandermonde_matrix(deg, r)
        dVf = dvandermonde_matrix(deg, np.array([-1.0, 1.0]))
        self.dll = np.linalg.solve(V1.T, dVf[0, :])
        self.dlr = np.linalg.solve(V1.T, dVf[1, :])
        self.dhl, self.dhr = dradau(deg, r)
        r2d = np.repeat(r[:, None], n, axis=1)
        s2d = np.repeat(r[None, :], n, axis=0)
        self.V = vandermonde_matrix(Quad, deg, r2d.ravel(order="F"), s2d.ravel(order="F"))
        self.iV = inv(self.V)

    @property
    def nsp(self) -> int:
        return self.deg + 1

    @property
    def Ji(self) -> np.ndarray:
        """The Jacobians the reference forms at its interface points (struct.jl:144-158), shape
        (nx+2g, ny+2g, 4, n, 2, 2).  rs_jacobi(ri, si, vertices) there takes the linear index f over
        the first axis of the 4 x n point matrices only (geo_jacobi.jl:93-94), so Ji[i, j][f, q] is
        the Jacobian at (ri[f, 1], si[f, 1]) for every q: (r_1, 0), (1, r_1), (r_n, 1), (0, r_n).
        Computed on first use; no kernel reads it."""
        if getattr(self, "_Ji", None) is None:
            r, n, v = self.xpl, self.deg + 1, self.base.vertices
            pts = ((r[0], 0.0), (1.0, r[0]), (r[-1], 1.0), (0.0, r[-1]))
            Ji = np.zeros(v.shape[:2] + (4, n, 2, 2), order="F")
            for f, (rr, ss) in enumerate(pts):
                J, _ = rs_jacobi_mesh(v, np.array([rr, ss]))
                # the tensor grid's point (k, l) = (0, 1) is (r[0], r[1]) = (rr, ss)
                Ji[:, :, f, :] = J[:, :, 0, 1][:, :, None]
            self._Ji = Ji
        return self._Ji
