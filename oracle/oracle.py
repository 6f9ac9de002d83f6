"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the reference results for tests/ and __graft_entry__.smoke(), and the
cpu_baseline leg of bench.py.  The product package never imports this module.
Arrays follow the reference's Julia column-major layout (see fr_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FLUX = {"hll": 0, "rusanov": 1, "roe": 2}
BC = {"ghost": 0, "periodic": 1}
CORR = {"radau": 0, "sd": 1, "huynh": 2}

_dp = C.POINTER(C.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_legendre_p.restype = C.c_double
        L.or_dlegendre.restype = C.c_double
        L.or_jacobi_p.restype = C.c_double
        L.or_djacobi_p.restype = C.c_double
        L.or_jacobi_p.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_djacobi_p.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_legendre_p.argtypes = [C.c_int, C.c_double]
        L.or_dlegendre.argtypes = [C.c_int, C.c_double]
        L.or_fr2d_len.restype = C.c_size_t
        L.or_shock_detector.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double]
        L.or_shock_detector.restype = C.c_int
        for name in ("or_L1_error", "or_L2_error", "or_Linf_error"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [_dp, _dp, C.c_size_t, C.c_double]
        L.or_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


# ----------------------------------------------------------------- operators
def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_gauss_legendre(C.c_int(n), _p(x), _p(w))
    return x, w


def dlegendre(n: int, x: float) -> float:
    return lib().or_dlegendre(n, float(x))


def dcorrection(kind: str, p: int, x):
    x = _f64(x)
    dgl = np.zeros_like(x)
    dgr = np.zeros_like(x)
    lib().or_dcorrection(C.c_int(CORR[kind]), C.c_int(p), _p(x), C.c_int(x.size), _p(dgl), _p(dgr))
    return dgl, dgr


def lagrange_point(sp, x: float):
    sp = _f64(sp)
    out = np.zeros_like(sp)
    lib().or_lagrange_point(_p(sp), C.c_int(sp.size), C.c_double(x), _p(out))
    return out


def dlagrange(sp):
    sp = _f64(sp)
    n = sp.size
    out = np.zeros(n * n)
    lib().or_dlagrange(_p(sp), C.c_int(n), _p(out))
    return out.reshape((n, n), order="F")


def jacobi_p(x: float, a: float, b: float, N: int) -> float:
    return lib().or_jacobi_p(float(x), float(a), float(b), int(N))


def vandermonde_1d(N: int, r):
    r = _f64(r)
    V = np.zeros(r.size * (N + 1))
    lib().or_vandermonde_1d(C.c_int(N), _p(r), C.c_int(r.size), _p(V))
    return V.reshape((r.size, N + 1), order="F")


def dvandermonde_1d(N: int, r):
    r = _f64(r)
    V = np.zeros(r.size * (N + 1))
    lib().or_dvandermonde_1d(C.c_int(N), _p(r), C.c_int(r.size), _p(V))
    return V.reshape((r.size, N + 1), order="F")


def vandermonde_quad(N: int, r, s):
    r = _f64(r)
    s = _f64(s)
    V = np.zeros(r.size * (N + 1) ** 2)
    lib().or_vandermonde_quad(C.c_int(N), _p(r), _p(s), C.c_int(r.size), _p(V))
    return V.reshape((r.size, (N + 1) ** 2), order="F")


def inv(A):
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    a = _f64(A)
    out = np.zeros(n * n)
    rc = lib().or_inv(C.c_int(n), _p(a), _p(out))
    assert rc == 0
    return out.reshape((n, n), order="F")


def rs_jacobi_quad(r: float, s: float, vert):
    v = _f64(np.asarray(vert, dtype=np.float64))
    J = np.zeros(4)
    lib().or_rs_jacobi_quad(C.c_double(r), C.c_double(s), _p(v), _p(J))
    return J.reshape((2, 2), order="F")


class Ops:
    """Operator set of FRPSpace1D/2D computed by the oracle (src/struct.jl:40-85,130-227)."""

    def __init__(self, deg: int, correction: str = "radau"):
        n = deg + 1
        self.deg = deg
        self.nsp = n
        self.r, self.w = gauss_legendre(n)
        self.ll = lagrange_point(self.r, -1.0)
        self.lr = lagrange_point(self.r, 1.0)
        self.dl = dlagrange(self.r)
        self.dhl, self.dhr = dcorrection(correction, deg, self.r)
        self.V1 = vandermonde_1d(deg, self.r)
        self.iV1 = inv(self.V1)
        r2 = np.repeat(self.r[:, None], n, axis=1)  # r2d[:, j] = r
        s2 = np.repeat(self.r[None, :], n, axis=0)  # s2d[i, :] = r
        self.V = vandermonde_quad(deg, r2.ravel(order="F"), s2.ravel(order="F"))
        self.iV = inv(self.V)
        self.wp = np.outer(self.w, self.w)


# ----------------------------------------------------------------- physics
def riemann2(kind: str, wL, wR, g: float):
    a = _f64(wL)
    b = _f64(wR)
    out = np.zeros(4)
    lib().or_riemann2(C.c_int(FLUX[kind]), _p(a), _p(b), C.c_double(g), _p(out))
    return out


def riemann1(kind: str, wL, wR, g: float):
    a = _f64(wL)
    b = _f64(wR)
    out = np.zeros(3)
    lib().or_riemann1(C.c_int(FLUX[kind]), _p(a), _p(b), C.c_double(g), _p(out))
    return out


def euler_flux2(w, g: float):
    a = _f64(w)
    F = np.zeros(4)
    G = np.zeros(4)
    lib().or_euler_flux2(_p(a), C.c_double(g), _p(F), _p(G))
    return F, G


def prim_conserve2(prim, g: float):
    a = _f64(prim)
    out = np.zeros(4)
    lib().or_prim_conserve2(_p(a), C.c_double(g), _p(out))
    return out


def prim_conserve1(prim, g: float):
    out = np.zeros(3)
    lib().or_prim_conserve1(_p(_f64(prim)), C.c_double(g), _p(out))
    return out


def conserve_prim2(w, g: float):
    a = _f64(w)
    out = np.zeros(4)
    lib().or_conserve_prim2(_p(a), C.c_double(g), _p(out))
    return out


def solution_field2d(u, xpg, g: float):
    """example/shock-vortex.jl:333-352: for every interior element (i, j) and solution point
    (k, l), x[idx, idy] = xpg[i, j, k, l, 1], y likewise, sol[idx, idy, :] = conserve_prim(u) with
    the last entry inverted, idx = (i-1)*nsp + k, idy = (j-1)*nsp + l (1-based in the reference).
    u and xpg carry the ghost layer (i, j = 0 .. n+1).  Loops: small inputs only."""
    nx, ny, n = u.shape[0] - 2, u.shape[1] - 2, u.shape[2]
    x = np.zeros((nx * n, ny * n))
    y = np.zeros((nx * n, ny * n))
    sol = np.zeros((nx * n, ny * n, 4))
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            idx0, idy0 = (i - 1) * n, (j - 1) * n
            for k in range(n):
                for l in range(n):
                    x[idx0 + k, idy0 + l] = xpg[i, j, k, l, 0]
                    y[idx0 + k, idy0 + l] = xpg[i, j, k, l, 1]
                    prim = conserve_prim2(u[i, j, k, l, :], g)
                    prim[3] = 1.0 / prim[3]
                    sol[idx0 + k, idy0 + l, :] = prim
    return x, y, sol


# ----------------------------------------------------------------- 2D residual
class OrFr2d(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nsp", C.c_int),
        ("gamma", C.c_double),
        ("flux", C.c_int),
        ("bc", C.c_int),
        ("iJ", C.c_double * 4),
        ("ll", _dp),
        ("lr", _dp),
        ("dl", _dp),
        ("dhl", _dp),
        ("dhr", _dp),
        ("geom", C.c_int),
        ("n1", C.c_double * 2),
        ("n2", C.c_double * 2),
        ("iJp", _dp),
    ]


class Fr2d:
    """Oracle 2D Euler residual on a uniform grid (ghost layer of one cell).

    Rectangular cells (example/shock-vortex.jl:26-118) by default.  With `iJ` (the uniform 2x2
    inverse Jacobian [[r_x, r_y], [s_x, s_y]] of parallelogram cells) and the face normals n1, n2,
    the interface terms of dev/parallelogram.jl:116-162 (fluxes solved in the face-normal frames,
    scaled by (iJ n1)_1 and (iJ n2)_2).  With `iJp`, the inverse Jacobian of every solution point,
    shape (nx+2, ny+2, nsp, nsp, 2, 2) (src/struct.jl:134-142 for any quad mesh): the residual of
    example/shock-vortex.jl:26-118 with the metric of each point (volume flux iJ[k, l] [F, G],
    face fluxes of point (k, l) scaled by its own iJ11 / iJ22); dx, dy are then unused."""

    def __init__(self, nx, ny, deg, dx, dy, gamma=1.4, flux="hll", bc="ghost", ops: Ops | None = None,
                 iJ=None, n1=None, n2=None, iJp=None):
        self.ops = ops or Ops(deg)
        self.nx, self.ny, self.nsp = nx, ny, deg + 1
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (self.ops.ll, self.ops.lr, self.ops.dl.ravel(order="F"), self.ops.dhl, self.ops.dhr)]
        d = OrFr2d()
        d.nx, d.ny, d.nsp = nx, ny, deg + 1
        d.gamma = gamma
        d.flux = FLUX[flux]
        d.bc = BC[bc]
        if iJ is None:
            d.iJ[0] = 1.0 / (dx / 2.0)
            d.iJ[1] = 0.0
            d.iJ[2] = 0.0
            d.iJ[3] = 1.0 / (dy / 2.0)
        else:
            iJ = np.asarray(iJ, dtype=np.float64)
            d.iJ[0], d.iJ[1], d.iJ[2], d.iJ[3] = iJ[0, 0], iJ[1, 0], iJ[0, 1], iJ[1, 1]
        d.geom = 0
        if iJ is not None or n1 is not None or n2 is not None:
            d.geom = 1
            n1 = (1.0, 0.0) if n1 is None else n1
            n2 = (0.0, 1.0) if n2 is None else n2
            d.n1[0], d.n1[1] = float(n1[0]), float(n1[1])
            d.n2[0], d.n2[1] = float(n2[0]), float(n2[1])
        d.ll, d.lr, d.dl, d.dhl, d.dhr = [_p(a) for a in self._keep]
        if iJp is not None:
            if d.geom != 0:
                raise ValueError("point metrics take the rectangle face terms (no iJ / n1 / n2)")
            iJp = np.asfortranarray(iJp, dtype=np.float64)
            if iJp.shape != (nx + 2, ny + 2, deg + 1, deg + 1, 2, 2):
                raise ValueError("iJp must have shape (nx+2, ny+2, nsp, nsp, 2, 2)")
            self._iJp = iJp
            d.iJp = _p(iJp)
        self.d = d
        self.shape = (nx + 2, ny + 2, deg + 1, deg + 1, 4)

    def _u(self, u):
        assert u.shape == self.shape and u.dtype == np.float64 and u.flags["F_CONTIGUOUS"]
        return u

    def rhs(self, u):
        u = self._u(u)
        du = np.zeros(self.shape, order="F")
        lib().or_fr2d_rhs(C.byref(self.d), _p(u), _p(du))
        return du

    def fill_periodic(self, u):
        lib().or_fr2d_fill_periodic(C.byref(self.d), _p(self._u(u)))

    def ssprk3(self, u, dt, nsteps):
        lib().or_fr2d_ssprk3(C.byref(self.d), _p(self._u(u)), C.c_double(dt), C.c_int(nsteps))

    def rk4(self, u, dt, nsteps):
        lib().or_fr2d_rk4(C.byref(self.d), _p(self._u(u)), C.c_double(dt), C.c_int(nsteps))

    def midpoint(self, u, dt, nsteps):
        lib().or_fr2d_midpoint(C.byref(self.d), _p(self._u(u)), C.c_double(dt), C.c_int(nsteps))

    def tsit5(self, u, dt, nsteps):
        """Fixed-step Tsit5 (the module's tsit5 loop over this residual; du = 0 on the ghost layer,
        which the steps therefore keep), in place; periodic grids refresh their ghost layer from
        the interior at the end, as or_fr2d_ssprk3 does."""
        u[...] = tsit5(self.rhs, u, dt, nsteps)
        if self.d.bc == BC["periodic"]:
            self.fill_periodic(u)

    def stage(self, u0, us, a0, a1, b):
        out = np.zeros(self.shape, order="F")
        lib().or_fr2d_stage(C.byref(self.d), _p(self._u(u0)), _p(self._u(us)), _p(out),
                            C.c_double(a0), C.c_double(a1), C.c_double(b))
        return out

    def sensor_filter(self, u, lam=5e-4, S0=None, kappa=9.0):
        deg = self.nsp - 1
        if S0 is None:
            S0 = -3.0 * np.log10(deg)
        V = np.ascontiguousarray(self.ops.V.ravel(order="F"))
        iV = np.ascontiguousarray(self.ops.iV.ravel(order="F"))
        flags = np.zeros((self.nx + 2) * (self.ny + 2), dtype=np.int32)
        n = lib().or_fr2d_sensor_filter(C.byref(self.d), _p(V), _p(iV), _p(self._u(u)), C.c_double(lam),
                                         C.c_double(S0), C.c_double(kappa),
                                         flags.ctypes.data_as(C.POINTER(C.c_int)))
        return n, flags.reshape((self.nx + 2, self.ny + 2), order="F")

    def filter_cells(self, u, lam):
        """L2 modal filter with per-cell strength lam (shape (nx+2, ny+2)); returns #filtered."""
        V = np.ascontiguousarray(self.ops.V.ravel(order="F"))
        iV = np.ascontiguousarray(self.ops.iV.ravel(order="F"))
        lam = np.asfortranarray(lam, dtype=np.float64)
        return lib().or_fr2d_filter_cells(C.byref(self.d), _p(V), _p(iV), _p(self._u(u)), _p(lam))

    def positivity(self, u):
        wp4 = np.ascontiguousarray((self.ops.wp / 4.0).ravel(order="F"))
        return lib().or_fr2d_positivity(C.byref(self.d), _p(wp4), _p(self._u(u)))


def positive_limiter2d(ue, g, weights, ll, lr):
    """One element ue(k, l, m); returns (rc, limited copy)."""
    a = np.array(ue, dtype=np.float64, order="F")
    n = a.shape[0]
    w = _f64(weights)
    rc = lib().or_positive_limiter2d(_p(a), C.c_int(n), C.c_double(g), _p(w), _p(_f64(ll)), _p(_f64(lr)))
    return rc, a


def positive_limiter1d_scalar(u, weights, ll, lr):
    """One scalar element u(p) (src/dissipation.jl:38-59, positive_limiter for a scalar field).
    weights is an array or one number; as in Julia, sum(weights) of one number is that number, so
    the element mean um = sum(u .* w) / sum(w) is then sum(u).  Returns (rc, limited copy)."""
    a = np.array(u, dtype=np.float64)
    scalar = np.ndim(weights) == 0
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, dtype=np.float64), a.shape))
    wsum = float(weights) if scalar else float(np.sum(w))
    rc = lib().or_positive_limiter1d_scalar(_p(a), C.c_int(a.size), _p(w), C.c_double(wsum), _p(_f64(ll)),
                                            _p(_f64(lr)))
    return rc, a


def positive_limiter1d_euler(ue, g, weights, ll, lr):
    """One 1D Euler element ue(p, k) (nsp x 3); weights sum to one. Returns (rc, limited copy)."""
    a = np.array(ue, dtype=np.float64, order="F")
    rc = lib().or_positive_limiter1d_euler(_p(a), C.c_int(a.shape[0]), C.c_double(g), _p(_f64(weights)),
                                           _p(_f64(ll)), _p(_f64(lr)))
    return rc, a


def fr1d_positivity(u, g, weights, ll, lr):
    """The 1D limiter over every cell of u(i, p, k) in place; returns the failed-cell count."""
    assert u.flags.f_contiguous and u.dtype == np.float64
    return lib().or_fr1d_positivity(C.c_int(u.shape[0]), C.c_int(u.shape[1]), C.c_double(g), _p(_f64(weights)),
                                    _p(_f64(ll)), _p(_f64(lr)), _p(u))


def fr1d_sensor_filter(u, ops: "Ops", dt):
    """example/euler_highlevel.jl:37-52 on u(i, p, k) in place; returns the flagged-cell count."""
    assert u.flags.f_contiguous and u.dtype == np.float64
    return lib().or_fr1d_sensor_filter(C.c_int(u.shape[0]), C.c_int(u.shape[1]), _p(_f64(ops.V1)), _p(_f64(ops.iV1)),
                                       C.c_double(dt), _p(u))


def shock_detector(Se, deg, S0=None, kappa=4.0) -> bool:
    if S0 is None:
        S0 = -3.0 * np.log10(deg)
    return bool(lib().or_shock_detector(float(Se), int(deg), float(S0), float(kappa)))


def filter_exp1d(N, s, Nc=0):
    d = np.zeros(N + 1)
    lib().or_filter_exp1d(C.c_int(N), C.c_int(s), C.c_int(Nc), _p(d))
    return d


def filter_exp2d(N, s, Nc=0):
    d = np.zeros((N + 1) * (N + 2) // 2)
    lib().or_filter_exp2d(C.c_int(N), C.c_int(s), C.c_int(Nc), _p(d))
    return d


def filter_exp(N, s, V, Nc=0, invV=None):
    """src/Polynomial/poly_filter.jl:11-24; None where the reference leaves the filter undefined."""
    V = np.asfortranarray(V, dtype=np.float64)
    nv = V.shape[0]
    iv = np.asfortranarray(inv(V) if invV is None else invV, dtype=np.float64)
    F = np.zeros((nv, nv), order="F")
    rc = lib().or_filter_exp(C.c_int(N), C.c_int(s), C.c_int(Nc), _p(V), C.c_int(nv), _p(iv), _p(F))
    return None if rc else F


def modal_filter_l2(u, lam):
    a = _f64(u).copy()
    lib().or_modal_filter_l2(_p(a), C.c_int(a.size), C.c_double(lam))
    return a


# ----------------------------------------------------------------- 1D residuals
def fr1d_euler_rhs(u, J, ops: Ops, g, flux="hll", bc="dirichlet"):
    u = np.asfortranarray(u, dtype=np.float64)
    ncell, nsp, _ = u.shape
    du = np.zeros_like(u, order="F")
    Jv = _f64(J)
    dl = _f64(ops.dl)
    lib().or_fr1d_euler_rhs(C.c_int(ncell), C.c_int(nsp), _p(Jv), _p(_f64(ops.ll)), _p(_f64(ops.lr)), _p(dl),
                            _p(_f64(ops.dhl)), _p(_f64(ops.dhr)), C.c_double(g), C.c_int(FLUX[flux]),
                            C.c_int(0 if bc == "dirichlet" else 1), _p(u), _p(du))
    return du


def fr1d_advection_rhs(u, J, ops: Ops, a=1.0, bc="period"):
    u = np.asfortranarray(u, dtype=np.float64)
    ncell, nsp = u.shape
    du = np.zeros_like(u, order="F")
    lib().or_fr1d_advection_rhs(C.c_int(ncell), C.c_int(nsp), _p(_f64(J)), _p(_f64(ops.ll)), _p(_f64(ops.lr)),
                                _p(_f64(ops.dl)), _p(_f64(ops.dhl)), _p(_f64(ops.dhr)), C.c_double(a),
                                C.c_int(0 if bc == "dirichlet" else 1), _p(u), _p(du))
    return du


# ----------------------------------------------------------------- Tsit5 (fixed step)
# Tsitouras' 5(4) pair (Ch. Tsitouras, "Runge-Kutta pairs of order 5(4) satisfying only the first
# column simplifying assumption", Comput. Math. Appl. 62 (2011) 770-775, Table 1, to double
# precision): OrdinaryDiffEq's Tsit5, which example/advection_highlevel.jl:26 and
# example/euler1d_convergence.jl:133 step with adaptive = false.  TSIT5_C: nodes c_1..c_7;
# TSIT5_A[i]: a_{i+2, 1..i+1}; TSIT5_B: b_1..b_6 (b_7 = 0, the seventh stage is the next step's
# first).  Pinned by the order conditions through order 5 (tests/test_oracle_tsit5.py).
TSIT5_C = (0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0)
TSIT5_A = (
    (0.161,),
    (-0.008480655492356989, 0.335480655492357),
    (2.897153057105493, -6.359448489975075, 4.3622954328695815),
    (5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525),
    (5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383),
)
TSIT5_B = (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
           2.324710524099774)


def tsit5(L, u, dt, nsteps):
    """nsteps fixed steps of Tsit5 on du/dt = L(u): k_1 = L(u), k_s = L(u + dt sum_j a_sj k_j),
    u <- u + dt sum_j b_j k_j (the stage sums in the order j = 1, 2, ..)."""
    u = np.array(u, dtype=np.float64, order="F")
    for _ in range(nsteps):
        k = [L(u)]
        for row in TSIT5_A:
            s = row[0] * k[0]
            for a, kj in zip(row[1:], k[1:]):
                s = s + a * kj
            k.append(L(u + dt * s))
        s = TSIT5_B[0] * k[0]
        for b, kj in zip(TSIT5_B[1:], k[1:]):
            s = s + b * kj
        u = u + dt * s
    return u


def L1_error(u, ue, dx):
    a = _f64(u)
    b = _f64(ue)
    return lib().or_L1_error(_p(a), _p(b), a.size, dx)


def L2_error(u, ue, dx):
    a = _f64(u)
    b = _f64(ue)
    return lib().or_L2_error(_p(a), _p(b), a.size, dx)


def Linf_error(u, ue, dx):
    a = _f64(u)
    b = _f64(ue)
    return lib().or_Linf_error(_p(a), _p(b), a.size, dx)


# ----------------------------------------------------------------- 2D BGK
class OrBgk2d(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nsp", C.c_int), ("nu", C.c_int), ("nv", C.c_int),
                ("iJx", C.c_double), ("iJy", C.c_double), ("tau", C.c_double), ("gamma", C.c_double),
                ("uv", _dp), ("vv", _dp), ("wu", _dp), ("wv", _dp),
                ("ll", _dp), ("lr", _dp), ("dl", _dp), ("dhl", _dp), ("dhr", _dp)]


def bgk2d_rhs(f, nx, ny, deg, u, wu, v, wv, tau, Lx=1.0, Ly=1.0, gamma=2.0):
    ops = Ops(deg)
    keep = [_f64(a) for a in (u, v, wu, wv, ops.ll, ops.lr, ops.dl, ops.dhl, ops.dhr)]
    d = OrBgk2d()
    d.nx, d.ny, d.nsp, d.nu, d.nv = nx, ny, deg + 1, len(u), len(v)
    d.iJx, d.iJy, d.tau, d.gamma = 2.0 / (Lx / nx), 2.0 / (Ly / ny), tau, gamma
    d.uv, d.vv, d.wu, d.wv, d.ll, d.lr, d.dl, d.dhl, d.dhr = [_p(a) for a in keep]
    f = _f64(f)
    df = np.zeros_like(f)
    lib().or_bgk2d_rhs(C.byref(d), _p(f), _p(df))
    return df


# ----------------------------------------------------------------- 1D BGK (bgk_wave.jl as written)
class OrBgk1d(C.Structure):
    _fields_ = [("nx", C.c_int), ("nu", C.c_int), ("nsp", C.c_int),
                ("dx", C.c_double), ("tau", C.c_double), ("gamma", C.c_double),
                ("uv", _dp), ("wv", _dp), ("ll", _dp), ("lr", _dp), ("dl", _dp), ("dhl", _dp), ("dhr", _dp)]


class Bgk1d:
    """Oracle of example/bgk_wave.jl:69-129 (mol!): state u(i, j, k) of nx periodic cells, the
    velocities u / weights w of the velocity space and nsp = deg + 1 points, column-major; tau the
    relaxation time (1e-2 in the script, inf for free transport), gamma = 3 as the script passes."""

    def __init__(self, nx, deg, uv, wv, dx, tau=1e-2, gamma=3.0):
        self.ops = Ops(deg)
        self.nx, self.nu, self.nsp = nx, len(uv), deg + 1
        self._keep = [_f64(a) for a in (uv, wv, self.ops.ll, self.ops.lr, self.ops.dl, self.ops.dhl, self.ops.dhr)]
        d = OrBgk1d()
        d.nx, d.nu, d.nsp = nx, len(uv), deg + 1
        d.dx, d.tau, d.gamma = dx, tau, gamma
        d.uv, d.wv, d.ll, d.lr, d.dl, d.dhl, d.dhr = [_p(a) for a in self._keep]
        self.d = d
        self.shape = (nx, len(uv), deg + 1)

    def rhs(self, f):
        f = np.asfortranarray(f, dtype=np.float64).reshape(self.shape, order="F")
        df = np.zeros(self.shape, order="F")
        lib().or_bgk1d_rhs(C.byref(self.d), _p(f), _p(df))
        return df

    def moments(self, f):
        f = np.asfortranarray(f, dtype=np.float64).reshape(self.shape, order="F")
        w = np.zeros((self.nx, self.nsp, 3), order="F")
        lib().or_bgk1d_moments(C.byref(self.d), _p(f), _p(w))
        return w

    def stage(self, u0, us, a0, a1, b):
        """out = a0 u0 + a1 us + b L(us), the stage convention of the device steppers."""
        return a0 * u0 + a1 * us + b * self.rhs(us)

    def step(self, f, dt, nsteps, scheme="midpoint"):
        u = np.asfortranarray(f, dtype=np.float64).reshape(self.shape, order="F").copy(order="F")
        for _ in range(nsteps):
            if scheme == "midpoint":  # OrdinaryDiffEq Midpoint (bgk_wave.jl:136)
                u1 = self.stage(u, u, 0.0, 1.0, 0.5 * dt)
                u = self.stage(u, u1, 1.0, 0.0, dt)
            elif scheme == "ssprk3":
                u1 = self.stage(u, u, 0.0, 1.0, dt)
                u2 = self.stage(u, u1, 0.75, 0.25, 0.25 * dt)
                u = self.stage(u, u2, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt)
            elif scheme == "euler":
                u = self.stage(u, u, 0.0, 1.0, dt)
            else:
                raise ValueError(scheme)
        return u
