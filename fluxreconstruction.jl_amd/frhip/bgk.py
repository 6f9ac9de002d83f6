"""Discrete-velocity BGK on libfrhip's gfx950 kernels: the 1D model of example/bgk_wave.jl:69-129
as the script writes it (include/fr.h, "1D discrete-velocity BGK"), its tensorisation to periodic
quads (2D), and batched ensembles of both that differ in the relaxation time."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


class FrBgkDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("deg", C.c_int), ("nu", C.c_int), ("nv", C.c_int),
                ("nmem", C.c_int), ("dx", C.c_double), ("dy", C.c_double), ("gamma", C.c_double),
                ("u", _dp), ("v", _dp), ("wu", _dp), ("wv", _dp), ("tau", _dp)]


_lib.SIGNATURES.update({
    "fr_bgk_create": (C.c_int, [C.POINTER(_vp), C.POINTER(FrBgkDesc)]),
    "fr_bgk_destroy": (C.c_int, [_vp]),
    "fr_bgk_state_len": (C.c_size_t, [_vp]),
    "fr_bgk_rhs": (C.c_int, [_vp, _dp, _dp]),
    "fr_bgk_set_state": (C.c_int, [_vp, _dp]),
    "fr_bgk_set_state_dev": (C.c_int, [_vp, _vp]),
    "fr_bgk_get_state": (C.c_int, [_vp, _dp]),
    "fr_bgk_step": (C.c_int, [_vp, C.c_double, C.c_int]),
    "fr_bgk_moments": (C.c_int, [_vp, _dp]),
    "fr_gauss_hermite": (C.c_int, [C.c_int, _dp, _dp]),
})


class FrBgk1dDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("deg", C.c_int), ("nu", C.c_int), ("nmem", C.c_int), ("dx", C.c_double),
                ("gamma", C.c_double), ("u", _dp), ("w", _dp), ("tau", _dp)]


_lib.SIGNATURES.update({
    "fr_bgk1d_create": (C.c_int, [C.POINTER(_vp), C.POINTER(FrBgk1dDesc)]),
    "fr_bgk1d_destroy": (C.c_int, [_vp]),
    "fr_bgk1d_state_len": (C.c_size_t, [_vp]),
    "fr_bgk1d_rhs": (C.c_int, [_vp, _dp, _dp]),
    "fr_bgk1d_rhs_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "fr_bgk1d_set_state": (C.c_int, [_vp, _dp]),
    "fr_bgk1d_get_state": (C.c_int, [_vp, _dp]),
    "fr_bgk1d_step": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int]),
    "fr_bgk1d_step_graph": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int]),
    "fr_bgk1d_set_path": (C.c_int, [_vp, C.c_int]),
    "fr_bgk1d_get_path": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "fr_bgk1d_moments": (C.c_int, [_vp, _dp]),
})


def _bind():
    L = _lib.lib()
    for name in ("fr_bgk_create", "fr_bgk_destroy", "fr_bgk_state_len", "fr_bgk_rhs", "fr_bgk_set_state",
                 "fr_bgk_set_state_dev",
                 "fr_bgk_get_state", "fr_bgk_step", "fr_bgk_moments", "fr_gauss_hermite",
                 "fr_bgk1d_create", "fr_bgk1d_destroy", "fr_bgk1d_state_len", "fr_bgk1d_rhs", "fr_bgk1d_rhs_dev",
                 "fr_bgk1d_set_state", "fr_bgk1d_get_state", "fr_bgk1d_step", "fr_bgk1d_step_graph", "fr_bgk1d_set_path",
                 "fr_bgk1d_get_path", "fr_bgk1d_moments"):
        res, args = _lib.SIGNATURES[name]
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def gauss_hermite(n: int):
    """Nodes and weights of the n-point Gauss-Hermite rule for exp(-x^2) (ascending)."""
    x = np.zeros(n)
    w = np.zeros(n)
    check(_bind().fr_gauss_hermite(n, dptr(x), dptr(w)))
    return x, w


def velocity_grid(n: int):
    """Discrete velocities for integrals over the real line: Gauss-Hermite nodes and weights
    w e^{x^2}, exact for Maxwellians with lambda = 1 (KitBase VSpace of Gauss-Hermite type)."""
    x, w = gauss_hermite(n)
    return x, w * np.exp(x * x)


class BGK2D:
    """Periodic nx x ny quads of degree deg on [0, Lx] x [0, Ly], velocity grid (u, wu) x (v, wv),
    one relaxation time per ensemble member."""

    def __init__(self, nx, ny, deg, u, wu, v, wv, tau, Lx=1.0, Ly=1.0, gamma=2.0):
        L = _bind()
        self.tau = np.atleast_1d(np.asarray(tau, dtype=np.float64))
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, wu, wv, self.tau)]
        d = FrBgkDesc()
        d.nx, d.ny, d.deg, d.nu, d.nv, d.nmem = nx, ny, deg, len(u), len(v), self.tau.size
        d.dx, d.dy, d.gamma = Lx / nx, Ly / ny, gamma
        d.u, d.v, d.wu, d.wv, d.tau = [dptr(a) for a in self._keep]
        h = C.c_void_p()
        check(L.fr_bgk_create(C.byref(h), C.byref(d)))
        self.h = h
        self.nx, self.ny, self.nsp, self.nu, self.nv, self.nmem = nx, ny, deg + 1, len(u), len(v), self.tau.size
        self.Lx, self.Ly = Lx, Ly
        self.shape = (nx, ny, deg + 1, deg + 1, len(u), len(v))
        self.len = L.fr_bgk_state_len(h)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr_bgk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, f):
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel(order="F"))
        df = np.zeros_like(f)
        check(_lib.lib().fr_bgk_rhs(self.h, dptr(f), dptr(df)))
        return df

    def set_state(self, f):
        """Host array (any shape, column-major order) or a flat float64 torch GPU tensor."""
        if type(f).__module__.startswith("torch"):
            assert f.is_cuda and str(f.dtype) == "torch.float64" and f.is_contiguous() and f.numel() == self.len
            check(_lib.lib().fr_bgk_set_state_dev(self.h, C.c_void_p(f.data_ptr())))
            return
        flat = np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel(order="F"))
        check(_lib.lib().fr_bgk_set_state(self.h, dptr(flat)))

    def get_state(self):
        f = np.zeros(self.len)
        check(_lib.lib().fr_bgk_get_state(self.h, dptr(f)))
        return f

    def step(self, dt, nsteps=1):
        check(_lib.lib().fr_bgk_step(self.h, float(dt), int(nsteps)))

    def moments(self):
        w = np.zeros(self.nmem * 4 * self.nx * self.ny * self.nsp ** 2)
        check(_lib.lib().fr_bgk_moments(self.h, dptr(w)))
        return w.reshape((self.nmem, 4, -1))


def maxwellian_state(nx, ny, deg, u, v, Lx=1.0, Ly=1.0, amp=0.1, lam=1.0, nmem=1):
    """Initial Maxwellian with density 1 + amp sin(2 pi x) sin(2 pi y), zero mean velocity and
    lambda = rho / (2 p) = lam, at the Gauss solution points; flat, members concatenated."""
    r, _ = np.polynomial.legendre.leggauss(deg + 1)
    xc = (np.arange(nx) + 0.5) * Lx / nx
    yc = (np.arange(ny) + 0.5) * Ly / ny
    x = xc[:, None, None, None] + 0.5 * Lx / nx * r[None, None, :, None]
    y = yc[None, :, None, None] + 0.5 * Ly / ny * r[None, None, None, :]
    rho = 1.0 + amp * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y)
    U, V = np.meshgrid(u, v, indexing="ij")
    f = rho[..., None, None] * (lam / np.pi) * np.exp(-lam * (U ** 2 + V ** 2))[None, None, None, None]
    one = np.asfortranarray(f).ravel(order="F")
    return np.tile(one, nmem)


# ----------------------------------------------------------------- 1D BGK (bgk_wave.jl as written)
class VSpace1D:
    """Uniform velocity space of nu midpoints on [u0, u1] with weights du (KitBase VSpace1D as
    bgk_wave.jl:25 builds it: fields u, du, weights).  KitBase is not part of the reference tree,
    so the midpoint placement follows its documented rectangle rule, not pinned against KitBase."""

    def __init__(self, u0, u1, nu):
        self.u0, self.u1, self.nu = float(u0), float(u1), int(nu)
        du = (self.u1 - self.u0) / nu
        self.u = self.u0 + (np.arange(nu) + 0.5) * du
        self.du = np.full(nu, du)
        self.weights = np.full(nu, du)


def maxwellian1d(u, prim):
    """KitBase maxwellian(u, prim) in 1D: rho sqrt(lambda / pi) exp(-lambda (u - U)^2)."""
    rho, U, lam = prim
    return rho * np.sqrt(lam / np.pi) * np.exp(-lam * (np.asarray(u) - U) ** 2)


def bgk_wave_initial(ps, vs):
    """The initial distribution of example/bgk_wave.jl:34-40: at every solution point the
    Maxwellian of rho = 1 + 0.1 sin(2 pi x), U = 1, lambda = 1 / T with T = 2 * 0.5 / rho;
    shape (nx, nu, nsp), column-major."""
    nx, nsp = ps.xpg.shape
    f = np.zeros((nx, vs.nu, nsp), order="F")
    for i in range(nx):
        for k in range(nsp):
            rho = 1.0 + 0.1 * np.sin(2.0 * np.pi * ps.xpg[i, k])
            T = 2 * 0.5 / rho
            f[i, :, k] = maxwellian1d(vs.u, (rho, 1.0, 1.0 / T))
    return f


class BGK1D:
    """mol! of example/bgk_wave.jl:69-129 on the device: the FRPSpace1D ps (periodic cells, Radau
    correction), the velocity space vs (u, weights), the relaxation time tau (the script: 1e-2; a
    sequence gives an ensemble of members that differ in tau) and gamma of conserve_prim (3)."""

    def __init__(self, ps, vs, tau=1e-2, gamma=3.0):
        L = _bind()
        self.tau = np.atleast_1d(np.asarray(tau, dtype=np.float64))
        u = np.asarray(vs.u, dtype=np.float64)
        w = np.asarray(vs.weights, dtype=np.float64)
        self._keep = [np.ascontiguousarray(a) for a in (u, w, self.tau)]
        dx = np.asarray(ps.base.dx if hasattr(ps, "base") else ps.dx, dtype=np.float64)
        if not np.allclose(dx, dx.flat[0], rtol=1e-12, atol=0.0):
            raise ValueError("BGK1D: cells of one size")
        d = FrBgk1dDesc()
        d.nx, d.deg, d.nu, d.nmem = ps.nx, ps.deg, u.size, self.tau.size
        d.dx, d.gamma = float(dx.flat[0]), float(gamma)
        d.u, d.w, d.tau = [dptr(a) for a in self._keep]
        h = C.c_void_p()
        check(L.fr_bgk1d_create(C.byref(h), C.byref(d)))
        self.h = h
        self.nx, self.nu, self.nsp, self.nmem = ps.nx, u.size, ps.deg + 1, self.tau.size
        self.shape = (ps.nx, u.size, ps.deg + 1)
        self.len = L.fr_bgk1d_state_len(h)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr_bgk1d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _flat(self, f):
        flat = np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel(order="F"))
        if flat.size != self.len:
            raise ValueError(f"state must hold {self.len} values")
        return flat

    def rhs(self, f):
        """du of every member for a host state (nx, nu, nsp) per member, column-major."""
        flat = self._flat(f)
        df = np.zeros_like(flat)
        check(_lib.lib().fr_bgk1d_rhs(self.h, dptr(flat), dptr(df)))
        return df.reshape(self.shape + ((self.nmem,) if self.nmem > 1 else ()), order="F")

    def mol(self, du, u, p=None, t=0.0):
        """mol!(du, u, p, t): host arrays, or flat float64 torch GPU tensors on torch's current
        stream."""
        if type(u).__module__.startswith("torch"):
            import torch

            for a in (u, du):
                if not (a.is_cuda and str(a.dtype) == "torch.float64" and a.is_contiguous() and a.numel() == self.len):
                    raise ValueError("device arrays must be contiguous float64 GPU tensors of the state's size")
            st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
            check(_lib.lib().fr_bgk1d_rhs_dev(self.h, C.c_void_p(u.data_ptr()), C.c_void_p(du.data_ptr()), st))
            return du
        du[...] = self.rhs(u).reshape(du.shape, order="F")
        return du

    def set_state(self, f):
        check(_lib.lib().fr_bgk1d_set_state(self.h, dptr(self._flat(f))))

    def get_state(self):
        f = np.zeros(self.len)
        check(_lib.lib().fr_bgk1d_get_state(self.h, dptr(f)))
        return f.reshape(self.shape + ((self.nmem,) if self.nmem > 1 else ()), order="F")

    def step(self, dt, nsteps=1, scheme="midpoint"):
        check(_lib.lib().fr_bgk1d_step(self.h, _lib.RK[scheme], float(dt), int(nsteps)))

    PATHS = {"auto": 0, "stream": 1, "resident": 2}

    def set_path(self, path):
        """'stream' (two launches per stage over HBM), 'resident' (a workgroup per member holding
        it in LDS for up to 256 steps per launch; FRError at step time if a member does not fit)
        or 'auto' (resident whenever it fits).  Every path gives the same bits."""
        check(_lib.lib().fr_bgk1d_set_path(self.h, self.PATHS[path]))

    def path(self, scheme="midpoint"):
        """The path step() takes for `scheme` under the current setting."""
        v = C.c_int()
        check(_lib.lib().fr_bgk1d_get_path(self.h, _lib.RK[scheme], C.byref(v)))
        return {1: "stream", 2: "resident"}[v.value]

    def step_graph(self, dt, nsteps, scheme="midpoint", chunk=100):
        """nsteps steps as replays of a captured graph of `chunk` steps (kept by the context while
        scheme, dt and chunk are unchanged) and the remainder as eager steps: the bits of step()."""
        chunk = max(1, min(int(chunk), int(nsteps))) if nsteps > 0 else 1
        nrep, rest = divmod(int(nsteps), chunk)
        if nrep:
            check(_lib.lib().fr_bgk1d_step_graph(self.h, _lib.RK[scheme], float(dt), chunk, nrep))
        if rest:
            self.step(dt, rest, scheme)

    def moments(self):
        """(rho, rho U, E) of the device state at every point: shape (nmem, 3, nx, nsp)."""
        w = np.zeros(self.nmem * 3 * self.nx * self.nsp)
        check(_lib.lib().fr_bgk1d_moments(self.h, dptr(w)))
        return w.reshape((self.nmem, 3, self.nsp, self.nx)).transpose(0, 1, 3, 2)
