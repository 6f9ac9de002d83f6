"""2D discrete-velocity BGK on periodic quads (example/bgk_wave.jl:69-129 tensorised) and its
batched ensembles, on libfrhip's gfx950 kernels (include/fr.h, "2D discrete-velocity BGK")."""
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


def _bind():
    L = _lib.lib()
    for name in ("fr_bgk_create", "fr_bgk_destroy", "fr_bgk_state_len", "fr_bgk_rhs", "fr_bgk_set_state",
                 "fr_bgk_set_state_dev",
                 "fr_bgk_get_state", "fr_bgk_step", "fr_bgk_moments", "fr_gauss_hermite"):
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
