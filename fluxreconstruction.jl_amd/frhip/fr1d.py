"""1D FR problems (FREulerProblem / FRAdvectionProblem of src/Equation/eq_euler.jl:1-27 and
src/Equation/eq_advection.jl:1-53) and batched ensembles of independent 1D Euler problems,
on libfrhip's gfx950 kernels (include/fr.h, "1D" and "batched 1D Euler ensembles")."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr
from .space import FRPSpace1D

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

_lib.SIGNATURES.update({
    "fr1d_create": (C.c_int, [C.POINTER(_vp), C.POINTER(_lib.Fr1dDesc)]),
    "fr1d_destroy": (C.c_int, [_vp]),
    "fr1d_state_len": (C.c_size_t, [_vp]),
    "fr1d_rhs": (C.c_int, [_vp, _dp, _dp, C.c_double]),
    "fr1d_rhs_dev": (C.c_int, [_vp, _vp, _vp, C.c_double]),
    "fr1d_set_stream": (C.c_int, [_vp, _vp]),
    "fr1d_set_state": (C.c_int, [_vp, _dp]),
    "fr1d_get_state": (C.c_int, [_vp, _dp]),
    "fr1d_step": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int]),
    "fr1d_set_path": (C.c_int, [_vp, C.c_int]),
    "fr1d_get_path": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
    "fr_ens_set_path": (C.c_int, [_vp, C.c_int]),
    "fr_ens_create": (C.c_int, [C.POINTER(_vp), C.c_int, _ip, _ip, _dp, C.c_double, C.c_int, C.c_int]),
    "fr_ens_destroy": (C.c_int, [_vp]),
    "fr_ens_state_len": (C.c_size_t, [_vp]),
    "fr_ens_offset": (C.c_size_t, [_vp, C.c_int]),
    "fr_ens_set_state": (C.c_int, [_vp, _dp]),
    "fr_ens_get_state": (C.c_int, [_vp, _dp]),
    "fr_ens_step": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int]),
    "fr_ens_step_graph": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "fr_ens_autotune": (C.c_int, [_vp, C.c_double, _ip]),
    "fr1d_sensor_filter": (C.c_int, [_vp, C.c_double, _ip]),
    "fr1d_positivity": (C.c_int, [_vp, _ip]),
    "fr_ens_sensor_filter": (C.c_int, [_vp, C.c_double, _ip]),
    "fr_ens_positivity": (C.c_int, [_vp, _ip]),
    "fr_ens_run": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _ip]),
    "fr_ens_run_graph": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _ip]),
})

ENS_FILTER, ENS_LIMIT, ENS_MIDPOINT = 1, 2, 4


def _opts(filter, limit, scheme):
    if scheme not in ("ssprk3", "midpoint"):
        raise ValueError("ensemble loops step with ssprk3 or midpoint")
    return (ENS_FILTER if filter else 0) | (ENS_LIMIT if limit else 0) | (ENS_MIDPOINT if scheme == "midpoint" else 0)


def _bind():
    L = _lib.lib()
    for name in ("fr1d_create", "fr1d_destroy", "fr1d_state_len", "fr1d_rhs", "fr1d_rhs_dev", "fr1d_set_stream",
                 "fr1d_set_state",
                 "fr1d_get_state", "fr1d_step", "fr_ens_create", "fr_ens_destroy", "fr_ens_state_len", "fr_ens_offset",
                 "fr_ens_set_state", "fr_ens_get_state", "fr_ens_step", "fr_ens_step_graph", "fr_ens_autotune",
                 "fr1d_sensor_filter", "fr1d_positivity", "fr_ens_sensor_filter", "fr_ens_positivity", "fr_ens_run",
                 "fr_ens_run_graph", "fr1d_set_path", "fr1d_get_path", "fr_ens_set_path"):
        res, args = _lib.SIGNATURES[name]
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


class FR1D:
    """One 1D FR problem.  equation "euler" (u(i, p, k), 3 variables, frode_euler!) or
    "advection" (u(i, p), frode_advection!); bc "dirichlet" or "period"."""

    def __init__(self, ps: FRPSpace1D, equation: str = "euler", gamma: float = 1.4, a: float = 1.0,
                 flux: str = "hll", bc: str = "dirichlet"):
        L = _bind()
        self.ps = ps
        self.equation = equation
        n = ps.deg + 1
        self.shape = (ps.nx, n, 3) if equation == "euler" else (ps.nx, n)
        d = _lib.Fr1dDesc()
        d.equation = _lib.EQ[equation]
        d.ncell = ps.nx
        d.deg = ps.deg
        d.dx = float(ps.base.dx[0])
        d.gamma = float(gamma)
        d.a = float(a)
        d.flux = _lib.FLUX[flux]
        d.bc = _lib.BC1[bc]
        d.correction = _lib.CORR[getattr(ps, "correction", "radau")]
        h = C.c_void_p()
        check(L.fr1d_create(C.byref(h), C.byref(d)))
        self.h = h
        self.len = L.fr1d_state_len(h)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr1d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dudt(self, du, u, p=None, t: float = 0.0):
        """In-place du = L(u): numpy (host) arrays in the reference layout, or contiguous float64
        torch GPU tensors of fr1d_state_len values, launched on torch's current stream."""
        if type(u).__module__.startswith("torch"):
            import torch

            for a in (u, du):
                if str(a.dtype) != "torch.float64" or not a.is_cuda or not a.is_contiguous() or a.numel() != self.len:
                    raise ValueError("device arrays must be contiguous float64 GPU tensors of the state size")
            check(_lib.lib().fr1d_set_stream(self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
            check(_lib.lib().fr1d_rhs_dev(self.h, C.c_void_p(u.data_ptr()), C.c_void_p(du.data_ptr()), float(t)))
            return None
        check(_lib.lib().fr1d_rhs(self.h, dptr(u), dptr(du), float(t)))

    def rhs(self, u):
        u = np.asfortranarray(u, dtype=np.float64)
        du = np.zeros(self.shape, order="F")
        self.dudt(du, u)
        return du

    def set_state(self, u):
        check(_lib.lib().fr1d_set_state(self.h, dptr(np.asfortranarray(u, dtype=np.float64))))

    def get_state(self):
        out = np.zeros(self.shape, order="F")
        check(_lib.lib().fr1d_get_state(self.h, dptr(out)))
        return out

    def step(self, dt: float, nsteps: int = 1, scheme: str = "ssprk3"):
        check(_lib.lib().fr1d_step(self.h, _lib.RK[scheme], float(dt), int(nsteps)))

    PATHS = {"auto": 0, "batched": 1, "member": 2}

    def set_path(self, path: str):
        """'batched' (a launch per stage), 'member' (one workgroup runs the whole call's stages in
        a few launches) or 'auto' (member when the problem has at most two work items, i.e. up to
        508 cells at 256 threads).  Every path gives the same bits."""
        check(_lib.lib().fr1d_set_path(self.h, self.PATHS[path]))

    def path(self, scheme: str = "ssprk3") -> str:
        """The path step() takes for `scheme` under the current setting."""
        v = C.c_int()
        check(_lib.lib().fr1d_get_path(self.h, _lib.RK[scheme], C.byref(v)))
        return {1: "batched", 2: "member"}[v.value]

    def sensor_filter(self, dt: float) -> int:
        """example/euler_highlevel.jl:37-52 on the device state; returns the flagged-cell count."""
        n = C.c_int(0)
        check(_lib.lib().fr1d_sensor_filter(self.h, float(dt), C.byref(n)))
        return n.value

    def positivity(self) -> int:
        """src/dissipation.jl:61-123 on every cell of the device state; returns the cells changed."""
        n = C.c_int(0)
        check(_lib.lib().fr1d_positivity(self.h, C.byref(n)))
        return n.value


def _problem(args, kwargs, eq, param, default_bc):
    """(ps, param[, bc]) or the reference's (u, tspan, ps, param, bc): the latter sets u as the
    device state and keeps tspan (the reference returns ODEProblem(frode!, u, tspan, p))."""
    u = tspan = None
    if args and not isinstance(args[0], FRPSpace1D):
        u, tspan, *args = args
    ps, value, *rest = args
    bc = str(rest[0] if rest else kwargs.pop("bc", default_bc)).lstrip(":")
    opts = {param: value, "bc": bc}
    if eq == "euler":
        opts["flux"] = rest[1] if len(rest) > 1 else kwargs.pop("flux", "hll")
    if kwargs:
        raise TypeError(f"unexpected arguments {sorted(kwargs)}")
    pr = FR1D(ps, eq, **opts)
    if u is not None:
        pr.set_state(u)
        pr.tspan = tuple(tspan)
    return pr


def FREulerProblem(*args, **kwargs) -> FR1D:
    """src/Equation/eq_euler.jl:1-27: FREulerProblem(ps, gamma[, bc[, flux]]) or the reference's
    FREulerProblem(u, tspan, ps, gamma, bc) (bc "dirichlet" / "period", a leading ":" allowed),
    which also makes u the device state."""
    return _problem(args, kwargs, "euler", "gamma", "dirichlet")


def FRAdvectionProblem(*args, **kwargs) -> FR1D:
    """src/Equation/eq_advection.jl:1-26: FRAdvectionProblem(ps, a[, bc]) or the reference's
    FRAdvectionProblem(u, tspan, ps, a, bc), which also makes u the device state."""
    return _problem(args, kwargs, "advection", "a", "period")


class Ensemble1D:
    """Independent 1D Euler problems advanced together (one launch per SSP-RK3 stage).
    Member b: ncell[b] cells of degree deg[b] on a uniform grid of spacing dx[b]."""

    def __init__(self, ncell, deg, dx, gamma: float = 1.4, flux: str = "rusanov", bc: str = "dirichlet"):
        L = _bind()
        self.ncell = np.ascontiguousarray(ncell, dtype=np.int32)
        self.deg = np.ascontiguousarray(deg, dtype=np.int32)
        self.dx = np.ascontiguousarray(dx, dtype=np.float64)
        h = C.c_void_p()
        check(L.fr_ens_create(C.byref(h), int(self.ncell.size), self.ncell.ctypes.data_as(_ip),
                              self.deg.ctypes.data_as(_ip), dptr(self.dx), float(gamma), _lib.FLUX[flux],
                              _lib.BC1[bc]))
        self.h = h
        self.len = L.fr_ens_state_len(h)
        self.offsets = np.array([L.fr_ens_offset(h, b) for b in range(self.ncell.size)], dtype=np.int64)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def member(self, flat, b):
        n, p = int(self.ncell[b]), int(self.deg[b]) + 1
        o = int(self.offsets[b])
        return flat[o:o + n * p * 3].reshape((n, p, 3), order="F")

    def pack(self, members):
        flat = np.zeros(self.len)
        for b, u in enumerate(members):
            o = int(self.offsets[b])
            flat[o:o + u.size] = np.asarray(u, dtype=np.float64).ravel(order="F")
        return flat

    def set_state(self, flat):
        check(_lib.lib().fr_ens_set_state(self.h, dptr(np.ascontiguousarray(flat, dtype=np.float64))))

    def get_state(self):
        out = np.zeros(self.len)
        check(_lib.lib().fr_ens_get_state(self.h, dptr(out)))
        return out

    def step(self, dt: float, nsteps: int = 1, looped: bool = False):
        check(_lib.lib().fr_ens_step(self.h, 1 if looped else 0, float(dt), int(nsteps)))

    def set_path(self, path: str):
        """Execution path of step() and run(): 'batched', 'member' (a workgroup per member) or
        'auto' (member when no member has more than two work items); same bits on every path."""
        check(_lib.lib().fr_ens_set_path(self.h, FR1D.PATHS[path]))

    def step_graph(self, dt: float, nsteps: int, nrep: int = 1):
        check(_lib.lib().fr_ens_step_graph(self.h, float(dt), int(nsteps), int(nrep)))

    def run(self, dt: float, nsteps: int, filter: bool = True, limit: bool = False, looped: bool = False,
            scheme: str = "ssprk3"):
        """The reference's per-step 1D loop (example/euler_highlevel.jl:37-54) over every member:
        sensor + filter before every step (filter), positivity limiter on the input of every stage
        (limit); steps of SSP-RK3 or of the script's own Midpoint (scheme "midpoint").  Returns
        (filtered cells, limited cells) summed over the loop."""
        nf, nl = C.c_int(0), C.c_int(0)
        opts = _opts(filter, limit, scheme)
        check(_lib.lib().fr_ens_run(self.h, 1 if looped else 0, opts, float(dt), int(nsteps), C.byref(nf),
                                    C.byref(nl)))
        return nf.value, nl.value

    def run_graph(self, dt: float, nsteps: int, nrep: int = 1, filter: bool = True, limit: bool = False,
                  scheme: str = "ssprk3"):
        """run() of nsteps steps captured once in a hipGraph (kept by the context) and replayed nrep
        times; nrep = 0 only captures."""
        nf, nl = C.c_int(0), C.c_int(0)
        opts = _opts(filter, limit, scheme)
        check(_lib.lib().fr_ens_run_graph(self.h, opts, float(dt), int(nsteps), int(nrep), C.byref(nf), C.byref(nl)))
        return nf.value, nl.value

    def sensor_filter(self, dt: float) -> int:
        n = C.c_int(0)
        check(_lib.lib().fr_ens_sensor_filter(self.h, float(dt), C.byref(n)))
        return n.value

    def positivity(self) -> int:
        n = C.c_int(0)
        check(_lib.lib().fr_ens_positivity(self.h, C.byref(n)))
        return n.value

    def autotune(self, dt: float) -> int:
        c = C.c_int(-1)
        check(_lib.lib().fr_ens_autotune(self.h, float(dt), C.byref(c)))
        return c.value
