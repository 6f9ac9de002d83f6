"""2D Euler FR solver on uniform quads: the drop-in right-hand side and the device-resident
integrator of libfrhip (include/fr.h, "2D Euler on uniform quads").

The reference's user-level right-hand side is ``dudt!(du, u, p, t)`` of
example/shock-vortex.jl:26-118 with the scratch tuple ``p`` of :138-153; here the scratch lives in
the library context and ``FREuler2D.dudt(du, u, p, t)`` keeps the same argument meaning.
"""
from __future__ import annotations

import contextlib
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr
from .space import FRPSpace2D


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class FREuler2D:
    """Context of one 2D Euler problem.

    ps      FRPSpace2D with one ghost layer (ngx = ngy = 1)
    gamma   ratio of specific heats
    flux    "hll" (reference default), "rusanov" or "roe"
    bc      "ghost" (ghost layer is data, refreshed by the caller as the reference scripts do)
            or "periodic" (neighbours wrap inside the kernels)
    """

    def __init__(self, ps: FRPSpace2D, gamma: float = 1.4, flux: str = "hll", bc: str = "ghost"):
        if getattr(ps, "ngx", 1) != 1 or getattr(ps, "ngy", 1) != 1:
            raise ValueError("FREuler2D needs an FRPSpace2D with one ghost layer (ngx = ngy = 1)")
        self.ps = ps
        self.gamma = float(gamma)
        self.flux = flux
        self.bc = bc
        n = ps.deg + 1
        self.nsp = n
        self.shape = (ps.nx + 2, ps.ny + 2, n, n, 4)
        self._ops = [np.ascontiguousarray(a, dtype=np.float64) for a in (
            ps.ll, ps.lr, np.asarray(ps.dl).ravel(order="F"), ps.dhl, ps.dhr,
            np.asarray(ps.V).ravel(order="F"), np.asarray(ps.iV).ravel(order="F"))]
        d = _lib.Fr2dDesc()
        d.nx, d.ny, d.deg = ps.nx, ps.ny, ps.deg
        d.gamma = self.gamma
        d.flux = _lib.FLUX[flux]
        d.bc = _lib.BC[bc]
        point = getattr(ps, "iJ_elem", 0) is None  # per-point metrics (an FRPSpace2D over any mesh)
        iJ = np.eye(2) if point else np.asarray(ps.iJ_elem, dtype=np.float64)
        rect = iJ[0, 1] == 0.0 and iJ[1, 0] == 0.0
        d.iJ11 = float(iJ[0, 0]) if rect else 1.0
        d.iJ22 = float(iJ[1, 1]) if rect else 1.0
        d.ll, d.lr, d.dl, d.dhl, d.dhr, d.V, d.iV = [dptr(a) for a in self._ops]
        h = C.c_void_p()
        check(_lib.lib().fr2d_create(C.byref(h), C.byref(d)))
        self.h = h
        self.len = _lib.lib().fr2d_state_len(h)
        if point:
            self.set_metrics(ps.iJ)
        elif not rect:
            self.set_geometry(iJ)

    def set_geometry(self, iJ, n1=None, n2=None):
        """Uniform parallelogram cells (dev/parallelogram.jl:78-162): iJ is the 2x2 inverse Jacobian
        [[r_x, r_y], [s_x, s_y]] of every cell, n1 / n2 the normals of the x / y faces (None:
        grad r / grad s).  Fluxes are solved in the face frames and scaled by (iJ n)."""
        iJ = np.asarray(iJ, dtype=np.float64).reshape(2, 2)
        arrs = [np.ascontiguousarray(iJ.ravel(order="F"))]
        arrs += [None if n is None else np.ascontiguousarray(n, dtype=np.float64).reshape(2) for n in (n1, n2)]
        self._geom = arrs  # kept alive for the duration of the call only; the library copies them
        check(_lib.lib().fr2d_set_geometry(self.h, *[None if a is None else dptr(a) for a in arrs]))

    def set_metrics(self, iJ):
        """Point-varying metrics: iJ[i, j, k, l] is the 2x2 inverse Jacobian of solution point
        (k, l) of cell (i, j), an array of shape (nx+2, ny+2, nsp, nsp, 2, 2) (FRPSpace2D.iJ over a
        curvilinear mesh; src/struct.jl:134-142).  The residual is example/shock-vortex.jl:26-118
        with each point's metric (fr2d_set_metrics)."""
        iJ = np.asarray(iJ, dtype=np.float64)
        want = (self.ps.nx + 2, self.ps.ny + 2, self.nsp, self.nsp, 2, 2)
        if iJ.shape != want:
            raise ValueError(f"iJ must have shape {want}")
        flat = np.ascontiguousarray(iJ.ravel(order="F"))
        check(_lib.lib().fr2d_set_metrics(self.h, dptr(flat)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- streams
    def use_torch_stream(self, stream=None):
        """Launch on a torch CUDA/HIP stream (default: torch's current stream; its default stream
        is HIP's null stream, handle 0)."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        check(_lib.lib().fr2d_set_stream(self.h, C.c_void_p(s.cuda_stream)))

    def use_own_stream(self):
        """Launch on the context's own non-blocking stream again (the default)."""
        check(_lib.lib().fr2d_set_stream(self.h, C.c_void_p(2 ** (8 * C.sizeof(C.c_void_p)) - 1)))

    def stream_handle(self) -> int:
        return _lib.lib().fr2d_get_stream(self.h) or 0

    @contextlib.contextmanager
    def _torch_ordered(self):
        """Orders a call on torch tensors with torch's current stream: the context's stream waits
        for the work queued so far on torch's current stream (which wrote the inputs), and torch's
        current stream waits for the call's work (which writes the outputs).  Nothing is added when
        the context already runs on torch's current stream."""
        import torch

        cur = torch.cuda.current_stream()
        mine = self.stream_handle()
        if mine == cur.cuda_stream:
            yield
            return
        ctx = torch.cuda.default_stream() if mine == 0 else torch.cuda.ExternalStream(mine)
        ctx.wait_stream(cur)
        yield
        cur.wait_stream(ctx)

    # ---------------------------------------------------------------- right-hand side
    def _check_host(self, a):
        if a.shape != self.shape or a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
            raise ValueError(f"expected a Fortran-ordered float64 array of shape {self.shape}")

    def _check_dev(self, a):
        if tuple(a.shape) != (self.len,) and tuple(a.shape) != self.shape[::-1]:
            if a.numel() != self.len:
                raise ValueError("device array has the wrong size")
        if str(a.dtype) != "torch.float64" or not a.is_cuda or not a.is_contiguous():
            raise ValueError("device arrays must be contiguous float64 torch tensors on the GPU")

    def dudt(self, du, u, p=None, t: float = 0.0):
        """In-place du = L(u) with du = 0 on the ghost layer (example/shock-vortex.jl:26-118).
        numpy (host) arrays in the reference layout, or flat/reversed-shape torch GPU tensors."""
        if _is_torch(u):
            self._check_dev(u)
            self._check_dev(du)
            with self._torch_ordered():
                check(_lib.lib().fr2d_euler_rhs_dev(self.h, C.c_void_p(u.data_ptr()), C.c_void_p(du.data_ptr()),
                                                    float(t)))
        else:
            self._check_host(u)
            self._check_host(du)
            check(_lib.lib().fr2d_euler_rhs(self.h, dptr(u), dptr(du), float(t)))
        return None

    __call__ = dudt

    def rhs(self, u):
        du = np.zeros(self.shape, order="F")
        self.dudt(du, np.asfortranarray(u, dtype=np.float64))
        return du

    # ---------------------------------------------------------------- integrator
    def set_state(self, u):
        if _is_torch(u):
            self._check_dev(u)
            with self._torch_ordered():
                check(_lib.lib().fr2d_set_state_dev(self.h, C.c_void_p(u.data_ptr())))
        else:
            u = np.asfortranarray(u, dtype=np.float64)
            self._check_host(u)
            check(_lib.lib().fr2d_set_state(self.h, dptr(u)))

    def get_state(self):
        out = np.zeros(self.shape, order="F")
        check(_lib.lib().fr2d_get_state(self.h, dptr(out)))
        return out

    def state_ptr(self) -> int:
        """Device pointer of the state, in the padded layout of state_pitch()."""
        return _lib.lib().fr2d_state_dev(self.h) or 0

    def state_pitch(self):
        """(pitch, plane): element (i, j) of plane p of the device state is at
        i + pitch*j + plane*p (rows padded to a multiple of 16 elements)."""
        p, q = C.c_int(0), C.c_longlong(0)
        check(_lib.lib().fr2d_state_pitch(self.h, C.byref(p), C.byref(q)))
        return p.value, q.value

    def stagger(self) -> int:
        """Staggered-start delay of the RK stage launches picked on the device by the context's
        first stepping call (units of 512 clocks per workgroup slot; 0 off, -1 not tuned yet)."""
        v = C.c_int(0)
        check(_lib.lib().fr2d_get_stagger(self.h, C.byref(v)))
        return v.value

    def set_stagger(self, stagger: int) -> None:
        """Fix the staggered-start delay (0..16; 0 off) instead of the first stepping call's device
        timing, or (-1) have the next stepping call time the candidates again."""
        check(_lib.lib().fr2d_set_stagger(self.h, int(stagger)))

    def step(self, dt: float, nsteps: int = 1, scheme: str = "ssprk3"):
        check(_lib.lib().fr2d_step(self.h, _lib.RK[scheme], float(dt), int(nsteps)))

    def step_graph(self, dt: float, nsteps: int, nrep: int = 1, scheme: str = "ssprk3"):
        check(_lib.lib().fr2d_step_graph(self.h, _lib.RK[scheme], float(dt), int(nsteps), int(nrep)))

    def fill_ghost(self, rule: str = "periodic"):
        """Ghost refresh of the device state: "periodic", "outflow", "inflow_left"
        (shock-vortex.jl:322-325), "wave" / "wave_y" (euler2d_wave.jl:127-132 / :153-158)."""
        check(_lib.lib().fr2d_fill_ghost(self.h, _lib.GHOST[rule]))

    def sensor_filter(self, mode: str = "modal", lam: float = 5e-4, S0: float | None = None, kappa: float = 9.0):
        """Modal sensor + L2 filter of example/shock-vortex.jl:308-321; returns #filtered cells."""
        if S0 is None:
            S0 = -3.0 * np.log10(self.ps.deg)
        n = C.c_int(0)
        check(_lib.lib().fr2d_sensor_filter(self.h, _lib.SENSOR[mode], float(lam), float(S0), float(kappa), C.byref(n)))
        return n.value

    def filter_refresh_limit(self, mode: str = "modal", lam: float = 5e-4, S0: float | None = None,
                             kappa: float = 9.0, rule: str = "inflow_left") -> int:
        """Sensor + filter, ghost refresh `rule` and positivity limiter in one pass (the end of one
        shock-vortex.jl:296-326 iteration and the start of the next); returns #filtered cells."""
        if S0 is None:
            S0 = -3.0 * np.log10(self.ps.deg)
        n = C.c_int(0)
        check(_lib.lib().fr2d_filter_refresh_limit(self.h, _lib.SENSOR[mode], float(lam), float(S0), float(kappa),
                                                   _lib.GHOST[rule], C.byref(n)))
        return n.value

    def run_host_loop(self, dt: float, nsteps: int, mode: str = "modal", lam: float = 5e-4,
                      S0: float | None = None, kappa: float = 9.0, rule: str = "inflow_left",
                      scheme: str = "ssprk3") -> int:
        """nsteps iterations of the shock-vortex.jl:296-326 host loop enqueued natively with one
        synchronisation at the end (fr2d_run_host_loop); returns #filtered cells over the loop."""
        if S0 is None:
            S0 = -3.0 * np.log10(self.ps.deg)
        n = C.c_int(0)
        check(_lib.lib().fr2d_run_host_loop(self.h, _lib.RK[scheme], float(dt), int(nsteps), _lib.SENSOR[mode],
                                            float(lam), float(S0), float(kappa), _lib.GHOST[rule], C.byref(n)))
        return n.value

    def positivity(self):
        """Positivity limiter of src/dissipation.jl:125-206 on every interior cell."""
        check(_lib.lib().fr2d_positivity(self.h))

    def check_state(self):
        check(_lib.lib().fr2d_check_state(self.h))

    def primitives(self, out=None):
        """Post-processing of example/shock-vortex.jl:333-352 on the device state.

        Returns (x, y, sol): solution-point coordinates and the primitive variables
        (rho, U, V, 1/lambda), lambda = rho/(2p), of every solution point on the global grid of
        shape (nx*nsp, ny*nsp) (sol has a trailing axis of 4), element-major along each axis like
        the reference.  With `out` (a torch GPU tensor of nx*nsp*ny*nsp*4 doubles) the field is
        written on the device instead and (x, y, out) is returned."""
        nx, ny, n = self.ps.nx, self.ps.ny, self.ps.nsp
        xpg = np.asarray(self.ps.xpg)
        if xpg.shape[0] == nx + 2:
            xpg = xpg[1:-1, 1:-1]
        x = np.transpose(xpg[..., 0], (0, 2, 1, 3)).reshape(nx * n, ny * n)
        y = np.transpose(xpg[..., 1], (0, 2, 1, 3)).reshape(nx * n, ny * n)
        if out is not None:
            if not _is_torch(out) or out.numel() != nx * n * ny * n * 4:
                raise ValueError("primitives: out must be a GPU tensor of nx*nsp*ny*nsp*4 doubles")
            if str(out.dtype) != "torch.float64" or not out.is_cuda or not out.is_contiguous():
                raise ValueError("primitives: out must be a contiguous float64 tensor on the GPU")
            with self._torch_ordered():
                check(_lib.lib().fr2d_primitives_dev(self.h, C.c_void_p(out.data_ptr())))
            return x, y, out
        sol = np.zeros((nx * n, ny * n, 4), order="F")
        check(_lib.lib().fr2d_primitives(self.h, dptr(sol)))
        return x, y, sol

    def synchronize(self):
        import torch

        torch.cuda.synchronize()
