"""Learned troubled-cell sensor: an MLP (16 normalised density modal coefficients -> 32 -> 32 -> 1,
ReLU, BF16 weights and inputs, FP32 accumulation) that stands in for the modal-decay sensor of
example/shock-vortex.jl:308-321.  It is trained on the GPU by libfrhip's own MFMA forward and
backward kernels (include/fr.h, "learned troubled-cell sensor") on cell features and labels that
the plain solver produces on seeded shock / vortex runs; the modal sensor remains the fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib, ics
from ._lib import FRError, check, dptr
from .euler2d import FREuler2D
from .space import FRPSpace2D

_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

_lib.SIGNATURES.update({
    "fr_mlp_nparam": (C.c_int, []),
    "fr_mlp_create": (C.c_int, [C.POINTER(_vp)]),
    "fr_mlp_destroy": (C.c_int, [_vp]),
    "fr_mlp_set_params": (C.c_int, [_vp, _fp]),
    "fr_mlp_get_params": (C.c_int, [_vp, _fp]),
    "fr_mlp_forward": (C.c_int, [_vp, _vp, _vp, C.c_longlong]),
    "fr_mlp_grad": (C.c_int, [_vp, _vp, _vp, C.c_longlong, _fp, _fp]),
    "fr_mlp_train": (C.c_int, [_vp, _vp, _vp, C.c_longlong, C.c_int, C.c_int, C.c_float, _fp]),
    "fr_mfma_selftest": (C.c_int, [_fp]),
    "fr2d_set_sensor_mlp": (C.c_int, [_vp, _vp]),
    "fr2d_sensor_features": (C.c_int, [_vp, _vp]),
    "fr2d_sensor_eval": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, _vp]),
})

NIN, NH = 16, 32
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sensor_mlp.npz")


def _bind():
    L = _lib.lib()
    for name in ("fr_mlp_nparam", "fr_mlp_create", "fr_mlp_destroy", "fr_mlp_set_params", "fr_mlp_get_params",
                 "fr_mlp_forward", "fr_mlp_grad", "fr_mlp_train", "fr_mfma_selftest", "fr2d_set_sensor_mlp",
                 "fr2d_sensor_features", "fr2d_sensor_eval"):
        res, args = _lib.SIGNATURES[name]
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def unpack(p):
    """Parameter vector -> (W1 [32,16], b1, W2 [32,32], b2, W3 [32], b3)."""
    p = np.asarray(p, dtype=np.float32)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN)
    o += NH * NIN
    b1 = p[o:o + NH]
    o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH)
    o += NH * NH
    b2 = p[o:o + NH]
    o += NH
    W3 = p[o:o + NH]
    o += NH
    return W1, b1, W2, b2, W3, p[o]


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (output bias 0)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / NIN)
    W2 = rng.standard_normal((NH, NH)) * np.sqrt(2.0 / NH)
    W3 = rng.standard_normal(NH) * np.sqrt(1.0 / NH)
    return np.concatenate([W1.ravel(), np.zeros(NH), W2.ravel(), np.zeros(NH), W3, [0.0]]).astype(np.float32)


def mfma_selftest() -> float:
    L = _bind()
    e = C.c_float(-1.0)
    check(L.fr_mfma_selftest(C.byref(e)))
    return e.value


class SensorMLP:
    """Device-resident MLP: parameters, Adam state, forward / gradient / training launches.
    Feature and label arrays are torch GPU tensors (features: torch.bfloat16 of shape (n, 16))."""

    def __init__(self, params=None):
        L = _bind()
        self.nparam = L.fr_mlp_nparam()
        h = C.c_void_p()
        check(L.fr_mlp_create(C.byref(h)))
        self.h = h
        self.set_params(init_params() if params is None else params)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _lib.lib().fr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        check(_lib.lib().fr_mlp_set_params(self.h, p.ctypes.data_as(_fp)))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparam, dtype=np.float32)
        check(_lib.lib().fr_mlp_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    @staticmethod
    def _chk(x, y=None):
        import torch

        assert x.is_cuda and x.dtype == torch.bfloat16 and x.is_contiguous() and x.shape[1] == NIN
        if y is not None:
            assert y.is_cuda and y.dtype == torch.float32 and y.is_contiguous() and y.shape[0] == x.shape[0]

    def forward(self, x):
        import torch

        self._chk(x)
        out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        check(_lib.lib().fr_mlp_forward(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), x.shape[0]))
        return out

    def grad(self, x, y):
        self._chk(x, y)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float(0.0)
        check(_lib.lib().fr_mlp_grad(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), x.shape[0],
                                     g.ctypes.data_as(_fp), C.byref(loss)))
        return g, loss.value

    def train(self, x, y, batch=4096, nsteps=1000, lr=2e-3) -> float:
        self._chk(x, y)
        loss = C.c_float(0.0)
        check(_lib.lib().fr_mlp_train(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), x.shape[0],
                                      int(batch), int(nsteps), float(lr), C.byref(loss)))
        return loss.value

    def save(self, path=WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path)
        return cls(z["params"])


def attach(solver: FREuler2D, mlp: SensorMLP | None):
    check(_bind().fr2d_set_sensor_mlp(solver.h, mlp.h if mlp is not None else None))
    solver._mlp = mlp  # the context keeps a pointer: keep the network alive with it


def features(solver: FREuler2D):
    """BF16 features of every cell ((nx+2)(ny+2) rows of 16, torch GPU tensor)."""
    import torch

    n = (solver.ps.nx + 2) * (solver.ps.ny + 2)
    x = torch.empty((n, NIN), dtype=torch.bfloat16, device="cuda")
    check(_bind().fr2d_sensor_features(solver.h, C.c_void_p(x.data_ptr())))
    return x


def sensor_eval(solver: FREuler2D, mode="modal", S0=None, kappa=9.0):
    """Per-cell strength multiplier (torch GPU float32 vector over the (nx+2)(ny+2) layout)."""
    import torch

    if S0 is None:
        S0 = -3.0 * np.log10(solver.ps.deg)
    n = (solver.ps.nx + 2) * (solver.ps.ny + 2)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    check(_bind().fr2d_sensor_eval(solver.h, _lib.SENSOR[mode], float(S0), float(kappa), C.c_void_p(out.data_ptr())))
    return out


# ------------------------------------------------------------------ shock / vortex runs
SHOCK_VORTEX = dict(x1=2.0, y1=1.0, nx=512, ny=256, deg=3, shock_mach=1.1, shock_x=0.5, vortex_mach=0.9, vortex_r=0.075,
          vortex_x=0.25, vortex_y=0.5, gamma=1.4, lam=5e-4, kappa=9.0, cfl=0.05)


def shock_vortex_solver(nx, ny, deg=3, gamma=1.4, flux="hll", **ic):
    ps = FRPSpace2D(0.0, SHOCK_VORTEX["x1"], nx, 0.0, SHOCK_VORTEX["y1"], ny, deg, 1, 1)
    u0 = ics.shock_vortex(ps, gamma, **ic)
    s = FREuler2D(ps, gamma, flux, "ghost")
    s.set_state(u0)
    return ps, u0, s


def run_shock_vortex(s: FREuler2D, dt, nsteps, mode="modal", lam=SHOCK_VORTEX["lam"], kappa=SHOCK_VORTEX["kappa"], callback=None,
                     every=0, fused=True, native=False, scheme="ssprk3"):
    """Host loop of example/shock-vortex.jl:296-326 on the device state: positivity limiter on
    every cell (:298-303), SSP-RK3 step, sensor + filter (mode "modal" / "learned" / "off",
    :306-321), ghost refresh of the bottom, top and right sides with the inflow column kept
    (:322-325).  fused: the filter, the ghost refresh and the next iteration's limiter run as one
    pass over the state (fr2d_filter_refresh_limit, the same arithmetic); the last iteration ends
    with the filter and the ghost refresh.  native: the whole fused loop is enqueued by the
    library (fr2d_run_host_loop, no callback) with one synchronisation at the end.  scheme: the
    stepper of every iteration (SSP-RK3 by default; the script itself steps with Midpoint)."""
    S0 = -3.0 * np.log10(s.ps.deg)
    if native:
        if callback is not None or not fused:
            raise ValueError("run_shock_vortex: the native loop is the fused loop without callbacks")
        s.run_host_loop(dt, nsteps, mode, lam, S0, kappa, "inflow_left", scheme=scheme)
        return
    s.positivity()
    for n in range(nsteps):
        s.step(dt, 1, scheme)
        if fused and n + 1 < nsteps:
            s.filter_refresh_limit(mode, lam, S0, kappa, "inflow_left")
        else:
            if mode != "off":
                _filter(s, mode, lam, S0, kappa)
            s.fill_ghost("inflow_left")
            if n + 1 < nsteps:
                s.positivity()
        if callback is not None and every and (n + 1) % every == 0:
            callback(n + 1)


def _filter(s, mode, lam, S0, kappa):
    check(_lib.lib().fr2d_sensor_filter(s.h, _lib.SENSOR[mode], float(lam), float(S0), float(kappa), None))


def collect_training_data(nruns=200, nx=128, ny=64, nsteps=240, every=40, seed=1234, max_neg_ratio=1.0):
    """Cell features and modal-sensor labels from `nruns` seeded shock / vortex runs with the vortex
    position and strength drawn uniformly within +-20% of the shock-vortex case values and the shock Mach number
    in [1.05, 1.5].  Every `every` steps all troubled cells and an equal number of random smooth
    cells are kept.  Returns torch GPU tensors (x bf16 (n, 16), y float32 (n,))."""
    import torch

    rng = np.random.default_rng(seed)
    xs, ys = [], []
    failed = 0
    for r in range(nruns):
        f = lambda: rng.uniform(0.8, 1.2)  # noqa: E731
        ic = dict(shock_mach=rng.uniform(1.05, 1.5), vortex_mach=SHOCK_VORTEX["vortex_mach"] * f(),
                  vortex_r=SHOCK_VORTEX["vortex_r"] * f(), vortex_x=SHOCK_VORTEX["vortex_x"] * f(), vortex_y=SHOCK_VORTEX["vortex_y"] * f())
        ps, u0, s = shock_vortex_solver(nx, ny, **ic)
        dt, _ = ics.cfl_steps(u0, SHOCK_VORTEX["gamma"], ps.dx0, SHOCK_VORTEX["cfl"], 1.0)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(rng.integers(1 << 31)))
        run_x, run_y = [], []

        def snap(_n):
            x = features(s)
            y = sensor_eval(s, "modal", kappa=SHOCK_VORTEX["kappa"])
            pos = torch.nonzero(y > 0.5).squeeze(1)
            neg = torch.nonzero(y <= 0.5).squeeze(1)
            k = min(neg.numel(), max(int(max_neg_ratio * pos.numel()), 64))
            pick = neg[torch.randperm(neg.numel(), generator=gen, device="cuda")[:k]]
            idx = torch.cat([pos, pick])
            run_x.append(x[idx])
            run_y.append(y[idx])

        try:
            snap(0)
            # unfused loop: the snapshots see the state after the ghost refresh, before the limiter
            run_shock_vortex(s, dt, nsteps, "modal", callback=snap, every=every, fused=False)
            xs.extend(run_x)
            ys.extend(run_y)
        except FRError:
            failed += 1  # a run whose limiter failed contributes no samples
        s.close()
    x, y = torch.cat(xs), torch.cat(ys)
    keep = torch.isfinite(x.float()).all(dim=1) & torch.isfinite(y)
    collect_training_data.last_stats = dict(failed_runs=failed, dropped_samples=int((~keep).sum().item()))
    return x[keep].contiguous(), y[keep].contiguous()


def train_sensor(x, y, nsteps=3000, batch=4096, lr=2e-3, seed=0, val_frac=0.1):
    """Adam training on the GPU (MFMA forward / backward kernels); returns (mlp, stats)."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    perm = torch.randperm(x.shape[0], generator=g, device="cuda")
    x, y = x[perm].contiguous(), y[perm].contiguous()
    nv = int(val_frac * x.shape[0])
    xv, yv, xt, yt = x[:nv].contiguous(), y[:nv].contiguous(), x[nv:].contiguous(), y[nv:].contiguous()
    mlp = SensorMLP(init_params(seed))
    loss = mlp.train(xt, yt, batch=batch, nsteps=nsteps, lr=lr)
    pred = mlp.forward(xv)
    acc = float(((pred > 0.5) == (yv > 0.5)).float().mean().item()) if nv else float("nan")
    return mlp, dict(train_loss=loss, val_accuracy=acc, n_train=int(xt.shape[0]), n_val=nv)


# ------------------------------------------------------------------ evaluation on the shock-vortex case
def cell_means(ps, u, var=0):
    """Element means of one variable over the interior cells (Gauss weights wp / 4)."""
    return np.einsum("ijkl,kl->ij", u[1:-1, 1:-1, :, :, var], ps.wp / 4.0)


def restrict_means(fine_means, factor):
    nx, ny = fine_means.shape
    return fine_means.reshape(nx // factor, factor, ny // factor, factor).mean(axis=(1, 3))


def evaluate_shock_vortex(mlp: SensorMLP | None, nx=SHOCK_VORTEX["nx"], ny=SHOCK_VORTEX["ny"], t_end=0.4, fine=2, modes=("modal", "learned", "off"),
                reference=None):
    """Shock-vortex interaction (Mach-1.1 stationary shock at x = 0.5, Lamb-Oseen vortex of
    Mach 0.9 and radius 0.075 at x = 0.25) on nx x ny quads, p = 3, to t_end, with each sensor mode;
    density L1 error (src/tools.jl:43-45 on element means, sum |rho - rho_ref| dx dy) against a
    `fine`-times refined run with the modal sensor, and the wall time per step of each mode.
    Returns a dict; `reference` may carry a precomputed fine solution's element means."""
    import time

    import torch

    out = {}
    ic = {k: SHOCK_VORTEX[k] for k in ("shock_mach", "shock_x", "vortex_mach", "vortex_r", "vortex_x", "vortex_y")}
    if reference is None:
        psf, u0f, sf = shock_vortex_solver(nx * fine, ny * fine, **ic)
        dtf, nf = ics.cfl_steps(u0f, SHOCK_VORTEX["gamma"], psf.dx0, SHOCK_VORTEX["cfl"], t_end)
        run_shock_vortex(sf, dtf, nf, "modal")
        reference = restrict_means(cell_means(psf, sf.get_state()), fine)
        sf.close()
        out["fine_steps"] = nf
    for mode in modes:
        ps, u0, s = shock_vortex_solver(nx, ny, **ic)
        if mode == "learned":
            if mlp is None:
                continue
            attach(s, mlp)
        dt, n = ics.cfl_steps(u0, SHOCK_VORTEX["gamma"], ps.dx0, SHOCK_VORTEX["cfl"], t_end)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_shock_vortex(s, dt, n, mode, native=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        uf = s.get_state()
        finite = bool(np.all(np.isfinite(uf)))
        rho = cell_means(ps, uf)
        l1 = float(np.sum(np.abs(rho - reference)) * ps.dx0 * ps.dy0) if finite else float("inf")
        out[mode] = dict(L1_density=l1, ms_per_step=1e3 * el / n, steps=n, finite=finite)
        s.close()
    out["reference_means"] = reference
    return out
