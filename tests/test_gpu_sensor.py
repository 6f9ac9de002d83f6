"""GPU checks of the learned troubled-cell sensor: MFMA lane layout, forward and gradient parity
with the CPU oracle (oracle/mlp_oracle.py, same BF16 rounding points, FP32 accumulation), training
on the device, and the solver integration (features, learned filter)."""
import numpy as np
import pytest

import frhip
from frhip import ics
from frhip import sensor as S
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _features(n, seed=0):
    import torch

    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 16))
    x[:, ::3] *= 0.1
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    bits = M.bf16_bits(x)
    xt = torch.from_numpy(bits.view(np.int16).copy()).cuda().view(torch.bfloat16)
    return M.bf16(x), xt


def test_mfma_lane_layout():
    assert S.mfma_selftest() == 0.0


def test_forward_parity():
    p = S.init_params(3)
    xr, xt = _features(1000, seed=1)
    mlp = S.SensorMLP(p)
    got = mlp.forward(xt).cpu().numpy().astype(np.float64)
    ref = np.maximum(M.forward(p, xr)[0], 0.0)
    scale = np.max(np.abs(ref))
    d = np.abs(got - ref)
    # FP32 accumulation order can move an activation across a BF16 rounding boundary: rare cells
    # differ by up to one BF16 ulp of a hidden unit, everything else to FP32 round-off
    assert np.quantile(d, 0.99) <= 1e-5 * scale
    assert np.max(d) <= 2e-2 * scale


def test_gradient_parity():
    import torch

    p = S.init_params(5)
    xr, xt = _features(4096, seed=2)
    rng = np.random.default_rng(3)
    y = (rng.uniform(size=4096) > 0.7).astype(np.float32)
    yt = torch.from_numpy(y).cuda()
    mlp = S.SensorMLP(p)
    g, loss = mlp.grad(xt, yt)
    gr, lr_ = M.grad(p, xr, y.astype(np.float64), rounding=True)
    assert abs(loss - lr_) <= 1e-4 * lr_
    assert np.linalg.norm(g - gr) <= 2e-3 * np.linalg.norm(gr)
    # and it is the gradient of the loss: unrounded float64 gradient agrees to BF16 accuracy
    g64, _ = M.grad(p, xr, y.astype(np.float64), rounding=False)
    assert np.linalg.norm(g - g64) <= 3e-2 * np.linalg.norm(g64)


def test_training_reduces_loss():
    import torch

    xr, xt = _features(16384, seed=4)
    # separable target: the highest mode carries the label
    y = (np.abs(xr[:, 15]) > 0.25).astype(np.float32)
    yt = torch.from_numpy(y).cuda()
    mlp = S.SensorMLP(S.init_params(0))
    _, l0 = mlp.grad(xt, yt)
    l1 = mlp.train(xt, yt, batch=2048, nsteps=400, lr=3e-3)
    assert l1 < 0.5 * l0
    acc = ((mlp.forward(xt).cpu().numpy() > 0.5) == (y > 0.5)).mean()
    assert acc > 0.9


def test_solver_features_and_learned_filter():
    import torch

    ps = frhip.FRPSpace2D(0.0, 2.0, 40, 0.0, 1.0, 20, 3, 1, 1)
    u = ics.perturbed(ics.shock_vortex(ps, 1.4), seed=9, amplitude=0.02)
    s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
    s.set_state(u)
    x = S.features(s)
    # features: iV * rho normalised to unit norm, rounded to BF16
    ops = O.Ops(3)
    rho = np.stack([u[i, j, :, :, 0].ravel(order="F") for j in range(22) for i in range(42)])
    mod = rho @ ops.iV.T
    ref = M.bf16(mod / np.linalg.norm(mod, axis=1, keepdims=True))
    got = x.float().cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 4e-3  # one BF16 ulp at most
    mlp = S.SensorMLP(S.init_params(7))
    S.attach(s, mlp)
    lam = S.sensor_eval(s, "learned")
    assert torch.equal(lam, mlp.forward(x))
    # learned filter == oracle L2 filter with the per-cell strength 5e-4 * s
    strength = 5e-4 * lam.cpu().numpy().astype(np.float64).reshape((42, 22), order="F")
    s.sensor_filter("learned", 5e-4)
    uc = u.copy(order="F")
    orc = O.Fr2d(40, 20, 3, 2.0 / 40, 1.0 / 20, 1.4, "hll", "ghost")
    n = orc.filter_cells(uc, strength)
    assert n > 0
    ug = s.get_state()
    assert np.max(np.abs(ug - uc)) <= 1e-13 * np.max(np.abs(uc))


@pytest.mark.parametrize("bc,iters,scheme", [("ghost", 12, "ssprk3"), ("ghost", 7, "ssprk3"),
                                             ("periodic", 5, "ssprk3"), ("ghost", 5, "tsit5"),
                                             ("periodic", 4, "midpoint")])
@pytest.mark.parametrize("mode", ["off", "modal", "learned"])
def test_fused_filter_refresh_limit_equals_host_loop(mode, bc, iters, scheme):
    """fr2d_filter_refresh_limit (filter, inflow-keeping ghost refresh and the next limiter in one
    pass) and the natively enqueued loop fr2d_run_host_loop reproduce the unfused host loop of
    example/shock-vortex.jl:296-326 bitwise, for odd and even iteration counts and on periodic
    grids (whose ghost layer the loop refills after every step)."""
    import os

    ps = frhip.FRPSpace2D(0.0, 2.0, 48, 0.0, 1.0, 24, 3, 1, 1)
    u0 = ics.shock_vortex(ps, 1.4, shock_mach=1.3)
    dt, _ = ics.cfl_steps(u0, 1.4, ps.dx0, 0.05, 1.0)
    mlp = None
    if mode == "learned":
        if not os.path.exists(S.WEIGHTS):
            pytest.skip("no trained sensor weights in the package")
        mlp = S.SensorMLP.load()
    out = []
    for fused, native in ((False, False), (True, False), (True, True)):
        s = frhip.FREuler2D(ps, 1.4, "hll", bc)
        s.set_state(u0)
        if mlp is not None:
            S.attach(s, mlp)
        S.run_shock_vortex(s, dt, iters, mode, fused=fused, native=native, scheme=scheme)
        out.append(s.get_state())
        s.close()
    assert np.all(np.isfinite(out[0]))
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], out[2])  # fr2d_run_host_loop


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mode", ["off", "modal"])
@pytest.mark.parametrize("rule", ["periodic", "outflow", "inflow_left", "wave", "wave_y"])
def test_fused_pass_with_active_limiter(deg, mode, rule):
    """The fused pass on a state whose limiter corrects cells (density and energy pushed down at a
    few points) reproduces the unfused sensor_filter -> fill_ghost -> positivity sequence bitwise,
    for every degree and ghost rule."""
    nx, ny = 21, 13
    ps = frhip.FRPSpace2D(0.0, 1.0, nx, 0.0, 1.0, ny, deg, 1, 1)
    u = ics.isentropic_vortex(ps, 1.4, strength=5.0, center=(0.5, 0.5))
    u = ics.perturbed(u, seed=deg, amplitude=0.02)
    rng = np.random.default_rng(7 + deg)
    n = deg + 1
    for _ in range(8):
        i, j = rng.integers(1, nx + 1), rng.integers(1, ny + 1)
        k, l = rng.integers(0, n), rng.integers(0, n)
        u[i, j, k, l, 0] *= -0.2
        u[i, j, k, l, 3] *= 0.05
    out = {}
    for path in ("fused", "unfused", "no_limiter"):
        s = frhip.FREuler2D(ps, 1.4, "hll", "ghost")
        s.set_state(u)
        if path == "fused":
            s.filter_refresh_limit(mode, 5e-4, kappa=9.0, rule=rule)
        else:
            if mode != "off":
                s.sensor_filter(mode, 5e-4, kappa=9.0)
            s.fill_ghost(rule)
            if path == "unfused":
                s.positivity()
        out[path] = s.get_state()
        s.close()
    assert np.all(np.isfinite(out["fused"]))
    assert not np.array_equal(out["unfused"], out["no_limiter"])  # the limiter corrected cells
    assert np.array_equal(out["fused"], out["unfused"])
