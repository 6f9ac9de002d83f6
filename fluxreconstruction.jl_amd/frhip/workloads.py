"""The measured workloads besides the 2D headline stage, as functions that bench.py calls and as a
command line that runs one of them alone (for a kernel profile of that workload):

    python3 -m frhip.workloads ensemble       512 seeded 1D Euler Riemann problems, SSP-RK3
    python3 -m frhip.workloads bgk            2D BGK 128^2 p = 2 x 24^2 velocities, one member
    python3 -m frhip.workloads bgk64          the same, 64 members with log-uniform tau
    python3 -m frhip.workloads shock_vortex --mode learned --steps 600
                                              the shock-vortex host loop on 512 x 256, p = 3
    python3 -m frhip.workloads vortex_grid --n 2048
                                              the headline vortex stage on an n x n grid

Each prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import ics, sensor
from .bgk import BGK1D, BGK2D, VSpace1D, bgk_wave_initial, maxwellian_state, velocity_grid
from .euler2d import FREuler2D
from .fr1d import Ensemble1D
from .space import CSpace2D, FRPSpace1D, FRPSpace2D


def riemann_ensemble(steps=10, nprob=512, seed=2024):
    """nprob seeded 1D Euler Riemann problems (64..2048 cells, p in {2,3,4}, piecewise-constant
    data with the jump on a cell face), Rusanov, SSP-RK3, run as the reference's per-step loop
    (example/euler_highlevel.jl:37-54: sensor + filter before every step): one batched launch per
    pass (autotuned mapping; eager and as a captured hipGraph, captured during the warm-up) vs one
    launch per member per pass."""
    import torch

    ncell, deg, dx, members = ics.riemann_ensemble(nprob, seed=seed)
    e = Ensemble1D(ncell, deg, dx, 1.4, "rusanov", "dirichlet")
    flat = e.pack(members)
    lam = max(float(np.max(np.abs(m[..., 1] / m[..., 0]) + np.sqrt(1.4 * 0.4 * (m[..., 2] - 0.5 * m[..., 1] ** 2 / m[..., 0]) / m[..., 0]))) for m in members)
    dt = 0.05 * float(np.min(dx)) / lam
    dof = int(sum(int(n) * (int(p) + 1) * 3 for n, p in zip(ncell, deg)))
    e.set_state(flat)
    choice = e.autotune(dt)
    res, flagged = {}, {}
    for name, fn in (("looped", lambda: e.run(dt, steps, looped=True)),
                     ("batched", lambda: e.run(dt, steps)),
                     ("batched_graph", lambda: e.run_graph(dt, steps, 1))):
        e.set_state(flat)
        fn()  # warm-up; captures the graph of the graph variant
        e.set_state(flat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        flagged[name] = fn()[0]
        torch.cuda.synchronize()
        res[name] = dof * 3 * steps / (time.perf_counter() - t0)
    out = e.get_state()
    finite = sum(bool(np.all(np.isfinite(e.member(out, b)))) for b in range(len(members)))
    return {"workload": f"{nprob} 1D Euler Riemann problems, 64..2048 cells, p in {{2,3,4}}, Rusanov, SSP-RK3, "
                        "sensor + filter before every step",
            "dof": dof, "steps": steps, "dt": dt, "unit": "DOF·RK-stage/s", "looped": res["looped"],
            "batched": res["batched"], "batched_graph": res["batched_graph"],
            "speedup_vs_looped": max(res["batched"], res["batched_graph"]) / res["looped"],
            "autotune_choice": choice, "members_finite": finite, "filtered_cells": flagged["batched_graph"]}


def bgk(steps=3, members=1, seed=5):
    """2D BGK, 128^2 periodic quads p = 2 x 24^2 Gauss-Hermite velocities; one member with
    Kn = tau = 1e-2, or `members` members with tau log-uniform in [1e-3, 1e-1].  Algorithmic bytes
    per value and stage: 8 (moments pass) + 21.33 (transport: 16 in stage 1, 24 in stages 2-3)."""
    import torch

    u, wu = velocity_grid(24)
    dt = 0.05 * (1.0 / 128) / float(np.max(np.abs(u)))
    taus = np.array([1e-2]) if members == 1 else np.exp(np.random.default_rng(seed).uniform(np.log(1e-3), np.log(1e-1), members))
    g = BGK2D(128, 128, 2, u, wu, u, wu, taus)
    one = torch.from_numpy(maxwellian_state(128, 128, 2, u, u)).cuda()
    g.set_state(one.repeat(len(taus)))
    del one
    g.step(dt, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(dt, steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    vals = g.len
    g.close()
    return {"members": len(taus), "values": int(vals), "value": vals * 3 * steps / el,
            "alg_GBs": vals * 3 * steps * (8 + 64 / 3) / el / 1e9, "steps": steps}


def bgk1d(steps=20, members=1, nx=8192, nu=256):
    """The 1D BGK model of example/bgk_wave.jl as written (Radau FR p = 2, VSpace1D on [-5, 5],
    gamma = 3, tau = 1e-2, the script's initial Maxwellians) on nx cells x nu velocities, one member
    or `members` copies (one launch per stage for all); SSP-RK3, wall time over eager steps.
    Algorithmic bytes per value and stage: 8 (moments) + 21.33 (transport: 16 / 24 / 24)."""
    ps = FRPSpace1D(0.0, 1.0, nx, 2)
    vs = VSpace1D(-5.0, 5.0, nu)
    f0 = bgk_wave_initial(ps, vs)
    g = BGK1D(ps, vs, tau=[1e-2] * members)
    g.set_state(np.tile(f0.ravel(order="F"), members))
    dt = 0.1 * (1.0 / nx) / 5.0
    g.step(dt, 2, "ssprk3")
    t0 = time.perf_counter()
    g.step(dt, steps, "ssprk3")
    el = time.perf_counter() - t0
    vals = g.len
    g.close()
    stage = el / (3 * steps)
    return {"members": members, "values": int(vals), "stage_us": stage * 1e6, "value": vals / stage,
            "alg_GBs": vals * (8 + 64 / 3) / stage / 1e9, "steps": steps}


def bgk1d_script(steps=1000):
    """example/bgk_wave.jl as the script runs it: 20 cells, p = 2, 100 velocities on [-5, 5],
    tau = 1e-2, Midpoint with dt = cfl dx / u1 = 1e-3 for 1000 steps (t = 1).  Wall time of the
    whole run from the same initial state on the streaming path (two launches per stage) eagerly and
    as replays of a captured 100-step graph, and on the LDS-resident path (one launch per 256
    steps, the default for this size); all three end in the same bits."""
    ps = FRPSpace1D(0.0, 1.0, 20, 2)
    vs = VSpace1D(-5.0, 5.0, 100)
    f0 = bgk_wave_initial(ps, vs)
    g = BGK1D(ps, vs)
    dt = 0.1 * (1.0 / 20) / 5.0
    out, states = {}, []
    for mode in ("stream", "stream_graph", "resident"):
        g.set_path("resident" if mode == "resident" else "stream")
        g.set_state(f0)
        g.step(dt, 100, "midpoint")  # warm-up (and the graph's capture)
        g.step_graph(dt, 100, "midpoint", chunk=100)
        g.set_state(f0)
        t0 = time.perf_counter()
        if mode == "stream_graph":
            g.step_graph(dt, steps, "midpoint", chunk=100)
        else:
            g.step(dt, steps, "midpoint")
        el = time.perf_counter() - t0
        out[mode] = {"run_ms": el * 1e3, "stage_us": el / (2 * steps) * 1e6}
        states.append(g.get_state())
    g.close()
    # a tau ensemble of 256 such members (tau log-uniform in [3e-3, 1e-1]; Midpoint at this dt
    # breaks down below tau = 1.5e-3): one resident workgroup per member against the streaming
    # launches over all members
    taus = list(np.exp(np.random.default_rng(5).uniform(np.log(3e-3), np.log(1e-1), 256)))
    e = BGK1D(ps, vs, tau=taus)
    ens, est = {}, []
    for path in ("stream", "resident"):
        e.set_path(path)
        e.set_state(np.tile(f0.ravel(order="F"), len(taus)))
        e.step(dt, 20, "midpoint")
        t0 = time.perf_counter()
        e.step(dt, 200, "midpoint")
        el = time.perf_counter() - t0
        ens[path] = {"stage_us": el / 400 * 1e6, "values_stage_per_s": e.len * 400 / el}
        est.append(e.get_state())
    e.close()
    return {"steps": steps, "values": int(f0.size), **out,
            "resident_speedup": out["stream"]["run_ms"] / out["resident"]["run_ms"],
            "bitwise_equal": bool(all(np.array_equal(states[0], x) for x in states[1:])),
            "ensemble256_200_steps": {**ens, "resident_speedup": ens["resident"]["values_stage_per_s"] /
                                      ens["stream"]["values_stage_per_s"],
                                      "bitwise_equal": bool(np.array_equal(est[0], est[1])),
                                      "finite": bool(np.all(np.isfinite(est[0])))}}


def fr1d_small(steps=1000):
    """The reference's small 1D runs, wall time of the whole run on the batched path (a launch per
    stage) and on the member path (one workgroup runs every stage, a launch per 3072 stages), from
    the same state; both end in the same bits:
    - advection, 100 cells, p = 2, RK4, dt = 0.1 dx (example/advection_lowlevel.jl's setup, 1000
      steps to t = 1);
    - Euler (HLL, Dirichlet), Sod's problem on 100 cells, p = 2, Midpoint, dt = 1e-4 (1000 steps)."""
    from . import fr1d as _fr1d

    out = {}
    ps = FRPSpace1D(0.0, 1.0, 100, 2)
    cases = {
        "advection_100_p2_rk4": (lambda: _fr1d.FRAdvectionProblem(ps, 1.0, "period"),
                                 np.asfortranarray(np.sin(2 * np.pi * ps.xpg)), 0.1 / 100, "rk4", 4),
    }
    x = ps.xpg
    prim = np.stack([np.where(x < 0.5, 1.0, 0.125), np.zeros_like(x), np.where(x < 0.5, 0.5, 0.625)], -1)
    cases["sod_100_p2_midpoint"] = (lambda: _fr1d.FREulerProblem(ps, 1.4, "dirichlet", "hll"),
                                    np.asfortranarray(ics.prim_conserve1(prim, 1.4)), 1e-4, "midpoint", 2)
    for name, (mk, u0, dt, scheme, nst) in cases.items():
        res, states = {}, []
        for path in ("batched", "member"):
            g = mk()
            g.set_path(path)
            g.set_state(u0)
            g.step(dt, 50, scheme)
            g.set_state(u0)
            g.get_state()
            t0 = time.perf_counter()
            g.step(dt, steps, scheme)
            states.append(g.get_state())
            el = time.perf_counter() - t0
            res[path] = {"run_ms": el * 1e3, "stage_us": el / (nst * steps) * 1e6}
            g.close()
        out[name] = {**res, "member_speedup": res["batched"]["run_ms"] / res["member"]["run_ms"],
                     "bitwise_equal": bool(np.array_equal(states[0], states[1]))}
    return out


def polar_stage(n=256, steps=20):
    """Cells with point-varying metrics: the p = 3 SSP-RK3 stage on the polar mesh
    CSpace2D(1, 6, n, 0, pi, n, 1, 1) (fr2d_set_metrics), Rusanov, ghost layer as data, a smooth
    positive state, 20 graph-replayed steps after 300 eager ones (dt = 1e-7 keeps the run finite),
    HIP events on the launch stream.  Algorithmic bytes: 21.33 per DOF and stage + 8 for the metric."""
    import torch

    ps = FRPSpace2D(CSpace2D(1.0, 6.0, n, 0.0, np.pi, n, 1, 1), 3)
    x, y = ps.xpg[..., 0], ps.xpg[..., 1]
    rho = 1.0 + 0.2 * np.sin(x) * np.cos(y)
    uu, vv, p = 0.3 * np.cos(y), -0.2 * np.sin(x), 1.0 + 0.1 * np.cos(x + y)
    u0 = np.asfortranarray(np.stack([rho, rho * uu, rho * vv, p / 0.4 + 0.5 * rho * (uu * uu + vv * vv)], axis=-1))
    s = FREuler2D(ps, 1.4, "rusanov", "ghost")
    st = torch.cuda.Stream()
    s.use_torch_stream(st)
    s.set_state(u0)
    dt = 1e-7
    with torch.cuda.stream(st):
        s.step_graph(dt, steps, 1)
        s.set_state(u0)
        s.step(dt, 300)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        s.step_graph(dt, steps, 1)
        e1.record(st)
        torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / (3 * steps)
    dof = n * n * 16 * 4
    s.check_state()
    s.close()
    return {"grid": f"polar {n}x{n}", "dof": dof, "stage_us": us, "dof_stage_per_s": dof / (us * 1e-6),
            "alg_GBs": (64.0 / 3.0 + 8.0) * dof / (us * 1e-6) / 1e9}


def parallelogram_stage(n=256, steps=20, shear=0.4, stagger=None):
    """Uniform parallelogram cells (fr2d_set_geometry; dev/parallelogram.jl's face frames): the
    p = 3 SSP-RK3 stage on n x n cells of [0, 10]^2 sheared by `shear` (edges a = (dx, 0),
    b = (shear dx, dy), x faces in the frame of grad r, y faces in that of grad s), Rusanov,
    periodic, the isentropic vortex; 20 graph-replayed steps after 300 eager ones, HIP events on the
    launch stream.  Algorithmic bytes: 21.33 per DOF and stage (the metric is two scalars).
    stagger: a fixed staggered-start delay instead of the device timing (fr2d_set_stagger)."""
    import torch

    ps = FRPSpace2D(0.0, 10.0, n, 0.0, 10.0, n, 3, 1, 1)
    u0 = ics.isentropic_vortex(ps, 1.4, strength=5.0)
    dx, dy = 10.0 / n, 10.0 / n
    iJ = np.linalg.inv(np.array([[dx / 2, shear * dx / 2], [0.0, dy / 2]]))
    s = FREuler2D(ps, 1.4, "rusanov", "periodic")
    s.set_geometry(iJ)
    if stagger is not None:
        s.set_stagger(stagger)
    st = torch.cuda.Stream()
    s.use_torch_stream(st)
    s.set_state(u0)
    dt = 0.02 * dx
    with torch.cuda.stream(st):
        s.step_graph(dt, steps, 1)
        s.set_state(u0)
        s.step(dt, 300)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        s.step_graph(dt, steps, 1)
        e1.record(st)
        torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / (3 * steps)
    dof = n * n * 16 * 4
    s.check_state()
    stag = s.stagger()
    s.close()
    return {"grid": f"parallelogram {n}x{n} (shear {shear})", "dof": dof, "stage_us": us,
            "dof_stage_per_s": dof / (us * 1e-6), "alg_GBs": (64.0 / 3.0) * dof / (us * 1e-6) / 1e9, "stagger": stag}


def ops_rows(nrow=1 << 20, n=4, nbatch=16, reps=20):
    """interp_face! and poly_derivative! (src/interpolate.jl:1-10, src/derivative.jl:1-10) on device
    batches f(row, q, b) of nrow x n x nbatch doubles (the layout of a Julia Array{Float64,3}; 4 x
    16 = the 64 planes of a p = 3 2D state contracted along one point index), `reps` calls each,
    HIP events on the launch stream.  Algorithmic bytes per call: 8 (n + nout) per (row, batch)."""
    import torch

    from . import ops

    r, _ = np.polynomial.legendre.leggauss(n)
    ll = np.array([np.prod([(-1.0 - r[m]) / (r[q] - r[m]) for m in range(n) if m != q]) for q in range(n)])
    lr = np.array([np.prod([(1.0 - r[m]) / (r[q] - r[m]) for m in range(n) if m != q]) for q in range(n)])
    pdm = np.asfortranarray(np.random.default_rng(0).standard_normal((n, n)))
    f = torch.rand((nbatch, n, nrow), dtype=torch.float64, device="cuda")
    ff = torch.empty((nbatch, 2, nrow), dtype=torch.float64, device="cuda")
    df = torch.empty_like(f)
    out = {}
    for name, call, nout, dst in (("interp_face", lambda: ops.interp_face(ff, f, ll, lr), 2, ff),
                                  ("poly_derivative", lambda: ops.poly_derivative(df, f, pdm), n, df)):
        for _ in range(3):
            call()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(20_000_000)  # hold the stream while the calls are queued
        e0.record()
        for _ in range(reps):
            call()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        nbytes = 8.0 * (n + nout) * nrow * nbatch
        out[name] = {"us_per_call": us, "alg_GBs": nbytes / (us * 1e-6) / 1e9, "bytes": nbytes}
    out["shape"] = [nrow, n, nbatch]
    return out


def shock_vortex_loop(mode="modal", steps=600):
    """`steps` iterations of the shock-vortex host loop (example/shock-vortex.jl:296-326) on the
    512 x 256, p = 3 case: positivity, SSP-RK3, sensor + filter, ghost refresh."""
    import torch

    c = sensor.SHOCK_VORTEX
    ic = {k: c[k] for k in ("shock_mach", "shock_x", "vortex_mach", "vortex_r", "vortex_x", "vortex_y")}
    ps, u0, s = sensor.shock_vortex_solver(c["nx"], c["ny"], **ic)
    if mode == "learned":
        sensor.attach(s, sensor.SensorMLP.load())
    dt, _ = ics.cfl_steps(u0, c["gamma"], ps.dx0, c["cfl"], 0.4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sensor.run_shock_vortex(s, dt, steps, mode, native=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    s.close()
    return {"mode": mode, "steps": steps, "ms_per_step": 1e3 * el / steps}


def vortex_grid(n=1024, steps=60, flux="rusanov"):
    """The isentropic-vortex stage of the headline case on an n x n periodic grid, p = 3 (same
    [0,10]^2 domain, CFL 0.05): fused SSP-RK3 stage launches replayed from a graph, timed with HIP
    events on the launch stream; algorithmic bytes 21.33 per DOF and stage."""
    import torch

    ps = FRPSpace2D(0.0, 10.0, n, 0.0, 10.0, n, 3, 1, 1)
    u0 = ics.isentropic_vortex(ps, 1.4, strength=5.0, center=(5.0, 5.0))
    dt, _ = ics.cfl_steps(u0, 1.4, ps.dx0, 0.05, 2.0)
    s = FREuler2D(ps, 1.4, flux, "periodic")
    st = torch.cuda.Stream()
    s.use_torch_stream(st)
    s.set_state(u0)
    del u0
    chunk = 10
    with torch.cuda.stream(st):
        s.step_graph(dt, chunk, 0)
        s.step(dt, 3)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        s.step_graph(dt, chunk, max(1, steps // chunk))
        e1.record(st)
        torch.cuda.synchronize()
    launches = 3 * chunk * max(1, steps // chunk)
    us = e0.elapsed_time(e1) * 1e3 / launches
    dof = n * n * 16 * 4
    s.check_state()
    s.close()
    return {"grid": f"{n}x{n}", "dof": dof, "stage_us": us, "dof_stage_per_s": dof / (us * 1e-6),
            "alg_GBs": 64.0 / 3.0 * dof / (us * 1e-6) / 1e9}


def main():
    ap = argparse.ArgumentParser(description="run one measured workload and print one JSON line")
    ap.add_argument("workload", choices=["ensemble", "bgk", "bgk64", "bgk1d", "bgk1d64", "bgk1d_script", "fr1d_small", "polar", "parallelogram",
                                         "ops", "shock_vortex", "vortex_grid"])
    ap.add_argument("--n", type=int, default=None, help="vortex_grid (1024) / polar, parallelogram (256): cells per side")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--mode", default="modal", choices=["modal", "learned", "off"])
    ap.add_argument("--stagger", type=int, default=None,
                    help="parallelogram: fixed staggered-start delay (e.g. under a profiler)")
    a = ap.parse_args()
    if a.workload == "ensemble":
        r = riemann_ensemble(a.steps or 10)
    elif a.workload == "bgk":
        r = bgk(a.steps or 3)
    elif a.workload == "bgk64":
        r = bgk(a.steps or 1, members=64)
    elif a.workload == "bgk1d":
        r = bgk1d(a.steps or 20)
    elif a.workload == "fr1d_small":
        r = fr1d_small(a.steps or 1000)
    elif a.workload == "bgk1d_script":
        r = bgk1d_script(a.steps or 1000)
    elif a.workload == "bgk1d64":
        r = bgk1d(a.steps or 4, members=64)
    elif a.workload == "polar":
        r = polar_stage(a.n or 256, a.steps or 20)
    elif a.workload == "ops":
        r = ops_rows()
    elif a.workload == "parallelogram":
        r = parallelogram_stage(a.n or 256, a.steps or 20, stagger=a.stagger)
    elif a.workload == "vortex_grid":
        r = vortex_grid(a.n or 1024, a.steps or 60)
    else:
        r = shock_vortex_loop(a.mode, a.steps or 600)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
