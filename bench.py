"""Benchmark of the 2D Euler FR hot path on one MI355X.

Workload: 2D Euler (gamma = 1.4) on the periodic square [0,10]^2, 256 x 256
uniform quads, p = 3 (16 Gauss points per element, 4,194,304 DOF), isentropic vortex (strength 5,
centred) on rho = 1, u = v = 1, p = 1, Rusanov flux, SSP-RK3 at CFL 0.05 on the wave speed
(dt = 0.05 dx / max(|u|+c, |v|+c) = 2/2999), FP64.
One step = one SSP-RK3 step = three fused residual+stage launches on the device-resident state,
replayed from a captured hipGraph of 100 steps; the default step count (2999) is the full run to
t = 2.

value = DOF * RK stages / s over the timed steps.  roofline: fused stage kernel, algorithmic bytes
per launch = 21.33 B/DOF (stage 1 reads u and writes u1: 16 B/DOF; stages 2-3 read u_s and u^n and
write: 24 B/DOF) x DOF, divided by the average launch duration measured with HIP events on the
stream the kernels run on.  cpu_baseline: the CPU oracle (C/OpenMP restatement of the reference
dudt!) timed on this host over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "fluxreconstruction.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "DOF·RK-stage/s 2D Euler p=3 (%HBM peak); learned vs modal sensor error & cost"
HBM_PEAK_GBS = 8000.0
NX = NY = 256
DEG = 3
GAMMA = 1.4
CFL = 0.05
GRAPH_CHUNK = 100  # SSP-RK3 steps per captured graph
SETTLE_STEPS = 600  # warm-up steps in all before timing (about 30 ms of stage launches)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fr2d_stage_256x256_p3_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (calibrated on a plain device copy), or None when the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def hold_stream(stream, torch, cycles=2_000_000):
    """Keeps the stream busy (a bounded spin kernel of `cycles` clock cycles, 2e6 about a
    millisecond) so that the launches of the timed region are queued behind it while the device
    works, and the region's start event is taken when the device, not the host's submission,
    reaches it (the blocking-kernel method of NVBench).  The spin kernel runs before the start
    event and is not timed.  Regions of eager calls from Python (about 15 us of host time each)
    take a longer hold, so that all of them are queued before the device reaches the first."""
    with torch.cuda.stream(stream):
        torch.cuda._sleep(cycles)


def cpu_baseline(u0, dt, budget_s=12.0):
    """Oracle SSP-RK3 steps on the host cores for about budget_s seconds."""
    import numpy as np

    from oracle import oracle as O

    orc = O.Fr2d(NX, NY, DEG, 10.0 / NX, 10.0 / NY, GAMMA, "rusanov", "periodic")
    u = np.array(u0, order="F", copy=True)
    t0 = time.perf_counter()
    orc.ssprk3(u, dt, 1)
    one = time.perf_counter() - t0
    n = max(1, int(budget_s / max(one, 1e-3)))
    t0 = time.perf_counter()
    orc.ssprk3(u, dt, n)
    el = time.perf_counter() - t0
    dof = NX * NY * (DEG + 1) ** 2 * 4
    return {
        "value": dof * 3 * n / el,
        "unit": "DOF·RK-stage/s",
        "cores": int(O.lib().or_num_threads()),
        "kind": "port",
        "sample": f"{n} SSP-RK3 steps ({3 * n} residual stages) of the same 256x256 p=3 vortex case, "
                  f"C/OpenMP oracle, {el:.1f} s",
    }


def bench_ensemble(steps=10):
    """512 seeded 1D Euler Riemann problems, batched vs looped (frhip.workloads)."""
    from frhip import workloads

    return workloads.riemann_ensemble(steps)


def bench_bgk(steps=3, ens_steps=1):
    """2D BGK, 128^2 periodic quads p = 2 x 24^2 Gauss-Hermite velocities, Kn = tau = 1e-2;
    plus the 64-member ensemble with tau log-uniform in [1e-3, 1e-1] (frhip.workloads)."""
    from frhip import workloads

    return {"workload": "2D BGK 128x128 p=2 x 24x24 Gauss-Hermite velocities, SSP-RK3", "unit": "values·RK-stage/s",
            "single": workloads.bgk(steps), "ensemble64": workloads.bgk(ens_steps, members=64)}


def bench_bgk1d():
    """The 1D BGK model of example/bgk_wave.jl as written, 8192 cells x 256 velocities p = 2, one
    member and 64 members, and the script's own 20-cell run to t = 1 eager and graph-replayed
    (frhip.workloads)."""
    from frhip import workloads

    return {"workload": "1D BGK (bgk_wave.jl) 8192 cells x 256 velocities, p=2, SSP-RK3", "unit": "values·RK-stage/s",
            "single": workloads.bgk1d(20), "ensemble64": workloads.bgk1d(4, members=64),
            "script_case_20x100_midpoint_1000_steps": workloads.bgk1d_script()}


def bench_fr1d_small():
    """The reference's small 1D runs (100 cells, p = 2, 1000 steps: advection RK4, Sod Midpoint)
    with a launch per stage and on the per-member path (frhip.workloads)."""
    from frhip import workloads

    return {"workload": "1D FR 100 cells p=2, 1000 steps", "unit": "ms per run", **workloads.fr1d_small()}


def bench_polar():
    """The p = 3 stage on a 256 x 256 polar mesh with point-varying metrics (frhip.workloads)."""
    from frhip import workloads

    return workloads.polar_stage(256, 20)


def bench_parallelogram():
    """The p = 3 stage on 256 x 256 uniform parallelogram cells (frhip.workloads)."""
    from frhip import workloads

    return workloads.parallelogram_stage(256, 20)


def bench_grids(sizes=(1024, 2048), steps=60):
    """The same vortex stage on larger periodic grids (frhip.workloads.vortex_grid)."""
    from frhip import workloads

    return [workloads.vortex_grid(n, steps) for n in sizes]


def bench_sensor():
    """Shock-vortex 512x256 p = 3 to t = 0.4 with the modal sensor, the learned sensor (trained
    weights frhip/data/sensor_mlp.npz) and no sensor; density L1 vs a 2x refined modal run."""
    import os as _os

    from frhip import sensor as S

    mlp = S.SensorMLP.load() if _os.path.exists(S.WEIGHTS) else None
    r = S.evaluate_shock_vortex(mlp)
    r.pop("reference_means", None)
    if "learned" in r and "modal" in r:
        r["L1_learned_over_modal"] = r["learned"]["L1_density"] / r["modal"]["L1_density"]
        r["cost_learned_over_modal"] = r["learned"]["ms_per_step"] / r["modal"]["ms_per_step"]
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the full run to t = 2)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--stagger", type=int, default=None,
                    help="fix the stage launches' staggered-start delay (units of 512 clocks) instead of the "
                         "library's device timing, e.g. under a profiler, whose dispatch handling skews that timing")
    ap.add_argument("--no-extras", action="store_true", help="only the headline 2D Euler measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 10 plain device copies of 64 MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import frhip
    from frhip import ics

    ps = frhip.FRPSpace2D(0.0, 10.0, NX, 0.0, 10.0, NY, DEG, 1, 1)
    u0 = ics.isentropic_vortex(ps, GAMMA, strength=5.0, center=(5.0, 5.0))
    dt, n_full = ics.cfl_steps(u0, GAMMA, ps.dx0, CFL, 2.0)
    if args.steps is None:
        args.steps = n_full
    dof = NX * NY * (DEG + 1) ** 2 * 4

    stream = torch.cuda.Stream()
    solver = frhip.FREuler2D(ps, GAMMA, "rusanov", "periodic")
    if args.stagger is not None:
        solver.set_stagger(args.stagger)
    solver.use_torch_stream(stream)
    solver.set_state(u0)
    # The timed steps replay a hipGraph of CHUNK SSP-RK3 steps that the context captured and
    # replayed once during the warm-up, so that no first-replay cost falls in the timed region
    # (the remainder runs as eager launches).  Warm-up: the capture and one replay (the context's
    # first stepping call also tunes its staggered start), the W eager steps, then eager steps up
    # to SETTLE_STEPS in all, about 30 ms of sustained launches, so that the clocks have settled
    # as they have in any longer run.  The state is reset on the device before timing.
    chunk = max(1, min(GRAPH_CHUNK, args.steps))
    u0d = torch.from_numpy(np.asfortranarray(u0).ravel(order="F").copy()).cuda()
    with torch.cuda.stream(stream):
        solver.step_graph(dt, chunk, 1)  # capture and one warm replay
        solver.step(dt, args.warmup)
        solver.step(dt, max(0, SETTLE_STEPS - chunk - args.warmup))
        torch.cuda.synchronize()
        solver.set_state(u0d)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        hold_stream(stream, torch)
        t0 = time.perf_counter()
        ev0.record(stream)
        if args.steps // chunk:
            solver.step_graph(dt, chunk, args.steps // chunk)
        if args.steps % chunk:
            solver.step(dt, args.steps % chunk)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    uf = solver.get_state()
    solver.check_state()
    launches = 3 * args.steps
    per_launch_s = ms * 1e-3 / launches
    alg_bytes = (16.0 + 24.0 + 24.0) / 3.0 * dof
    achieved = alg_bytes / per_launch_s / 1e9
    value = dof * launches / (ms * 1e-3)

    # residual-only drop-in dudt! on device arrays (16 B/DOF)
    ud = torch.from_numpy(uf.ravel(order="F").copy()).cuda()
    dud = torch.empty_like(ud)
    with torch.cuda.stream(stream):
        for _ in range(5):
            solver.dudt(dud, ud)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        nres = 200
        hold_stream(stream, torch, 20_000_000)
        e0.record(stream)
        for _ in range(nres):
            solver.dudt(dud, ud)
        e1.record(stream)
        torch.cuda.synchronize()
    res_s = e0.elapsed_time(e1) * 1e-3 / nres

    # practical streaming rate of this working set on this GPU: c = a + b over three state-sized
    # arrays (24 B per value, torch's elementwise kernel), the traffic pattern of stages 2-3
    sa, sb = torch.rand_like(ud), torch.rand_like(ud)
    sc = torch.empty_like(ud)
    with torch.cuda.stream(stream):
        for _ in range(5):
            torch.add(sa, sb, out=sc)
        torch.cuda.synchronize()
        hold_stream(stream, torch, 20_000_000)
        e0.record(stream)
        for _ in range(nres):
            torch.add(sa, sb, out=sc)
        e1.record(stream)
        torch.cuda.synchronize()
    stream_gbs = 24.0 * ud.numel() / (e0.elapsed_time(e1) * 1e-3 / nres) / 1e9
    del sa, sb, sc

    if args.copy_calibration:
        src = torch.ones(8 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None if args.no_cpu else cpu_baseline(u0, dt)
    mass0 = float(np.einsum("ijkl,kl->", u0[1:-1, 1:-1, :, :, 0], ps.wp))
    mass1 = float(np.einsum("ijkl,kl->", uf[1:-1, 1:-1, :, :, 0], ps.wp))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF·RK-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "2D Euler isentropic vortex, [0,10]^2 periodic, 256x256 quads, p=3, Rusanov, "
                        "SSP-RK3 CFL 0.05 (dt = 0.05 dx / max(|u|+c,|v|+c)), fused residual+stage kernel, "
                        "device-resident state",
            "dof": dof,
            "dt": dt,
            "t_end": dt * args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "kernel": "fr2d_kernel<4,RUSANOV,STAGE>",
            "stagger": solver.stagger(),
            "avg_launch_us": per_launch_s * 1e6,
            "alg_bytes_per_launch": alg_bytes,
            "stream_add_GBs": stream_gbs,
            "frac_of_stream_add": achieved / stream_gbs,
        },
        "cpu_baseline": cpu,
        "residual_only": {
            "kernel": "fr2d_kernel<4,RUSANOV,RESIDUAL>",
            "avg_launch_us": res_s * 1e6,
            "dof_per_s": dof / res_s,
            "hbm_gbs_alg": 16.0 * dof / res_s / 1e9,
        },
        "wall_s_timed": wall,
        "mass_drift_rel": abs(mass1 - mass0) / abs(mass0),
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    if not args.no_extras:
        for key, fn in (("vortex_larger_grids", bench_grids), ("ensemble_riemann512", bench_ensemble),
                        ("bgk_128sq_p2_v24", bench_bgk), ("bgk1d_8192_v256", bench_bgk1d),
                        ("fr1d_small_100cells", bench_fr1d_small),
                        ("polar_256sq_p3", bench_polar), ("parallelogram_256sq_p3", bench_parallelogram),
                        ("sensor_shock_vortex", bench_sensor)):
            try:
                out[key] = fn()
            except Exception as exc:  # an extra leg must not hide the headline line
                out[key] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
