"""Train the learned troubled-cell sensor on the GPU and store its parameters.

    python -m frhip.train_sensor [--runs 200] [--steps 3000]

Data: cell features and modal-sensor labels from seeded shock / vortex runs (frhip.sensor.
collect_training_data); training: libfrhip's MFMA forward / backward kernels with Adam
(frhip.sensor.train_sensor); output: frhip/data/sensor_mlp.npz with the parameter vector and the
training statistics.
"""
import argparse
import json
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=200)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--run-steps", type=int, default=240)
    ap.add_argument("--every", type=int, default=40)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    from frhip import sensor as S

    t0 = time.time()
    x, y = S.collect_training_data(args.runs, args.nx, args.ny, args.run_steps, args.every, args.seed)
    t1 = time.time()
    print(f"train_sensor: {x.shape[0]} samples from {args.runs} runs ({t1 - t0:.1f} s), "
          f"positive fraction {float(y.mean().item()):.3f}", flush=True)
    mlp, stats = S.train_sensor(x, y, nsteps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed)
    t2 = time.time()
    stats.update(getattr(S.collect_training_data, "last_stats", {}))
    stats.update(dict(runs=args.runs, grid=[args.nx, args.ny], run_steps=args.run_steps, every=args.every,
                      steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed, samples=int(x.shape[0]),
                      data_s=t1 - t0, train_s=t2 - t1))
    print("train_sensor: " + json.dumps(stats), flush=True)
    mlp.save(args.out or S.WEIGHTS, stats=json.dumps(stats))


if __name__ == "__main__":
    main()
