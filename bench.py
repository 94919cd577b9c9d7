"""Benchmark: active voxel-steps/s of the sparse-brick solver step on one MI355X.

Workload: four fire spheres at seeded positions in a 6 m x 4 m x 6 m region, curl-noise
turbulence (strength 1, scale 2), one sphere collider of radius 0.5 m, 0.025 m voxels,
8-cell bricks, plain FP32, 300 steps at 1/60 s. A step is one solver step of the whole
simulation. Warmup steps advance the scene untimed (by default to step 240); the timed
steps follow (steps 240-299). Inputs are synthetic and already resident on the device; the
host only hands each step's plain-data frame over.

Prints one JSON line with the metric, p50/p99 step latency, the dominant kernel's roofline
(algorithmic bytes over its event-timed duration) and a CPU baseline (the FP32 restatement
on one host core, over a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "threejs-fire-pro_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per active field cell of each stage (plain FP32, velocityDivisor 2):
# compulsory traffic of the arrays the stage reads and writes, once each.
S_F, S_V, D3 = 4, 4, 8
STAGE_BYTES_PER_FIELD_CELL = {
    "advect_scalars": 8 * S_F + 3 * S_V / D3,     # fields in, predictor out, velocity
    "evolve_fields": 13 * S_F + 3 * S_V / D3,     # predictor, fields -> fields, expansion
    "advect_velocity": 6 * S_V / D3,
    "correct_velocity": 9 * S_V / D3,
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=60)
    p.add_argument("--warmup", type=int, default=240)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    return p.parse_args()


def cpu_baseline(scene_fn, seconds: float) -> dict:
    """The FP32 restatement on one core over the first steps of the same scene, stopping
    after about `seconds` of work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleSimulation
    sc = scene_fn()
    sim = OracleSimulation.create(sc.grid(), 32)
    voxels, elapsed, steps = 0, 0.0, 0
    for frame, g in sc.frames():
        sim.apply_grid(g)
        t0 = time.perf_counter()
        sim.step(frame)
        elapsed += time.perf_counter() - t0
        voxels += sim.stats().active_voxels
        steps += 1
        if elapsed > seconds:
            break
    return {"value": voxels / elapsed, "unit": "active voxel-steps/s", "cores": 1, "kind": "port",
            "sample": f"first {steps} steps of the same scene, FP32 CPU restatement, single thread, "
                      f"{elapsed:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from firepro import abi, scenes
    from firepro.engine import FluidSimulation
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    scene_fn = lambda: scenes.four_fires(seed=args.seed, steps=args.warmup + args.steps)  # noqa: E731
    sc = scene_fn()
    frames = [(abi.PackedFrame(f), g) for f, g in sc.frames()]
    sim = FluidSimulation.create(sc.grid())

    def step(frame, g):
        sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
        sim.step(frame)

    for frame, g in frames[: args.warmup]:
        step(frame, g)
    sim.profile(True)
    sync()
    latencies, voxels = [], 0
    t0 = time.perf_counter()
    for frame, g in frames[args.warmup:]:
        s0 = time.perf_counter()
        step(frame, g)  # returns once the step's readback has landed
        latencies.append(time.perf_counter() - s0)
        voxels += sim.stats().active_voxels
    sync()
    wall = time.perf_counter() - t0
    stages = sim.profile_read()
    st = sim.stats()
    bricks = st.active_bricks
    value = voxels / wall
    # The dominant kernel: the stage with the largest share of the step.
    cells = voxels / args.steps  # mean active field cells per step
    best = max((k for k in stages if k in STAGE_BYTES_PER_FIELD_CELL), key=lambda k: stages[k])
    achieved = STAGE_BYTES_PER_FIELD_CELL[best] * cells / (stages[best] * 1e-3) / 1e9
    line = {
        "metric": "active voxel-steps/s", "value": value, "unit": "active voxel-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "four-fires plain FP32 single simulation: 4 fire spheres in 6x4x6 m, "
                               "turbulence 1/2, sphere collider r 0.5, voxel 0.025 m, brick 8, "
                               f"steps {args.warmup}-{args.warmup + args.steps - 1} of 300",
                   "active_bricks": bricks, "active_voxels": bricks * 512},
        "p50_step_ms": float(np.percentile(latencies, 50) * 1e3),
        "p99_step_ms": float(np.percentile(latencies, 99) * 1e3),
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"kernel": best, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    sim.dispose()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lambda: scenes.four_fires(seed=args.seed), args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
