"""Benchmark: active voxel-steps/s of the sparse-brick solver step on one MI355X.

Workload: four fire spheres at seeded positions in a 6 m x 4 m x 6 m region, curl-noise
turbulence (strength 1, scale 2), one sphere collider of radius 0.5 m, 0.025 m voxels,
8-cell bricks, plain FP32, 300 steps at 1/60 s. A step is one solver step of the whole
simulation. The scene is first advanced untimed to step 240 (--state-step), whatever the
warmup, so the timed steps always see the late, large state (about 15.6 K active bricks);
then --warmup steps run untimed and --steps steps are timed (by default steps 245-299).
Inputs are synthetic and already resident on the device; the host only hands each step's
plain-data frame over.

Prints one JSON line with the metric, p50/p99 step latency (device time per step from HIP
events on the simulation's stream), the dominant kernel's roofline (its algorithmic bytes
over its event-timed duration), the metric's comparison (the serving stream batched with
BF16 multigrid levels against one simulation per launch in FP32, as it arrives and with
every request present at once) and a CPU baseline (the
FP32 restatement on all host threads and on one, over a bounded sample of the same
workload).
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
# Algorithmic bytes per active field cell of each stage kernel (plain FP32, velocityDivisor
# 2): the compulsory traffic of the arrays the reference's stage reads and writes, once each.
S_F, S_V, D3 = 4, 4, 8
STAGE_BYTES_PER_FIELD_CELL = {
    "advect_scalars": 8 * S_F + 3 * S_V / D3,    # fields -> predictor, velocity under the cell
    "evolve_fields": 13 * S_F + 3 * S_V / D3,    # predictor, fields, velocity -> fields, expansion
    "advect_velocity": 6 * S_V / D3,             # u -> advected u (per velocity cell, over d^3)
    "correct_velocity": 9 * S_V / D3,            # u, advected u -> corrected u
}
# Bytes beyond those the kernels move by design: the backtrace planes (three float4 per
# field cell, four per velocity cell) that the advection writes and the correction reads.
STAGE_OVERHEAD_PER_FIELD_CELL = {
    "advect_scalars": 12 * S_F, "evolve_fields": 12 * S_F,
    "advect_velocity": 16 * S_V / D3, "correct_velocity": 16 * S_V / D3,
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=55)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--brick", type=int, default=8, choices=(8, 16),
                   help="field brick edge of the four-fires workload (16 with --max-voxels 64e6: the "
                        "large-brick roofline variant whose working set exceeds the Infinity Cache)")
    p.add_argument("--max-voxels", type=float, default=16e6)
    p.add_argument("--state-step", type=int, default=240,
                   help="steps advanced untimed before the warmup: the state the timed steps see")
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="seconds of the CPU baseline on all host threads (one thread gets two thirds)")
    p.add_argument("--render-steps", type=int, default=6,
                   help="steps after the timed ones, each followed by a 1920x1080 render")
    p.add_argument("--serving", action="store_true",
                   help="the serving workload instead: continuous batching of many simulations against "
                        "one simulation per launch")
    p.add_argument("--serving-count", type=int, default=256)
    p.add_argument("--serving-extra", action="store_true",
                   help="with --serving: also every request at once and the single V-cycle policy")
    p.add_argument("--serving-no-graphs", action="store_true",
                   help="with --serving: launch the batched kernels one by one (for kernel tracing)")
    p.add_argument("--serving-rate", type=float, default=20.0)
    p.add_argument("--explosions", action="store_true", help="the brick-churn workload instead")
    p.add_argument("--presets", type=int, default=0, metavar="STEPS",
                   help="headless runs of the shipped presets (tests/golden/presets), STEPS each")
    p.add_argument("--mix", type=int, default=256,
                   help="requests of the serving stream (Poisson at --serving-rate) for the batched "
                        "mixed-precision comparison in the default line (0 skips it)")
    p.add_argument("--no-capacity", action="store_true",
                   help="skip the capacity leg of the comparison (the same requests all present at once)")
    p.add_argument("--profile-summary", metavar="DIR", default=None,
                   help="only summarize the rocprofv3 CSV output under DIR (kernel trace and/or "
                        "counters) over the last --tail-steps solver steps, and exit")
    p.add_argument("--tail-steps", type=int, default=20)
    p.add_argument("--traffic-summary", metavar="PREFIX", default=None,
                   help="only turn the FETCH_SIZE and WRITE_SIZE counter runs PREFIX_FETCH_SIZE and "
                        "PREFIX_WRITE_SIZE into HBM bytes per launch (--copy-dirs, --cells, --out), and exit")
    p.add_argument("--copy-dirs", default="gpurun_out/copy", metavar="PREFIX",
                   help="the copy probe's counter runs PREFIX_FETCH_SIZE and PREFIX_WRITE_SIZE")
    p.add_argument("--cells", type=float, default=0.0, help="active field cells per launch")
    p.add_argument("--out", default=None, help="JSON output of --traffic-summary")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="only copy this many MiB device to device 5 times (counter calibration)")
    return p.parse_args()


def profile_summary(root: str, tail: int) -> None:
    """Per-kernel averages from rocprofv3's CSV output under `root`, over the dispatches of
    the last `tail` solver steps (from the start of the tail-th last k_build_bricks, each
    step's first kernel, to the end): calls, mean and total duration from the kernel trace,
    and each counter's mean value per dispatch from a counter collection. Kernels of a run
    without solver steps (the copy probe) are summarized whole."""
    from collections import defaultdict

    def rows(pattern):
        return _profile_rows(root, pattern)

    def window(rs):
        # One start per dispatch (a counter collection has a row per counter).
        starts = sorted({(int(r["Start_Timestamp"]), r.get("Dispatch_Id", "")) for r in rs
                         if "k_build_bricks" in r["Kernel_Name"]})
        starts = [t for t, _ in starts]
        return starts[-tail] if len(starts) >= tail else (starts[0] if starts else 0)

    def short(name):
        return name.split("(")[0].replace("void ", "").replace("fp::", "")

    trace = rows("*kernel_trace.csv")
    if trace:
        t0 = window(trace)
        agg = defaultdict(list)
        vgpr = {}
        for r in trace:
            if int(r["Start_Timestamp"]) >= t0:
                agg[short(r["Kernel_Name"])].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
                vgpr[short(r["Kernel_Name"])] = r.get("Arch_VGPR_Count", r.get("VGPR_Count", ""))
        total = sum(sum(v) for v in agg.values())
        print(f"kernel trace, last {tail} steps: {sum(len(v) for v in agg.values())} dispatches, "
              f"{total / 1e6:.3f} ms of kernel time")
        print(f"{'kernel':48s} {'calls':>6s} {'avg us':>9s} {'total us':>10s} {'share':>6s} {'vgpr':>5s}")
        for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
            print(f"{k[:48]:48s} {len(v):6d} {sum(v) / len(v) / 1e3:9.2f} {sum(v) / 1e3:10.1f} "
                  f"{100 * sum(v) / total:5.1f}% {vgpr[k]:>5s}")
    agg = counter_means(root, tail)
    if agg:
        print(f"counters, last {tail} steps, mean per dispatch:")
        names = sorted({c for v in agg.values() for c in v if c != "calls"})
        print(f"{'kernel':48s} {'calls':>6s} " + " ".join(f"{c:>16s}" for c in names))
        for k, v in sorted(agg.items()):
            print(f"{k[:48]:48s} {v['calls']:6d} " + " ".join(f"{v.get(c, float('nan')):16.1f}" for c in names))


def _profile_rows(root: str, pattern: str):
    import csv
    import glob
    out = []
    for path in sorted(glob.glob(os.path.join(root, "**", pattern), recursive=True)):
        with open(path, newline="") as f:
            out.extend(csv.DictReader(f))
    return out


def counter_means(root: str, tail: int) -> dict:
    """{kernel: {counter: mean per dispatch, "calls": dispatches}} of a rocprofv3 counter
    collection under `root`, over the last `tail` solver steps (the whole run without any)."""
    from collections import defaultdict
    rs = _profile_rows(root, "*counter_collection.csv")
    starts = sorted({(int(r["Start_Timestamp"]), r.get("Dispatch_Id", "")) for r in rs
                     if "k_build_bricks" in r["Kernel_Name"]})
    t0 = starts[-tail][0] if len(starts) >= tail else (starts[0][0] if starts else 0)
    agg = defaultdict(lambda: defaultdict(list))
    for r in rs:
        if int(r["Start_Timestamp"]) >= t0:
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("fp::", "")
            agg[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {**{c: sum(x) / len(x) for c, x in v.items()}, "calls": max(len(x) for x in v.values())}
            for k, v in agg.items()}


def traffic_summary(prefix: str, copy: str, cells: float, tail: int, out: str) -> None:
    """HBM bytes per launch of each profiled kernel from two counter runs (prefix_FETCH_SIZE,
    prefix_WRITE_SIZE), each scaled by what the same counter reads for the plain copy probe
    (copy_FETCH_SIZE, copy_WRITE_SIZE: 256 MiB read and written per dispatch); per active
    field cell with `cells` cells a launch. Writes the table to stdout and JSON to `out`."""
    fetch = counter_means(prefix + "_FETCH_SIZE", tail)
    write = counter_means(prefix + "_WRITE_SIZE", tail)
    kib = 256 * 1024
    cf = counter_means(copy + "_FETCH_SIZE", tail)
    cw = counter_means(copy + "_WRITE_SIZE", tail)
    fscale = kib / next(v["FETCH_SIZE"] for k, v in cf.items() if "copyBuffer" in k)
    wscale = kib / next(v["WRITE_SIZE"] for k, v in cw.items() if "copyBuffer" in k)
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        if not k.startswith("k_"):
            continue
        f, w = fetch[k]["FETCH_SIZE"], write[k]["WRITE_SIZE"]
        b = (f * fscale + w * wscale) * 1024
        kernels[k[2:].split("<")[0]] = {
            "fetch_kib": round(f, 1), "write_kib": round(w, 1), "bytes_per_launch": b, "bytes_per_cell": b / cells}
    doc = {"cells": cells, "fetch_scale": fscale, "write_scale": wscale, "kernels": kernels}
    with open(out, "w") as fo:
        json.dump(doc, fo, indent=1)
    print(f"calibration: FETCH_SIZE x {fscale:.5f}, WRITE_SIZE x {wscale:.5f} (256 MiB copy probe)")
    print(f"{'kernel':34s} {'FETCH KiB':>11s} {'WRITE KiB':>11s} {'MB/launch':>10s} {'B/cell':>8s}")
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
        print(f"{k[:34]:34s} {v['fetch_kib']:11.1f} {v['write_kib']:11.1f} "
              f"{v['bytes_per_launch'] / 1e6:10.1f} {v['bytes_per_cell']:8.2f}")


def copy_probe(mib: int) -> None:
    """A plain device copy of known size: the calibration for HBM byte counters."""
    from firepro import abi
    n = mib << 20
    a, b = abi.DeviceBuffer(n), abi.DeviceBuffer(n)
    for _ in range(5):
        b.copy_from(a, n)
    b.synchronize()
    a.close()
    b.close()
    print(json.dumps({"copy_probe_bytes": n, "copies": 5}))


def host_threads() -> int:
    """Threads the host may use: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when it is set (a shared box gives each GPU its share)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_baseline(scene_fn, seconds: float) -> dict:
    """The FP32 restatement over the first steps of the same scene: on all host threads
    for about `seconds` of work (the reported value), then on one thread for two thirds of
    that. Its per-brick passes run on OpenMP threads with results independent of the count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from oracle import OracleSimulation

    def run(threads, budget):
        oracle.set_threads(threads)
        sc = scene_fn()
        sim = OracleSimulation.create(sc.grid(), 32)
        voxels, elapsed, steps, bricks = 0, 0.0, 0, 0
        for frame, g in sc.frames():
            sim.apply_grid(g)
            t0 = time.perf_counter()
            sim.step(frame)
            elapsed += time.perf_counter() - t0
            st = sim.stats()
            voxels += st.active_voxels
            bricks = st.active_bricks
            steps += 1
            if elapsed > budget:
                break
        return voxels / elapsed, steps, elapsed, bricks

    threads = host_threads()
    many, steps_m, el_m, bricks_m = run(threads, seconds)
    one, steps_1, el_1, bricks_1 = run(1, seconds * 2 / 3)
    oracle.set_threads(threads)
    return {"value": many, "unit": "active voxel-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "single_core_value": one,
            "sample": f"steps 0-{steps_m - 1} of the same scene (active bricks rising to {bricks_m}), FP32 CPU "
                      f"restatement, OpenMP over bricks on {threads} threads, {el_m:.1f} s; on one thread steps "
                      f"0-{steps_1 - 1} (to {bricks_1} bricks), {el_1:.1f} s"}


def serving_line(args) -> dict:
    """The serving workload: independent simulations (1-4 fires, voxel 0.025-0.1 m, brick 8,
    16-8192 bricks, 30-300 steps) arriving as a Poisson stream, admitted by a 200 GB HBM
    budget, each step's pressure solved by defect correction to a relative residual
    |b - Ax| / |b| of 1e-4 (at most 20 V-cycles).

    Headline: the stream served with continuous batching and multigrid levels in BF16,
    against the same stream served one simulation per launch with levels in FP32. Beside
    it: every request present at once (the servers' capacity, both policies), and both
    servers at the reference's single V-cycle."""
    from firepro import serving
    count = args.serving_count
    solve = (1e-4, 20)
    mixed, plain = ("bf16",) + solve, ("fp32",) + solve

    def say(line):
        print(line, file=sys.stderr, flush=True)

    def run(rate, batched, pressure):
        rep = serving.serve(serving.requests(count, rate), batched=batched, pressure=pressure, progress=say,
                            graphs=not args.serving_no_graphs)
        out = rep.summary()
        say(f"served: {out['mode']} {out['pressure']} rate {rate}: {out['voxel_steps_per_s']:.3e} voxel-steps/s")
        return out

    warm = serving.requests(count, 1e9)
    for r in warm:
        r.steps = 1
    serving.serve(warm, batched=True, pressure=mixed, graphs=not args.serving_no_graphs)  # blocks, streams
    stream_b, stream_s = run(args.serving_rate, True, mixed), run(args.serving_rate, False, plain)
    ratio = lambda a, b: a["voxel_steps_per_s"] / b["voxel_steps_per_s"] if b["voxel_steps_per_s"] else None  # noqa: E731
    line = {
        "metric": "serving active voxel-steps/s", "value": stream_b["voxel_steps_per_s"],
        "unit": "active voxel-steps/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "storage": "bf16 multigrid levels", "data": "synthetic",
        "config": {"workload": f"{count} fire simulations, Poisson arrivals at {args.serving_rate}/s, 1-4 fires, "
                               "voxel 0.025-0.1 m, brick 8, 16-8192 bricks, 30-300 steps, 200 GB HBM budget, "
                               "pressure to relative residual 1e-4 (at most 20 V-cycles)"},
        "speedup": ratio(stream_b, stream_s),
        "p50_step_ms": stream_b["p50_step_ms"], "p99_step_ms": stream_b["p99_step_ms"],
        "batched_bf16": stream_b, "one_sim_per_launch_fp32": stream_s,
        "graphs": not args.serving_no_graphs,
    }
    if args.serving_extra:
        burst_b, burst_s = run(1e9, True, mixed), run(1e9, False, plain)
        vcyc_b, vcyc_s = run(1e9, True, None), run(1e9, False, None)
        line["all_at_once"] = {"batched_bf16": burst_b, "one_sim_per_launch_fp32": burst_s,
                               "speedup": ratio(burst_b, burst_s)}
        line["single_vcycle_all_at_once"] = {"batched": vcyc_b, "one_sim_per_launch": vcyc_s,
                                             "speedup": ratio(vcyc_b, vcyc_s)}
    return line


def explosions_line(args) -> dict:
    """The brick-churn workload: 8 radial bursts in a 20 m cube (voxel 0.05 m, brick 8) at
    seeded steps 0-150, 400 steps, active bricks swinging with the bursts. Whole-run
    throughput and the allocation stage's share."""
    from firepro import abi, scenes
    from firepro.engine import FluidSimulation
    sc = scenes.explosions(seed=args.seed, steps=400)
    frames = [(abi.PackedFrame(f), g) for f, g in sc.frames()]
    sim = FluidSimulation.create(sc.grid())
    sim.profile(True)
    bricks, voxels = [], 0
    t0 = time.perf_counter()
    for frame, g in frames:
        sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
        sim.step(frame)
        st = sim.stats()
        bricks.append(st.active_bricks)
        voxels += st.active_voxels
    wall = time.perf_counter() - t0
    stages = sim.profile_read()
    sim.dispose()
    return {"metric": "active voxel-steps/s", "value": voxels / wall, "unit": "active voxel-steps/s",
            "n_gpus": 1, "steps": len(frames), "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "brick churn: 8 radial bursts in a 20 m cube, voxel 0.05 m, brick 8, 400 steps"},
            "ms_per_step": 1e3 * wall / len(frames), "active_bricks_min": min(bricks[150:]),
            "active_bricks_max": max(bricks), "active_bricks_last": bricks[-1],
            "stage_ms": {k: round(v, 4) for k, v in stages.items()}}


def presets_lines(steps: int):
    """Each shipped preset with sphere emitters, played headless (bursts fired at the start)
    for `steps` fixed steps: whole-run throughput and the step time."""
    import glob
    from firepro import abi, document
    from firepro.engine import FluidSimulation
    here = os.path.join(ROOT, "tests", "golden", "presets")
    for path in sorted(glob.glob(os.path.join(here, "*.json"))):
        with open(path) as f:
            doc = json.load(f)
        try:
            sc = document.preset_scene(doc, steps)
        except abi.FireProError as e:
            yield {"preset": doc["id"], "skipped": str(e)}
            continue
        frames = [(abi.PackedFrame(fr), g) for fr, g in sc.frames()]
        sim = FluidSimulation.create(sc.grid())
        voxels, peak, step_ms = 0, 0, []
        t0 = time.perf_counter()
        for frame, g in frames:
            s0 = time.perf_counter()
            sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
            sim.step(frame)
            step_ms.append(1e3 * (time.perf_counter() - s0))
            st = sim.stats()
            voxels += st.active_voxels
            peak = max(peak, st.active_bricks)
        wall = time.perf_counter() - t0
        sim.dispose()
        yield {"preset": doc["id"], "metric": "active voxel-steps/s", "value": voxels / wall,
               "steps": steps, "ms_per_step": 1e3 * wall / steps,
               "p50_step_ms": float(np.percentile(step_ms, 50)), "p99_step_ms": float(np.percentile(step_ms, 99)),
               "peak_active_bricks": peak, "grid": sc.grid()}


def mix_comparison(count: int, rate: float, capacity: bool = True) -> dict:
    """The metric's comparison on the serving workload: `count` requests arriving as a
    Poisson stream at `rate` per second (1-4 fires, voxel 0.025-0.1 m, brick 8, 16-8192
    bricks, 30-300 steps, admitted by a 200 GB HBM budget), each step's pressure solved to a
    relative residual |b - Ax| / |b| of 1e-4 (at most 20 V-cycles): served with continuous
    batching and BF16 multigrid levels, and one simulation per launch with FP32 levels. Step
    latencies are to device completion on both sides."""
    from firepro import serving
    solve = (1e-4, 20)
    warm = serving.requests(count, 1e9)
    for r in warm:
        r.steps = 1
    serving.serve(warm, batched=True, pressure=("bf16",) + solve)  # device blocks, streams, graphs
    say = lambda line: print(line, file=sys.stderr, flush=True)  # noqa: E731
    b = serving.serve(serving.requests(count, rate), batched=True, pressure=("bf16",) + solve,
                      progress=say).summary()
    s = serving.serve(serving.requests(count, rate), batched=False, pressure=("fp32",) + solve,
                      progress=say).summary()
    keep = ("voxel_steps_per_s", "wall_s", "rounds", "sim_steps", "requests", "p50_step_ms", "p99_step_ms",
            "p50_request_s", "p99_request_s", "peak_resident", "mean_cycles", "max_relative_residual",
            "steps_within_tolerance")
    out = {"sample": f"{count} serving requests, Poisson arrivals at {rate}/s (1-4 fires, voxel 0.025-0.1 m, "
                     "brick 8, 16-8192 bricks, 30-300 steps, 200 GB HBM budget, pressure to relative residual "
                     "1e-4 in at most 20 V-cycles)",
           "batched_bf16": {k: b[k] for k in keep}, "one_sim_per_launch_fp32": {k: s[k] for k in keep},
           "speedup": b["voxel_steps_per_s"] / s["voxel_steps_per_s"] if s["voxel_steps_per_s"] else None}
    if capacity:
        # The stream at 20/s is the batched server's offered load, which it keeps up with; the
        # same requests all present at once measure each server's capacity instead.
        bc = serving.serve(serving.requests(count, 1e9), batched=True, pressure=("bf16",) + solve).summary()
        sc = serving.serve(serving.requests(count, 1e9), batched=False, pressure=("fp32",) + solve,
                           progress=say).summary()
        out["capacity"] = {
            "sample": "the same requests all present at once",
            "batched_bf16": {k: bc[k] for k in keep}, "one_sim_per_launch_fp32": {k: sc[k] for k in keep},
            "speedup": bc["voxel_steps_per_s"] / sc["voxel_steps_per_s"] if sc["voxel_steps_per_s"] else None}
    return out


def render_phase(sim, frames, step) -> dict:
    """After the timed steps: a step, then the 1920x1080 ray-march with self-shadowing,
    several times, at half resolution (the reference's default: the march into depth
    slices at 960x540, then the full-resolution resume from them); the render stages'
    event-timed averages (preparation every step, the whole light every second step, the
    slices and march every frame), and the scene lights' totals with an anchor at each fire."""
    from firepro import abi
    w, h = 1920, 1080
    cam = abi.camera((0.0, 3.0, 9.0), (0.0, 1.5, 0.0), 50, w, h)
    rs = abi.render_settings()
    buf = abi.DeviceBuffer(w * h * 16)
    try:
        step(*frames[0])
        sim.render_device(cam, rs, buf.ptr.value)  # first build of the render pools
        buf.synchronize()
        sim.profile(True)
        walls = []
        counts = {}
        for frame, g in frames[1:]:
            step(frame, g)
            t0 = time.perf_counter()
            sim.render_device(cam, rs, buf.ptr.value)
            buf.synchronize()
            walls.append(time.perf_counter() - t0)
            for k, v in sim.render_stats().items():
                counts[k] = counts.get(k, 0) + v
            anchors = [tuple(frame.sources[i].position) for i in range(frame.frame.source_count)]
            sim.scene_lights(anchors, rs)
        stages = sim.profile_read()
    finally:
        buf.close()
    out = {"width": w, "height": h, "renders": len(walls),
           "wall_ms_mean": 1e3 * float(np.mean(walls)) if walls else None}
    for k in ("render_prepare", "render_light", "render_slices", "render_march", "scene_lights"):
        if k in stages:
            out[k + "_ms"] = round(stages[k], 4)
    # Samples per second of each march (the ray-march is bound by its gathers, not by HBM):
    # positions looked up in the occupancy and samples reconstructed and shaded, per render,
    # over the pass's event-timed mean.
    n = max(1, len(walls))
    for key, stage in (("slice", "render_slices"), ("march", "render_march")):
        visits, shades = counts.get(key + "_visits", 0) / n, counts.get(key + "_shades", 0) / n
        out[key + "_samples"] = {"visited": int(visits), "shaded": int(shades)}
        if stages.get(stage):
            out[key + "_samples"]["visited_per_s"] = visits / (stages[stage] * 1e-3)
            out[key + "_samples"]["shaded_per_s"] = shades / (stages[stage] * 1e-3)
    solver = {k: round(v, 4) for k, v in stages.items() if not k.startswith("render_") and k != "scene_lights"}
    return out, solver


def measured_traffic(kernel: str, cells: float, brick: int = 8):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of this
    same command (FETCH_SIZE scaled by the plain-copy calibration, plus WRITE_SIZE), per
    active cell, times this run's cells; None without a summary for the kernel."""
    path = os.path.join(ROOT, "profiles", f"four_fires_b{brick}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    per = t.get("kernels", {}).get(kernel)
    return per["bytes_per_cell"] * cells if per else None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.traffic_summary:
        traffic_summary(args.traffic_summary, args.copy_dirs, args.cells, args.tail_steps, args.out)
        return
    if args.profile_summary:
        profile_summary(args.profile_summary, args.tail_steps)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.serving:
        print(json.dumps(serving_line(args)))
        return
    if args.explosions:
        print(json.dumps(explosions_line(args)))
        return
    if args.presets:
        for line in presets_lines(args.presets):
            print(json.dumps(line), flush=True)
        return
    from firepro import abi, scenes
    from firepro.engine import FluidSimulation
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    first = args.state_step + args.warmup  # the first timed step
    total = first + args.steps + args.render_steps
    scene_fn = lambda: scenes.four_fires(seed=args.seed, steps=total, brick=args.brick,  # noqa: E731
                                         max_voxels=int(args.max_voxels))
    sc = scene_fn()
    frames = [(abi.PackedFrame(f), g) for f, g in sc.frames()]
    sim = FluidSimulation.create(sc.grid())

    def step(frame, g):
        sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
        sim.step(frame)

    for frame, g in frames[:first]:  # the untimed advance to the timed state, then the warmup
        step(frame, g)
    sim.profile(2)  # events around the dominant kernel only, inside the timed region
    timer = abi.StreamTimer(sim.stream())
    sync()
    voxels, bricks_seen = 0, []
    t0 = time.perf_counter()
    for frame, g in frames[first:first + args.steps]:
        timer.mark()
        step(frame, g)  # returns once the step's readback has landed; its tail stays queued
        timer.mark()
        st = sim.stats()
        voxels += st.active_voxels
        bricks_seen.append(st.active_bricks)
    sync()
    wall = time.perf_counter() - t0
    latencies = timer.elapsed()  # per step: device time to the completion of its last kernel
    timer.close()
    stages = sim.profile_read()
    value = voxels / wall
    # The dominant kernel: the stage with the largest event-timed share of the step.
    cells = voxels / args.steps  # mean active field cells per step
    best = max((k for k in stages if k in STAGE_BYTES_PER_FIELD_CELL), key=lambda k: stages[k])
    algorithmic = STAGE_BYTES_PER_FIELD_CELL[best] * cells
    achieved = algorithmic / (stages[best] * 1e-3) / 1e9
    traffic = measured_traffic(best, cells, args.brick)
    line = {
        "metric": "active voxel-steps/s", "value": value, "unit": "active voxel-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "four-fires plain FP32 single simulation: 4 fire spheres in 6x4x6 m, "
                               f"turbulence 1/2, sphere collider r 0.5, voxel 0.025 m, brick {args.brick}, "
                               f"maxVoxels {args.max_voxels:.3g}, "
                               f"timed steps {first}-{first + args.steps - 1} (after an untimed advance to "
                               f"step {args.state_step} and {args.warmup} warmup steps)",
                   "active_bricks": bricks_seen[-1], "active_bricks_min": min(bricks_seen),
                   "active_voxels": bricks_seen[-1] * args.brick ** 3},
        "p50_step_ms": float(np.percentile(latencies, 50)),
        "p99_step_ms": float(np.percentile(latencies, 99)),
        "step_latency": "device time per step on the simulation's stream, HIP events before and after each "
                        "step call (enqueue to completion of its last kernel)",
        "timed_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": {"kernel": best, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": algorithmic,
                     "algorithmic_bytes_per_cell": STAGE_BYTES_PER_FIELD_CELL[best],
                     "overhead_bytes_per_cell": STAGE_OVERHEAD_PER_FIELD_CELL[best],
                     "cells": cells, "kernel_ms": stages[best]},
    }
    if args.render_steps > 0:
        line["render"], line["stage_ms"] = render_phase(sim, frames[first + args.steps:], step)
        # Every stage kernel's fraction of peak over its algorithmic bytes, from the event-timed
        # steps right after the timed ones (near the same state).
        line["stage_roofline"] = {
            k: round(STAGE_BYTES_PER_FIELD_CELL[k] * cells / (line["stage_ms"][k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            for k in STAGE_BYTES_PER_FIELD_CELL if line["stage_ms"].get(k)}
    sim.dispose()
    if args.mix > 0:
        line["mixed_precision_batched"] = mix_comparison(args.mix, args.serving_rate, not args.no_capacity)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lambda: scenes.four_fires(seed=args.seed), args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
