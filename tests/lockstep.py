"""Runs one scene through the HIP engine and the CPU restatement side by side."""
from __future__ import annotations

import numpy as np

TABLE_KEYS = ("pages", "entries", "active", "zeroed", "free_slots")


def make_engines(scene, bits=32):
    from firepro.engine import FluidSimulation
    from oracle import OracleSimulation
    return FluidSimulation.create(scene.grid()), OracleSimulation.create(scene.grid(), bits)


def apply(gpu, cpu, g):
    gpu.cutoff, gpu.voxel_budget, gpu.ground = g["cutoff"], g["maxVoxels"], g["ground"]
    cpu.apply_grid(g)


def table_mismatches(ta, tb):
    out = [k for k in TABLE_KEYS if not np.array_equal(ta[k], tb[k])]
    act = ta["active"]
    if np.array_equal(ta["active"], tb["active"]) and not np.array_equal(ta["records"][act], tb["records"][act]):
        out.append("records")
    for k in ("free_count", "failures", "pool_slots"):
        if ta[k] != tb[k]:
            out.append(f"{k}: {ta[k]} != {tb[k]}")
    return out


def state_mismatches(a, b, keys=("fields", "velocity", "pressure", "expansion")):
    if not np.array_equal(a["ids"], b["ids"]):
        return ["active ids"]
    out = []
    for k in keys:
        d = np.abs(a[k].astype(np.float64) - b[k].astype(np.float64))
        if d.size and d.max() > 0:
            out.append(f"{k}: max |diff| {d.max():.3e} in {int((d > 0).sum())} of {d.size}")
    return out


def run_lockstep(scene, steps=None, check_every=1):
    """Step both engines; return the per-step list of mismatch descriptions."""
    gpu, cpu = make_engines(scene)
    report = []
    try:
        for k, (frame, g) in enumerate(scene.frames(steps)):
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            if k % check_every and k != (steps or scene.steps) - 1:
                continue
            msgs = table_mismatches(gpu.read_tables(), cpu.read_tables())
            msgs += state_mismatches(gpu.read_state(pressure=True, expansion=True),
                                     cpu.read_state(pressure=True, expansion=True))
            if scene.grid()["smokeDivisor"] > 1:
                msgs += state_mismatches(gpu.read_smoke(), cpu.read_smoke(), keys=("smoke",))
            if msgs:
                report.append((k, msgs))
    finally:
        gpu.dispose()
    return report
