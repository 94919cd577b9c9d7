"""Serving many independent simulations on one GPU with continuous batching.

Requests arrive over time (Poisson arrivals); each is a small fire scene that runs a given
number of steps and then leaves. A request is admitted as soon as it has arrived and its
device memory fits the HBM budget; every admitted simulation stays resident on the device.
Each round advances every resident simulation by one step:

* batched: one fp_batch_step over all of them, one launch per stage (firepro.engine.Batch);
* one simulation per launch: fp_sim_step for each in turn, the reference's way.

Requests that arrive while others run join at the next round, and finished ones leave at
once, so membership changes from round to round. The report gives aggregate active
voxel-steps/s, per-step and per-request latency percentiles and the peak memory held.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi
from .engine import Batch, FluidSimulation
from .library import FireSimulation
from .scenes import PCG32

BRICK = 8
DIVISOR = 2


def memory_estimate(target_bricks: int, brick: int = BRICK, divisor: int = DIVISOR) -> int:
    """Device bytes a simulation holds at its voxel budget: per pool slot its field,
    backtraces, expansion, velocity, curl, divergence, pressure and multigrid arrays and
    defect correction's x (two parts), r and right-hand side with 16-bit level copies, plus
    the tile tables (a tile per 16 bricks, generously)."""
    nf, nv = brick ** 3, (brick // divisor) ** 3
    per_slot = (nf * (16 * 6 + 4) + nf // 8 + nv * (4 * 3 * 4 + 16 + 4 + 8 + 64) + int(nv * 14 * 1.15)
                + nv * 4 * 4 + int(nv * 6 * 1.15))
    tiles = max(16, target_bricks // 16)
    per_tile = 64 * (4 * 3 + 28 * 4 + 4 * 3) + 32 * 4 + 64 * 14
    return int(target_bricks * per_slot + tiles * per_tile + (4 << 20))


@dataclass
class Request:
    index: int
    arrival: float
    voxel: float
    fires: int
    target_bricks: int
    steps: int
    frame: abi.PackedFrame
    grid: dict
    settings: dict
    memory: int
    # filled while serving
    sim: Optional[FluidSimulation] = None
    timer: Optional[abi.StreamTimer] = None  # one simulation per launch: its steps' device times
    done: int = 0
    admitted: float = 0.0
    finished: float = 0.0


def requests(count: int = 256, rate: float = 20.0, seed: int = 1) -> List[Request]:
    """The serving workload: Poisson arrivals at `rate` per second; per request 1-4 fire
    spheres, a voxel of 0.025-0.1 m, a brick budget log-uniform in 16-8192 (enforced as
    maxVoxels = bricks x 8^3) and 30-300 steps, all from one PCG32 stream."""
    rng = PCG32(seed, stream=91)
    out, t = [], 0.0
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        fires = rng.integer(1, 4)
        voxel = rng.uniform(0.025, 0.1)
        target = int(round(rng.log_uniform(16, 8192)))
        steps = rng.integer(30, 300)
        sim = FireSimulation({"voxelSize": voxel, "brickSize": BRICK, "velocityDivisor": DIVISOR, "seed": i + 1,
                              "grid": {"maxVoxels": max(4096, target * BRICK ** 3)},
                              "fuel": {"enabled": True, "ignitionHeat": 0.5, "burnRate": 4}})
        scale = voxel / 0.05
        for _ in range(fires):
            pos = (rng.uniform(-1.0, 1.0) * scale, rng.uniform(0.1, 0.6) * scale, rng.uniform(-1.0, 1.0) * scale)
            sim.add_fire({"position": pos, "radius": 0.3 * scale})
        o = sim.options
        grid = {"voxelSize": o["voxelSize"], "velocityDivisor": o["velocityDivisor"], "brickSize": o["brickSize"],
                "smokeDivisor": o["smokeDivisor"], "scalarMacCormack": o["scalarMacCormack"]}
        out.append(Request(i, t, voxel, fires, target, steps, abi.PackedFrame(sim.build_frame()), grid,
                           sim.grid_settings(), memory_estimate(target)))
    return out


@dataclass
class Report:
    mode: str
    requests: int = 0
    sim_steps: int = 0
    voxel_steps: int = 0
    wall_s: float = 0.0
    busy_s: float = 0.0
    rounds: int = 0
    step_ms: List[float] = field(default_factory=list)
    request_s: List[float] = field(default_factory=list)
    peak_resident: int = 0
    peak_reserved_bytes: int = 0
    phases: dict = field(default_factory=dict)  # batched: summed phase ms of the steps
    admit_s: float = 0.0  # creating admitted simulations
    account_s: float = 0.0  # per-round statistics and disposal of finished simulations
    pressure: Optional[tuple] = None
    cycles: int = 0  # defect correction's V-cycles over all steps
    max_residual: float = 0.0  # the largest final relative residual |b - Ax| / |b| of a step
    within_tolerance: int = 0  # steps whose solve met the tolerance

    def summary(self) -> dict:
        pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None  # noqa: E731
        return {
            "mode": self.mode, "requests": self.requests, "sim_steps": self.sim_steps,
            "active_voxel_steps": self.voxel_steps,
            "voxel_steps_per_s": self.voxel_steps / self.wall_s if self.wall_s else 0.0,
            "voxel_steps_per_busy_s": self.voxel_steps / self.busy_s if self.busy_s else 0.0,
            "wall_s": self.wall_s, "busy_s": self.busy_s, "rounds": self.rounds,
            "p50_step_ms": pct(self.step_ms, 50), "p99_step_ms": pct(self.step_ms, 99),
            "p50_request_s": pct(self.request_s, 50), "p99_request_s": pct(self.request_s, 99),
            "peak_resident": self.peak_resident, "peak_reserved_gb": self.peak_reserved_bytes / 1e9,
            "phase_ms": {k: round(v, 1) for k, v in self.phases.items()},
            "admit_s": self.admit_s, "account_s": self.account_s,
            "pressure": list(self.pressure) if self.pressure else "vcycle",
            "mean_cycles": self.cycles / self.sim_steps if self.pressure and self.sim_steps else None,
            "max_relative_residual": self.max_residual if self.pressure else None,
            "steps_within_tolerance": self.within_tolerance / self.sim_steps if self.pressure and self.sim_steps else None,
        }


def serve(reqs: List[Request], batched: bool = True, hbm_budget: float = 200e9, dt: float = 1 / 60,
          pressure: Optional[tuple] = None, progress=None, graphs: bool = True) -> Report:
    """Runs the request stream in real time and returns its report. `pressure`, as the
    arguments of FluidSimulation.set_pressure, is every simulation's pressure policy (the
    reference's single V-cycle when None). `progress`, if given, is called with a line of
    text every few seconds. `graphs` False launches the batched kernels one by one instead
    of replaying captured graphs."""
    rep = Report("batched" if batched else "one simulation per launch")
    rep.pressure = pressure
    said = time.perf_counter()
    pending = sorted(reqs, key=lambda r: r.arrival)
    resident: List[Request] = []
    reserved = 0
    batch = Batch() if batched else None
    if batch is not None and not graphs:
        batch.use_graphs(False)
    t0 = time.perf_counter()
    try:
        while pending or resident:
            now = time.perf_counter() - t0
            # Admission: arrived requests in order, while their memory fits.
            a0 = time.perf_counter()
            while pending and pending[0].arrival <= now and reserved + pending[0].memory <= hbm_budget:
                r = pending.pop(0)
                r.sim = FluidSimulation.create(r.grid)
                if not batched:
                    r.timer = abi.StreamTimer(r.sim.stream())
                g = r.settings
                r.sim.cutoff, r.sim.voxel_budget, r.sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                if pressure is not None:
                    r.sim.set_pressure(*pressure)
                r.admitted = time.perf_counter() - t0
                reserved += r.memory
                resident.append(r)
            rep.admit_s += time.perf_counter() - a0
            rep.peak_resident = max(rep.peak_resident, len(resident))
            rep.peak_reserved_bytes = max(rep.peak_reserved_bytes, reserved)
            if not resident:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            s0 = time.perf_counter()
            if batched:
                batch.step([r.sim for r in resident], [r.frame for r in resident], dt)
                ms = (time.perf_counter() - s0) * 1e3
                rep.step_ms.extend([ms] * len(resident))
                for k, v in batch.phases.items():
                    rep.phases[k] = rep.phases.get(k, 0.0) + v
            else:
                # Each step's latency is device time on the simulation's own stream, from the
                # call to the completion of its last kernel (the call itself returns once the
                # step's readback lands, with the tail still queued).
                for r in resident:
                    r.timer.mark()
                    r.sim.step(r.frame, dt)
                    r.timer.mark()
            rep.busy_s += time.perf_counter() - s0
            rep.rounds += 1
            c0 = time.perf_counter()
            still = []
            for r in resident:
                st = r.sim.stats()
                rep.voxel_steps += st.active_voxels
                if pressure is not None:
                    cycles, residual = r.sim.pressure_report()
                    rep.cycles += cycles
                    rep.max_residual = max(rep.max_residual, residual)
                    rep.within_tolerance += residual <= pressure[1]
                rep.sim_steps += 1
                r.done += 1
                if r.done >= r.steps:
                    if r.timer is not None:
                        rep.step_ms.extend(r.timer.elapsed())
                        r.timer.close()
                        r.timer = None
                    r.finished = time.perf_counter() - t0
                    rep.request_s.append(r.finished - r.arrival)
                    rep.requests += 1
                    r.sim.dispose()
                    r.sim = None
                    reserved -= r.memory
                else:
                    still.append(r)
            resident = still
            rep.account_s += time.perf_counter() - c0
            if progress is not None and time.perf_counter() - said > 5.0:
                said = time.perf_counter()
                progress(f"{rep.mode}: {rep.requests} of {len(reqs)} requests done, {len(resident)} resident, "
                         f"{rep.rounds} rounds, {said - t0:.0f} s")
        rep.wall_s = time.perf_counter() - t0
    finally:
        for r in resident:
            if r.timer is not None:
                r.timer.close()
                r.timer = None
            if r.sim is not None:
                r.sim.dispose()
        if batch is not None:
            batch.dispose()
    return rep
