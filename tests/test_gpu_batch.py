"""Batched stepping (the serving form) on the GPU: each simulation's tables and fields are
the same alone and in the mix, bit for bit, while simulations join and leave the batch
(continuous batching) and grow their pools."""
import numpy as np
import pytest

from lockstep import apply, state_mismatches, table_mismatches

pytestmark = pytest.mark.gpu


def _scenes():
    from firepro import scenes
    return [
        scenes.small(seed=1, steps=8, brick=8, divisor=2),
        scenes.small(seed=2, steps=8, brick=8, divisor=2, collider=False),
        scenes.small(seed=3, steps=8, brick=8, divisor=2, burst=False),
        scenes.campfire(steps=8),
    ]


# Pressure policies: the reference's single V-cycle, and defect correction with levels in
# FP32, BF16 and FP16 to per-simulation tolerances and cycle budgets (some steps stop on
# the tolerance, some on the budget). With x in FP32 these scenes' relative residuals level
# off between 1e-6 and 1e-4 (about the FP32 unit roundoff times the operator's condition
# number), so the tolerances that must be met sit at 1e-4.
POLICIES = [
    ("vcycle", None),
    ("fp32", [(1e-4, 12), (1e-5, 12), (1e-5, 2), (1e-6, 12)]),
    ("bf16", [(1e-4, 16), (1e-4, 8), (1e-6, 2), (1e-5, 16)]),
    ("fp16", [(1e-4, 16), (1e-5, 16), (1e-4, 16), (1e-6, 3)]),
]


@pytest.mark.parametrize("mode,tolerances", POLICIES, ids=[p[0] for p in POLICIES])
def test_batch_equals_alone_with_joins_and_leaves(mode, tolerances):
    from firepro import scenes
    from firepro.engine import Batch, FluidSimulation
    # The campfire uses 16-cell bricks; a batch holds one layout, so use 8-cell ones here.
    scs = _scenes()[:3] + [scenes.small(seed=4, steps=8, brick=8, divisor=2)]
    # Simulation i runs steps [start_i, start_i + len) of the batch's clock.
    start = [0, 0, 2, 3]
    frames = [list(sc.frames()) for sc in scs]
    alone = [FluidSimulation.create(sc.grid()) for sc in scs]
    mixed = [FluidSimulation.create(sc.grid()) for sc in scs]
    if tolerances:
        for i in range(len(scs)):
            for sim in (alone[i], mixed[i]):
                sim.set_pressure(mode, *tolerances[i])
    reports = [[] for _ in scs]
    batch = Batch()
    try:
        for i, sim in enumerate(alone):
            for frame, g in frames[i]:
                sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                sim.step(frame)
                reports[i].append(sim.pressure_report())
        for t in range(max(s + len(f) for s, f in zip(start, frames))):
            members, fr = [], []
            for i, sim in enumerate(mixed):
                k = t - start[i]
                if 0 <= k < len(frames[i]):
                    frame, g = frames[i][k]
                    sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                    members.append(sim)
                    fr.append(frame)
            print(f"case batch {mode} clock {t}: {len(members)} members", flush=True)
            batch.step(members, fr)
            for i, sim in enumerate(mixed):
                k = t - start[i]
                if 0 <= k < len(frames[i]):
                    assert sim.pressure_report() == reports[i][k], (i, k)
        if tolerances:
            # Some steps stop on the tolerance, some on the cycle budget.
            flat = [(c, r, tolerances[i]) for i in range(len(scs)) for c, r in reports[i]]
            assert any(r <= tol for c, r, (tol, m) in flat), reports
            assert any(c == m and r > tol for c, r, (tol, m) in flat), reports
        for i in range(len(scs)):
            msgs = table_mismatches(mixed[i].read_tables(), alone[i].read_tables())
            msgs += state_mismatches(mixed[i].read_state(pressure=True, expansion=True),
                                     alone[i].read_state(pressure=True, expansion=True))
            assert msgs == [], (i, msgs)
            assert mixed[i].stats().steps == alone[i].stats().steps
    finally:
        batch.dispose()
        for s in alone + mixed:
            s.dispose()


@pytest.mark.parametrize("mode", ["vcycle", "bf16"])
def test_batch_mixes_tile_paths(mode):
    """A member with over a thousand tile slots (its levels below the first tile level run
    level by level) batched with small ones (theirs run in one workgroup each): every
    member ends as it does alone."""
    from firepro import scenes
    from firepro.engine import Batch, FluidSimulation
    scs = [scenes.scattered(steps=4), scenes.small(seed=1, steps=4), scenes.small(seed=2, steps=4, collider=False)]
    frames = [list(sc.frames()) for sc in scs]
    alone = [FluidSimulation.create(sc.grid()) for sc in scs]
    mixed = [FluidSimulation.create(sc.grid()) for sc in scs]
    if mode != "vcycle":
        for sim in alone + mixed:
            sim.set_pressure(mode, 0.0, 3)
    batch = Batch()
    try:
        for i, sim in enumerate(alone):
            for frame, g in frames[i]:
                sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                sim.step(frame)
        for t in range(4):
            for i, sim in enumerate(mixed):
                g = frames[i][t][1]
                sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
            print(f"case mixed tile paths {mode} step {t}", flush=True)
            batch.step(mixed, [frames[i][t][0] for i in range(3)])
        assert mixed[0].stats().tile_capacity > 512, "the large member exceeds one workgroup's tile slots"
        assert all(m.stats().tile_capacity <= 512 for m in mixed[1:])
        for i in range(3):
            msgs = table_mismatches(mixed[i].read_tables(), alone[i].read_tables())
            msgs += state_mismatches(mixed[i].read_state(pressure=True, expansion=True),
                                     alone[i].read_state(pressure=True, expansion=True))
            assert msgs == [], (i, msgs)
    finally:
        batch.dispose()
        for s in alone + mixed:
            s.dispose()


def test_batch_pool_growth_and_budget():
    """A member that outgrows its pool and hits its voxel budget inside a batch."""
    from firepro import scenes
    from firepro.engine import Batch, FluidSimulation
    a = scenes.small(seed=1, steps=8)
    a.sim.options["grid"]["maxVoxels"] = 4096 * 3
    b = scenes.small(seed=5, steps=8)
    scs = [a, b]
    frames = [list(sc.frames()) for sc in scs]
    alone = [FluidSimulation.create(sc.grid()) for sc in scs]
    mixed = [FluidSimulation.create(sc.grid()) for sc in scs]
    batch = Batch()
    try:
        for k in range(8):
            fr = []
            for i in range(2):
                frame, g = frames[i][k]
                for sim in (alone[i], mixed[i]):
                    sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                alone[i].step(frame)
                fr.append(frame)
            batch.step(mixed, fr)
        for i in range(2):
            assert table_mismatches(mixed[i].read_tables(), alone[i].read_tables()) == []
            assert state_mismatches(mixed[i].read_state(pressure=True), alone[i].read_state(pressure=True),
                                    keys=("fields", "velocity", "pressure")) == []
        assert mixed[0].pool_limited
    finally:
        batch.dispose()
        for s in alone + mixed:
            s.dispose()


def test_batch_rejects_mixed_layouts():
    from firepro import scenes
    from firepro.engine import Batch, FluidSimulation
    a = FluidSimulation.create(scenes.small(steps=1, brick=8).grid())
    b = FluidSimulation.create(scenes.small(steps=1, brick=16).grid())
    batch = Batch()
    try:
        f = [next(iter(scenes.small(steps=1).frames()))[0]] * 2
        with pytest.raises(ValueError, match="brick size"):
            batch.step([a, b], f)
        with pytest.raises(ValueError, match="twice"):
            batch.step([a, a], f)
        c = FluidSimulation.create(scenes.small(steps=1, brick=8).grid())
        try:
            c.set_pressure("bf16", 1e-6, 8)
            with pytest.raises(ValueError, match="pressure policy"):
                batch.step([a, c], f)
        finally:
            c.dispose()
    finally:
        batch.dispose()
        a.dispose()
        b.dispose()


def test_serving_loop_completes():
    """A short request stream through the continuous-batching server and the
    one-simulation-per-launch loop: every request finishes its steps in both."""
    from firepro import serving
    for batched in (True, False):
        reqs = serving.requests(count=12, rate=200.0, seed=3)
        for r in reqs:
            r.steps = min(r.steps, 20)
        rep = serving.serve(reqs, batched=batched)
        print(f"case serving batched={batched}: {rep.summary()}", flush=True)
        assert rep.requests == 12
        assert rep.sim_steps == sum(r.steps for r in reqs)
        assert rep.voxel_steps > 0
