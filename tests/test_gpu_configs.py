"""The workloads at their own sizes: HIP against the FP32 restatement, bit for bit.

* campfire (plain method, small): 60 steps at dt 1/60, then the 256 x 256 full-resolution
  render (cubic filter, 256 ray steps, sun (.5, 1, .3) at intensity 3, lighting divisor 4);
* four fires (plain FP32, single large simulation): lockstep until at least 4096 bricks
  are active, then the 1920 x 1080 render with self-shadowing;
* brick churn (eight radial bursts in a 20 m cube): lockstep through the last burst, so
  every burst's allocation is compared;
* serving: eight requests of the serving stream batched together, joining at different
  rounds, their pressure solved by defect correction with BF16 levels at a fixed cycle
  count, each member against its own restatement.

The HIP kernels and the restatement evaluate the same float expression trees, so the bar
is equality. The restatement runs its per-brick passes on every host thread.
"""
import pytest

from lockstep import apply, make_engines, run_lockstep, state_mismatches, table_mismatches
from test_gpu_render import render_mismatches

pytestmark = pytest.mark.gpu


def test_campfire_60_steps_full_resolution_render():
    from firepro import abi, scenes
    sc = scenes.campfire(steps=60)
    gpu, cpu = make_engines(sc)
    try:
        for k, (frame, g) in enumerate(sc.frames()):
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            if k % 10 == 9:
                print(f"case campfire step {k}", flush=True)
                msgs = table_mismatches(gpu.read_tables(), cpu.read_tables())
                msgs += state_mismatches(gpu.read_state(pressure=True, expansion=True),
                                         cpu.read_state(pressure=True, expansion=True))
                assert msgs == [], (k, msgs)
        bricks = gpu.stats().active_bricks
        assert bricks > 0, bricks
        cam = abi.camera((-4.12, 1.96, 5.30), (0, 1.25, 0), 44, 256, 256)
        rs = abi.render_settings(half_resolution=False, filter="cubic", ray_steps=256, lighting_divisor=4,
                                 sun_direction=(0.5, 1.0, 0.3), sun_intensity=3.0)
        print(f"case campfire 256x256 full resolution, {bricks} bricks", flush=True)
        msgs, img = render_mismatches(gpu, cpu, cam, rs, "campfire 256")
        assert msgs == []
        assert (img[..., 3] > 0.01).mean() > 0.01, "the fire covers part of the image"
    finally:
        gpu.dispose()


def test_four_fires_past_4096_bricks_and_1080p_render():
    """Four fires to step 125 (past 4096 active bricks, about 2.1 M voxels), compared every
    25 steps and at the end, then the scene camera's 1920 x 1080 render (half resolution,
    the default) with the sun's self-shadowing."""
    from firepro import abi, scenes
    steps = 125
    sc = scenes.four_fires(steps=steps)
    gpu, cpu = make_engines(sc)
    try:
        for k, (frame, g) in enumerate(sc.frames()):
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            if k % 25 == 24:
                print(f"case four fires step {k}: {gpu.stats().active_bricks} bricks", flush=True)
                msgs = table_mismatches(gpu.read_tables(), cpu.read_tables())
                msgs += state_mismatches(gpu.read_state(pressure=True, expansion=True),
                                         cpu.read_state(pressure=True, expansion=True))
                assert msgs == [], (k, msgs)
        bricks = gpu.stats().active_bricks
        assert bricks >= 4096, bricks
        c = sc.camera
        cam = abi.camera(c["position"], c["target"], c["fov"], c["width"], c["height"])
        print(f"case four fires 1920x1080, {bricks} bricks", flush=True)
        msgs, img = render_mismatches(gpu, cpu, cam, abi.render_settings(), "four fires 1080p")
        assert msgs == []
        assert (img[..., 3] > 0.01).mean() > 0.01, "the fires cover part of the image"
    finally:
        gpu.dispose()


def test_explosions_through_every_burst():
    """The brick-churn workload to step 150: its eight bursts fire at seeded steps in
    0-150 (the last at step 136), each allocating and later freeing hundreds of bricks."""
    from firepro import scenes
    sc = scenes.explosions(steps=150)
    # The trigger steps, drawn from the scene's own PCG32 stream as scenes.explosions draws
    # them (step, position, fuel per burst): every one is inside the run.
    from firepro.scenes import PCG32
    rng = PCG32(1)
    triggers = []
    for _ in range(8):
        triggers.append(rng.integer(0, 150))
        rng.uniform(-10, 10), rng.uniform(0, 20), rng.uniform(-10, 10), rng.log_uniform(0.5, 5)
    assert max(triggers) < 150, triggers
    print(f"case explosions, bursts at steps {sorted(triggers)}", flush=True)
    assert run_lockstep(sc, check_every=10) == []


def test_serving_members_against_restatement():
    """Eight requests of the serving stream (1-4 fires, voxel 0.025-0.1 m, brick 8), the
    smaller ones so the restatement keeps up, batched with joins at rounds 0-3 and
    leaves when each has run its steps. Every member's pressure is solved by defect
    correction with BF16 levels at exactly 3 V-cycles a step (tolerance 0), and each
    member's tables, fields, velocity and pressure equal its own restatement's after
    every round it ran."""
    from firepro import abi, serving
    from firepro.engine import Batch, FluidSimulation
    from oracle import OracleSimulation
    reqs = [r for r in serving.requests(count=64, rate=20.0, seed=1) if r.target_bricks <= 1024][:8]
    assert len(reqs) == 8
    start = [0, 0, 1, 1, 2, 2, 3, 3]
    steps = [12, 10, 12, 8, 10, 12, 6, 9]
    gpus = [FluidSimulation.create(r.grid) for r in reqs]
    cpus = [OracleSimulation.create(r.grid, 32) for r in reqs]
    for r, gpu, cpu in zip(reqs, gpus, cpus):
        apply(gpu, cpu, r.settings)
        gpu.set_pressure("bf16", 0.0, 3)
        cpu.set_pressure(abi.PRESSURE_MODES["bf16"], 0.0, 3)
    batch = Batch()
    try:
        for t in range(max(s + n for s, n in zip(start, steps))):
            live = [i for i in range(8) if start[i] <= t < start[i] + steps[i]]
            print(f"case serving round {t}: members {live}", flush=True)
            batch.step([gpus[i] for i in live], [reqs[i].frame for i in live])
            for i in live:
                cpus[i].step(reqs[i].frame)
                msgs = table_mismatches(gpus[i].read_tables(), cpus[i].read_tables())
                msgs += state_mismatches(gpus[i].read_state(pressure=True, expansion=True),
                                         cpus[i].read_state(pressure=True, expansion=True))
                assert msgs == [], (t, i, msgs)
                assert gpus[i].pressure_report()[0] == 3
        assert sum(g.stats().active_bricks for g in gpus) > 0
    finally:
        batch.dispose()
        for g in gpus:
            g.dispose()
