"""HIP path against the FP32 restatement: integer tables and every field bit-exact.

The HIP kernels and the restatement evaluate the same float expression trees with IEEE
basic operations (no contraction, correctly rounded division and square root, exp from a
shared published method), so the bar is equality, not a tolerance.
"""
import numpy as np
import pytest

from lockstep import apply, make_engines, run_lockstep

pytestmark = pytest.mark.gpu


def _scenes():
    from firepro import scenes
    return scenes


@pytest.mark.parametrize("brick,divisor", [(8, 2), (16, 2), (8, 1), (16, 4)])
def test_small_scene_bit_exact(brick, divisor):
    sc = _scenes().small(steps=6, brick=brick, divisor=divisor)
    print(f"case small brick={brick} divisor={divisor}", flush=True)
    assert run_lockstep(sc) == []


def test_tilted_colliders_bit_exact():
    """A box turned about all three axes and a sphere in the plume: their solids, and the
    cells the solver treats as solid through them, bit-exact."""
    print("case small tilted colliders", flush=True)
    assert run_lockstep(_scenes().small(steps=8, tilted=True)) == []


@pytest.mark.parametrize("brick,divisor,smoke", [(8, 2, 2), (8, 1, 2), (8, 2, 4), (16, 2, 2), (16, 4, 2), (16, 4, 4)])
def test_coarse_smoke_bit_exact(brick, divisor, smoke):
    """smokeDivisor 2 and 4 (smoke.wgsl): the coarse density, the fine fields that stage
    its production, the velocity it drives and every table, bit-exact."""
    sc = _scenes().small(steps=8, brick=brick, divisor=divisor, smoke=smoke)
    print(f"case coarse smoke brick={brick} divisor={divisor} smoke={smoke}", flush=True)
    assert run_lockstep(sc) == []


def test_coarse_smoke_first_order_and_campfire():
    sc = _scenes().small(steps=5, smoke=2)
    sc.sim.options["scalarMacCormack"] = False
    print("case coarse smoke first order", flush=True)
    assert run_lockstep(sc) == []
    print("case campfire smoke=2", flush=True)
    assert run_lockstep(_scenes().campfire(steps=30, smoke=2), check_every=5) == []


def test_coarse_smoke_refuses_batch():
    from firepro.engine import Batch, FluidSimulation
    sc = _scenes().small(steps=2, smoke=2)
    gpu = FluidSimulation.create(sc.grid())
    batch = Batch()
    try:
        frames = list(sc.frames())
        gpu.step(frames[0][0])
        with pytest.raises(ValueError, match="smokeDivisor"):
            batch.step([gpu], [frames[1][0]])
        full = FluidSimulation.create(_scenes().small(steps=1).grid())
        with pytest.raises(ValueError, match="coarse smoke"):
            full.read_smoke()
        full.dispose()
    finally:
        batch.dispose()
        gpu.dispose()


@pytest.mark.parametrize("smoke,colliders,targeted", [(1, False, False), (2, False, False), (1, True, False),
                                                      (1, False, True)])
def test_mesh_emitters_bit_exact(smoke, colliders, targeted):
    """Mesh emitters (box, torus, disk): the rasterized records and brick tables drive
    allocation, emission and the source velocity identically on both sides; with mesh
    colliders (a tilted box, an octahedron) their occupancy draws the same solids; with a
    force targeting mesh emitters, their covered velocity cells feel it fully."""
    print(f"case meshes smoke={smoke} colliders={colliders} targeted={targeted}", flush=True)
    assert run_lockstep(_scenes().meshes(steps=10, smoke=smoke, colliders=colliders, targeted=targeted)) == []


def test_first_order_scalars_bit_exact():
    sc = _scenes().small(steps=4)
    sc.sim.options["scalarMacCormack"] = False
    assert run_lockstep(sc) == []


def test_campfire_bit_exact():
    sc = _scenes().campfire(steps=40)
    assert run_lockstep(sc, check_every=5) == []


def test_explosions_churn_tables():
    """Brick churn: bursts allocate and free thousands of bricks; tables stay bit-exact."""
    sc = _scenes().explosions(steps=40)
    assert run_lockstep(sc, check_every=4) == []


def test_four_fires_collider_turbulence():
    sc = _scenes().four_fires(steps=4)
    assert run_lockstep(sc, check_every=2) == []


def test_pool_growth_and_budget():
    """A tiny voxel budget forces failed requests, growth to the limit and poolLimited."""
    from firepro import scenes
    sc = scenes.small(steps=8)
    sc.sim.options["grid"]["maxVoxels"] = 4096 * 3
    gpu, cpu = make_engines(sc)
    try:
        for frame, g in sc.frames():
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            ta, tb = gpu.read_tables(), cpu.read_tables()
            for k in ("pages", "active", "free_slots"):
                assert np.array_equal(ta[k], tb[k]), k
            assert ta["failures"] == tb["failures"]
        sa, sb = gpu.stats(), cpu.stats()
        assert sa.pool_limited == sb.pool_limited == 1
        assert sa.pool_slots == sb.pool_slots
    finally:
        gpu.dispose()


def test_reset_restarts_identically():
    from firepro import scenes
    sc = scenes.small(steps=3)
    frames = list(sc.frames())
    gpu, _ = make_engines(sc)
    try:
        def run():
            for frame, g in frames:
                gpu.cutoff, gpu.voxel_budget, gpu.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                gpu.step(frame)
            return gpu.read_state(pressure=True)
        a = run()
        gpu.reset()
        assert gpu.stats().steps == 0
        b = run()
        for k in ("ids", "fields", "velocity", "pressure"):
            assert np.array_equal(a[k], b[k]), k
    finally:
        gpu.dispose()


def test_invalid_inputs_raise():
    from firepro import abi, scenes
    from firepro.engine import FluidSimulation
    with pytest.raises(ValueError, match="voxelSize"):
        FluidSimulation(0.0, 2, 16)
    with pytest.raises(ValueError, match="brickSize"):
        FluidSimulation(0.05, 2, 12)
    sc = scenes.small(steps=1)
    gpu = FluidSimulation.create(sc.grid())
    try:
        frame, g = next(sc.frames())
        with pytest.raises(ValueError, match="Simulation step"):
            gpu.step(frame, 0.1)
        frame.ramps = dict(frame.ramps)
        frame.ramps["heat"] = [(0, 0), (0.5, 1)]
        with pytest.raises(ValueError, match="span lifetime"):
            gpu.step(frame)
    finally:
        gpu.dispose()
    with pytest.raises(abi.FireProError):
        gpu.stats()


def test_solver_field_view():
    """showSolverField: the divergence view holds the step's right-hand side h^2 (div u -
    expansion), equal to the restatement's; switching to pressure copies at once and then
    after every step, batched steps included; with no field shown the read is refused."""
    from firepro import abi, scenes
    from firepro.engine import Batch
    sc = scenes.small(steps=6)
    frames = list(sc.frames())
    gpu, cpu = make_engines(sc)
    batch = Batch()
    try:
        gpu.show_solver_field("divergence")
        for frame, g in frames[:5]:
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        view = gpu.read_solver_field()
        assert np.array_equal(view["ids"], gpu.read_state()["ids"])
        assert np.array_equal(view["values"].astype(np.float64), cpu.rhs())
        gpu.show_solver_field("pressure")
        assert np.array_equal(gpu.read_solver_field()["values"], gpu.read_state(pressure=True)["pressure"])
        frame, g = frames[5]
        apply(gpu, cpu, g)
        batch.step([gpu], [frame])
        assert np.array_equal(gpu.read_solver_field()["values"], gpu.read_state(pressure=True)["pressure"])
        gpu.show_solver_field(None)
        with pytest.raises(abi.FireProError, match="no solver field"):
            gpu.read_solver_field()
        with pytest.raises(ValueError):
            gpu.show_solver_field("velocity")
        st = gpu.stats()
        assert st.memory_bytes > 0 and st.force_memory_bytes > 0 and st.mesh_memory_bytes >= 0
        rb = gpu.render_bricks
        assert rb["active_count"] == st.active_bricks and len(rb["boxes"]) >= 1
    finally:
        batch.dispose()
        gpu.dispose()


@pytest.mark.gpu
def test_scene_arrays_follow_changes_bit_exact():
    """The step skips the upload of a scene array whose bytes are unchanged since its last
    copy. Here the collider moves, the vortex's strength changes and the fire's position
    changes between steps (and then stay put), so each array is copied again only when it
    changed; the HIP steps must stay identical to the restatement's throughout."""
    from dataclasses import replace
    from firepro import scenes
    from lockstep import state_mismatches, table_mismatches
    sc = scenes.small(steps=10, burst=False)
    gpu, cpu = make_engines(sc)
    try:
        for k, (frame, g) in enumerate(sc.frames()):
            if k >= 3:
                frame.colliders = [replace(c, position=(c.position[0] + 0.05, c.position[1], c.position[2]))
                                   for c in frame.colliders]
            if k >= 5:
                frame.forces = [replace(f, strength=f.strength * 1.5) if f.type == "vortex" else f
                                for f in frame.forces]
            if k >= 7:
                frame.sources = [replace(s, position=(s.position[0], s.position[1], s.position[2] + 0.04))
                                 for s in frame.sources]
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            msgs = table_mismatches(gpu.read_tables(), cpu.read_tables())
            msgs += state_mismatches(gpu.read_state(pressure=True, expansion=True),
                                     cpu.read_state(pressure=True, expansion=True))
            assert not msgs, f"step {k}: {msgs}"
    finally:
        gpu.dispose()
