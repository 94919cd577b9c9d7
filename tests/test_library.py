"""The library layer's host logic (src/library mirror): option resolution and validation,
and the plain-data frames FireSimulation assembles each step."""
import math

import pytest

from firepro import library, scenes


def test_defaults_resolve_like_the_reference():
    o = library.resolve_simulation()
    assert o["voxelSize"] == 0.02 and o["brickSize"] == 16 and o["velocityDivisor"] == 2
    assert o["grid"] == {"maxVoxels": 8_000_000, "cutoff": 0.05, "ground": True}
    assert o["motion"] == {"buoyancy": 1.6, "smokeWeight": 0.12, "damping": 0.1, "vorticity": 3.2}
    assert o["fuel"] == {"enabled": False, "ignitionHeat": 0.5, "burnRate": 4}


@pytest.mark.parametrize("options,message", [
    ({"voxelSize": 0.001}, "voxelSize"),
    ({"brickSize": 12}, "brickSize"),
    ({"velocityDivisor": 3}, "velocityDivisor"),
    ({"bogus": 1}, "Unknown simulation option"),
    ({"grid": {"maxVoxels": 100}}, "grid.maxVoxels"),
    ({"motion": {"buoyancy": -1}}, "motion.buoyancy"),
    ({"rendering": {"filter": "sharp"}}, "rendering.filter"),
    ({"grid": {"cutoff": float("nan")}}, "grid.cutoff"),
    ({"seed": 1.5}, "seed must be an integer"),
])
def test_invalid_options_raise(options, message):
    with pytest.raises(ValueError, match=message):
        library.resolve_simulation(options)


def test_emitter_defaults_and_source_frame():
    sim = library.FireSimulation({"voxelSize": 0.05})
    fire = sim.add_fire({"position": (1, 2, 3)})
    src = fire.source()
    assert src.size == (0.38, 0.38, 0.38) and src.flame == 1 and src.heat_rate == 1.5
    assert src.velocity == pytest.approx((0, 1.4, 0)) and src.velocity_response == 60
    plain = sim.add_emitter({"emission": {"heatRate": 10, "fuelRate": 2}})
    s2 = plain.source()
    assert s2.velocity_response == 0 and s2.velocity == (0, 0, 0) and s2.fuel_rate == 2
    plain.active = False
    assert plain.source() is None


def test_explosion_frames_once_with_rates_per_step():
    """A triggered burst is a one-step sphere source whose rates are charge / FIXED_DT
    (FireSimulation.ts:471-493)."""
    sim = library.FireSimulation({"voxelSize": 0.05})
    x = sim.add_explosion({"radius": 0.4, "charge": {"heat": 2, "fuel": 1.5}, "outwardSpeed": 8})
    x.trigger((0, 1, 0))
    f1 = sim.build_frame()
    assert len(f1.sources) == 1
    s = f1.sources[0]
    assert s.heat_rate == pytest.approx(2 * 60) and s.fuel_rate == pytest.approx(1.5 * 60)
    assert s.velocity_response == 60 and s.outward_speed == 8 and s.variation == (0.5, 0)
    assert sim.build_frame().sources == []


def test_update_runs_at_most_one_fixed_step():
    sim = library.FireSimulation({"voxelSize": 0.05})
    steps = []
    sim.fluid = object()
    sim.step = lambda: steps.append(1)
    sim.update(1 / 120)
    assert steps == []
    sim.update(1 / 120)
    assert len(steps) == 1
    sim.update(0.1)  # behind by several steps: one step, the rest dropped
    assert len(steps) == 2
    assert sim.dropped_time == pytest.approx(math.floor((0.1 - 1 / 60 + 1e-12) * 60) / 60)


def test_force_frames():
    sim = library.FireSimulation({"voxelSize": 0.05})
    e = sim.add_fire()
    sim.add_force({"type": "wind", "direction": (3, 0, 4), "strength": 2})
    sim.add_force({"type": "vortex", "strength": 1, "radius": 2, "lift": 0.5}, targets=[e])
    f = sim.build_frame()
    assert f.forces[0].vector == pytest.approx((0.6, 0, 0.8)) and f.forces[0].sources is None
    assert f.forces[1].scale == 2 and f.forces[1].lift == 0.5 and len(f.forces[1].sources) == 1


def test_scene_frames_are_deterministic():
    a = [f.sources[0].position for f, _ in scenes.four_fires(steps=2).frames()]
    b = [f.sources[0].position for f, _ in scenes.four_fires(steps=2).frames()]
    assert a == b
    xs = [f for f, _ in scenes.explosions(steps=151).frames()]
    assert sum(len(f.sources) for f in xs) == 8
