"""Pins the oracle's coarse smoke grid (smokeDivisor 2 and 4, smoke.wgsl) with facts it does
not restate: conservation of the box average, exact exponential decay, the fine X channel
holding one step's production only, and agreement with the full-resolution grid.

With buoyancy, smoke weight, vorticity, expansion and forces all zero the velocity stays
zero, so transport is the identity and every change of the coarse density comes from the
production it adds and the decay it applies.
"""
import numpy as np
import pytest

import oracle
from firepro import scenes
from firepro.library import FireSimulation


def still_smoke(smoke_divisor: int, dissipation: float, divisor: int = 2, brick: int = 8):
    sim = FireSimulation({"voxelSize": 0.05, "brickSize": brick, "velocityDivisor": divisor,
                          "smokeDivisor": smoke_divisor, "seed": 3,
                          "grid": {"cutoff": 1e-6},
                          "motion": {"buoyancy": 0, "smokeWeight": 0, "damping": 0, "vorticity": 0},
                          "smoke": {"dissipation": dissipation}})
    e = sim.add_emitter({"position": (0.12, 0.55, -0.07), "shape": {"radius": 0.3},
                         "emission": {"flame": 0, "heatRate": 0, "smokeRate": 2}, "velocity": None})
    grid = {"voxelSize": 0.05, "velocityDivisor": divisor, "brickSize": brick,
            "smokeDivisor": smoke_divisor, "scalarMacCormack": True}
    return sim, e, oracle.OracleSimulation.create(grid)


def step(sim, orc):
    orc.apply_grid(sim.grid_settings())
    orc.step(sim.build_frame())


def by_id(state: dict, key: str) -> dict:
    return {int(i): row for i, row in zip(state["ids"], state[key])}


@pytest.mark.parametrize("smoke_divisor", [2, 4])
def test_coarse_smoke_conserves_production(smoke_divisor):
    """At rest, the coarse density times the cells it covers is the sum of the fine
    production of every step: the box average neither gains nor loses mass (a wrong
    divisor of the average, a dropped production term or a decay applied with zero
    dissipation would break it). The fine X channel holds the latest step's production
    only, the same every step, and the velocity stays zero."""
    sim, _, orc = still_smoke(smoke_divisor, 0.0)
    fine = []
    for k in range(5):
        step(sim, orc)
        st = orc.read_state()
        fine.append(float(st["fields"][..., 0].astype(np.float64).sum()))
        assert not np.any(st["velocity"])
        coarse = float(orc.read_smoke()["smoke"].astype(np.float64).sum()) * smoke_divisor ** 3
        assert coarse == pytest.approx(sum(fine), rel=2e-6)
    assert fine[0] > 1.0
    assert max(fine) - min(fine) <= 1e-6 * fine[0]


def test_coarse_smoke_decays_exactly():
    """With the sources off and no motion, each step multiplies every coarse cell by
    e^{-dissipation dt} (one float product), and cells that fall below the dormant cutoff
    0.001 become zero."""
    dissipation, dt = 0.7, 1 / 60
    sim, emitter, orc = still_smoke(2, dissipation)
    for _ in range(3):
        step(sim, orc)
    emitter.active = False
    loss = np.float32(oracle.fexp(float(np.float32(-np.float32(dissipation) * np.float32(dt)))))
    assert abs(float(loss) - np.exp(-dissipation * dt)) < 1e-6
    before = by_id(orc.read_smoke(), "smoke")
    assert max(float(v.max()) for v in before.values()) > 0.05
    for _ in range(4):
        step(sim, orc)
        after = by_id(orc.read_smoke(), "smoke")
        # Bricks the source alone kept leave once it stops; they held no smoke.
        assert set(after) <= set(before)
        assert not any(before[i].any() for i in set(before) - set(after))
        for i in after:
            expect = before[i] * loss
            expect[expect < np.float32(0.001)] = 0
            np.testing.assert_array_equal(after[i], expect)
        assert any(v.any() for v in after.values())
        assert not np.any(orc.read_state()["fields"][..., 0])
        before = after


def test_coarse_smoke_tracks_full_resolution():
    """On the small scene (fire with fuel, smoke source, burst, forces, collider) the coarse
    grid's smoke mass stays within 10% of the full-resolution grid's after 12 steps, while
    heat, which both grids carry at full resolution, agrees within 1%."""
    masses, heats = {}, {}
    for sd in (1, 2):
        sc = scenes.small(brick=8, divisor=2, smoke=sd, steps=12)
        orc = oracle.OracleSimulation.create(sc.grid())
        for frame, g in sc.frames():
            orc.apply_grid(g)
            orc.step(frame)
        st = orc.read_state()
        heats[sd] = float(st["fields"][..., 1].astype(np.float64).sum())
        if sd == 1:
            masses[sd] = float(st["fields"][..., 0].astype(np.float64).sum())
        else:
            masses[sd] = float(orc.read_smoke()["smoke"].astype(np.float64).sum()) * sd ** 3
    assert masses[1] > 10
    assert masses[2] == pytest.approx(masses[1], rel=0.10)
    assert heats[2] == pytest.approx(heats[1], rel=0.01)


def test_smoke_divisor_validation():
    """smokeDivisor is 1, 2 or 4 (FluidSimulation.ts:378-379); a full-resolution simulation
    has no coarse grid to read."""
    base = {"voxelSize": 0.05, "velocityDivisor": 2, "brickSize": 8, "scalarMacCormack": True}
    with pytest.raises(ValueError, match="smokeDivisor"):
        oracle.OracleSimulation.create(dict(base, smokeDivisor=3))
    full = oracle.OracleSimulation.create(dict(base, smokeDivisor=1))
    with pytest.raises(ValueError, match="coarse smoke"):
        full.read_smoke()
