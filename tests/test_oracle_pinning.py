"""Pins the CPU restatement (the checker) with facts it does not restate.

Each test fixes a part of the oracle by something independent of its own formula: closed
forms, brute force, invariants, or a published known-answer vector. See the per-test
docstrings for what a plausible mistake would trip.
"""
import math
import struct

import numpy as np
import pytest

import oracle
from firepro import scenes
from firepro.abi import DEFAULT_RAMPS, FRESH_ENTRY, FRESH_SLOT


def ulp_distance(a: float, b: float) -> int:
    ia = struct.unpack("<i", struct.pack("<f", a))[0]
    ib = struct.unpack("<i", struct.pack("<f", b))[0]
    return abs(ia - ib)


def test_fexp_error_bound():
    """exp from Cody-Waite reduction + Taylor polynomial: within 2 ulp of the correctly
    rounded single-precision exp over the range the solver uses, exact at 0."""
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-87, 88, 4000), rng.uniform(-1, 1, 4000), [0.0, -1e-6, 1e-6]])
    worst = 0
    for x in xs:
        xf = float(np.float32(x))
        ref = float(np.float32(math.exp(xf)))
        worst = max(worst, ulp_distance(oracle.fexp(xf), ref))
    assert worst <= 2
    assert oracle.fexp(0.0) == 1.0
    assert oracle.fexp(-200.0) == 0.0
    assert math.isinf(oracle.fexp(100.0))
    # subnormal results stay within 2 ulp of the subnormal grid
    assert abs(oracle.fexp(-100.0) - math.exp(-100.0)) <= 2 * 2.0 ** -149


def test_ramp_integral_is_exact_area():
    """rampIntegral is the exact area under the piecewise-linear ramp (flame-common.wgsl:
    170-184), checked against fine trapezoid integration of numpy.interp."""
    rng = np.random.default_rng(3)
    for ch, name in enumerate(("smoke", "heat", "expansion")):
        xs, ys = zip(*DEFAULT_RAMPS[name])
        for _ in range(40):
            lo, hi = sorted(rng.uniform(-0.1, 1.1, 2))
            grid = np.linspace(max(lo, 0.0), min(hi, 1.0), 200001)
            ref = np.trapezoid(np.interp(grid, xs, ys), grid) if grid[-1] > grid[0] else 0.0
            assert abs(oracle.ramp_integral(DEFAULT_RAMPS, ch, lo, hi) - ref) < 1e-9


def test_evaluate_flame_closed_form():
    """A flame of lifetime `before` loses dt/lifespan of it and yields rate * lifespan *
    the ramp's area over what it lost; expansion is averaged over the step."""
    dt, life = 1 / 60, 1.2
    for before in (0.05, 0.3, 0.5, 0.99, 1.0):
        after, smoke, heat, exp_ = oracle.evaluate_flame(DEFAULT_RAMPS, dt, life, 2.0, 4.0, 1.0, before)
        a = max(0.0, before - dt / life)
        assert after == pytest.approx(a, abs=1e-15)
        for ch, rate, got in ((0, 2.0, smoke), (1, 4.0, heat)):
            xs, ys = zip(*DEFAULT_RAMPS[("smoke", "heat")[ch]])
            g = np.linspace(a, before, 100001)
            assert got == pytest.approx(rate * life * np.trapezoid(np.interp(g, xs, ys), g), rel=1e-6, abs=1e-12)
        xs, ys = zip(*DEFAULT_RAMPS["expansion"])
        g = np.linspace(a, before, 100001)
        assert exp_ == pytest.approx(life * np.trapezoid(np.interp(g, xs, ys), g) / dt, rel=1e-6, abs=1e-12)
    assert list(oracle.evaluate_flame(DEFAULT_RAMPS, dt, life, 2, 4, 1, 0.0)) == [0, 0, 0, 0]


def test_emitter_weight_profile():
    """1 - smoothstep(.5, 1, r): full inside half the radius, half at r = .75, zero at and
    past the surface; boxes use the max norm."""
    pos, size = (0.0, 0.2, 0.0), (0.3, 0.3, 0.3)
    assert oracle.emitter_weight(pos, size, 0, pos) == 1.0
    assert oracle.emitter_weight(pos, size, 0, (0.15, 0.2, 0.0)) == 1.0
    assert oracle.emitter_weight(pos, size, 0, (0.225, 0.2, 0.0)) == pytest.approx(0.5)
    assert oracle.emitter_weight(pos, size, 0, (0.3, 0.2, 0.0)) == 0.0
    r = math.hypot(0.2, 0.2) / 0.3
    t = (r - 0.5) / 0.5
    assert oracle.emitter_weight(pos, size, 0, (0.2, 0.4, 0.0)) == pytest.approx(1 - t * t * (3 - 2 * t))
    # a box reaches its corners
    assert oracle.emitter_weight(pos, size, 1, (0.2, 0.4, 0.2)) > 0.0
    # the 0.03 m floor on the size
    assert oracle.emitter_weight(pos, (0.0, 0.0, 0.0), 0, (0.02, 0.2, 0.0)) > 0.0


def test_burst_perlin_vanishes_on_lattice():
    """Gradient noise is zero at every integer lattice point, for any seed."""
    for seed in (0, 1, 0x3F800000, 12345):
        for p in ((0, 0, 0), (3, -2, 7), (-5, 4, 1)):
            assert oracle.burst_perlin(seed, *p) == 0.0
        v = [oracle.burst_perlin(seed, x + 0.37, 0.61, -0.2) for x in range(8)]
        assert max(abs(x) for x in v) > 0.01


def test_curl_noise_table_is_divergence_free():
    """The periodic table is the curl of three noise fields: its central-difference
    divergence is small next to its derivatives (about 0.19 at four samples per noise
    lattice cell, from the central differences alone); swapping two components gives 1.5."""
    sim = oracle.OracleSimulation(0.05, 2, 16)
    t = sim.noise()[..., :3].astype(np.float64)  # [z, y, x, c]
    d = lambda c, axis: (np.roll(t[..., c], -1, axis) - np.roll(t[..., c], 1, axis)) / 2
    div = d(0, 2) + d(1, 1) + d(2, 0)
    grad = np.sqrt(np.mean([np.mean(d(c, a) ** 2) for c in range(3) for a in range(3)]))
    assert np.sqrt(np.mean(div ** 2)) < 0.3 * grad


def test_pcg32_known_answer():
    """pcg32_srandom_r(42, 54) yields 0xa15c02b7, 0x7b47f409, ... (pcg32-demo)."""
    rng = scenes.PCG32(42, 54)
    assert [rng.next_u32() for _ in range(6)] == [
        0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def _run(scene, bits=32, steps=None):
    sim = oracle.OracleSimulation.create(scene.grid(), bits)
    for frame, g in scene.frames(steps):
        sim.apply_grid(g)
        sim.step(frame)
    return sim


def test_pool_layout_constants():
    """types.ts arithmetic at library defaults: 34 slots a side and 136 initial slots for
    16-cell bricks at velocityDivisor 2; 68 a side for 8-cell bricks."""
    assert oracle.OracleSimulation(0.05, 2, 16).stats().pool_side == 34
    assert oracle.OracleSimulation(0.05, 2, 16).stats().pool_slots == 136
    assert oracle.OracleSimulation(0.025, 2, 8).stats().pool_side == 68


@pytest.mark.parametrize("name", ["small", "explosions"])
def test_allocation_invariants(name):
    """Slots: each owned by at most one brick; the free stack holds exactly the rest; the
    active list is the set of bricks with pages, in id order; fresh entries carry the
    FRESH flag exactly where the zeroed list says so; records hold the neighbors' entries."""
    sc = scenes.SCENES[name](steps=12 if name == "small" else 30)
    sim = oracle.OracleSimulation.create(sc.grid(), 32)
    for k, (frame, g) in enumerate(sc.frames()):
        sim.apply_grid(g)
        sim.step(frame)
        t = sim.read_tables()
        side = sim.stats().pool_side
        pages = t["pages"]
        owned = np.nonzero(pages >= 0)[0]
        assert np.array_equal(t["active"], owned)
        p = pages[owned].astype(np.int64)
        slots = ((p >> 16) * side + ((p >> 8) & 255)) * side + (p & 255)
        assert len(np.unique(slots)) == len(slots)
        free = t["free_slots"][: t["free_count"]].astype(np.int64)
        assert len(set(free.tolist()) | set(slots.tolist())) == t["pool_slots"]
        assert not (set(free.tolist()) & set(slots.tolist()))
        fresh = {int(i): int(pl) & 0xFFFFFF for i, pl in t["zeroed"] if int(pl) & FRESH_SLOT}
        for i in owned:
            e = int(t["entries"][i])
            assert (e & FRESH_ENTRY != 0) == (int(i) in fresh)
            assert e & 0xFFFFFF == pages[i]
        # records: 27 neighbor entries, by brute force over brick coordinates
        coords = sim.brick_coords(np.arange(t["brick_count"], dtype=np.int32))
        rec, _ = sim.tile_tables()
        live = rec[:, 3] != 0
        where = {}
        for i in range(t["brick_count"]):
            if live[i // 64]:
                where[tuple(coords[i])] = i
        for i in owned[:50]:
            c = coords[i]
            for kk in range(27):
                n = (c[0] + kk % 3 - 1, c[1] + (kk // 3) % 3 - 1, c[2] + kk // 9 - 1)
                want = t["entries"][where[n]] if n in where else -1
                assert t["records"][i, kk] == want


def test_tile_hash_and_records():
    """Every live tile is found by linear probing from its hash (TileDirectory.ts:156-167),
    and each record lists exactly its existing 26-neighbors."""
    sim = _run(scenes.explosions(steps=20))
    rec, hsh = sim.tile_tables()
    size = len(hsh)
    live = {tuple(r[:3]): s for s, r in enumerate(rec) if r[3] != 0}
    assert len(live) * 2 <= size or size == 256
    for c, slot in live.items():
        h = ((c[0] * 73856093) & 0xFFFFFFFF) ^ ((c[1] * 19349663) & 0xFFFFFFFF) ^ ((c[2] * 83492791) & 0xFFFFFFFF)
        h &= size - 1
        while tuple(hsh[h, :3]) != c:
            assert hsh[h, 3] >= 0, f"tile {c} not reachable"
            h = (h + 1) & (size - 1)
        assert hsh[h, 3] == slot
        for k in range(27):
            n = (c[0] + k % 3 - 1, c[1] + (k // 3) % 3 - 1, c[2] + k // 9 - 1)
            assert rec[slot, 4 + k] == live.get(n, -1)


def test_zero_velocity_step_is_exact_decay():
    """With still air, no buoyancy and no flame, transport is the identity and each
    smoke cell decays by exactly exp(-dissipation dt) (FP64 restatement)."""
    sc = scenes.campfire(steps=1)
    sc.sim.configure({"motion": {"buoyancy": 0, "smokeWeight": 0, "vorticity": 0, "damping": 0},
                      "fuel": {"enabled": False}})
    em = sc.sim._emitters[0]
    em.emission = {"flame": 0, "heatRate": 0, "smokeRate": 0, "fuelRate": 0}
    em.radius = 0.9
    sim = oracle.OracleSimulation.create(sc.grid(), 64)
    frame, g = next(sc.frames())
    sim.apply_grid(g)
    sim.step(frame)
    st = sim.read_state(dtype=np.float64)
    n, cells = st["fields"].shape[:2]
    rng = np.random.default_rng(5)
    f = np.zeros((n, cells, 4), np.float32)
    f[..., 0] = rng.uniform(0.2, 0.8, (n, cells))
    sim.write_state(f, np.zeros_like(st["velocity"], dtype=np.float32))
    sim.step(frame)
    out = sim.read_state(velocity=True, dtype=np.float64)
    assert np.array_equal(out["ids"], st["ids"])
    # The restatement takes its parameters as the GPU receives them: 32-bit floats.
    expect = f[..., 0].astype(np.float64) * math.exp(-float(np.float32(0.35)) * float(np.float32(1 / 60)))
    assert np.max(np.abs(out["fields"][..., 0] - expect) / expect) < 1e-12
    assert np.max(np.abs(out["velocity"])) == 0.0


def _dense(coords, arr, b):
    """Per-brick cell arrays [brick, comps, cells] onto a dense [comps, z, y, x] grid, NaN
    where no brick is active."""
    lo, hi = coords.min(0), coords.max(0)
    shape = (hi - lo + 1) * b
    comps = arr.shape[1]
    grid = np.full((comps, shape[2], shape[1], shape[0]), np.nan)
    for i, c in enumerate(coords):
        o = (c - lo) * b
        grid[:, o[2]:o[2] + b, o[1]:o[1] + b, o[0]:o[0] + b] = arr[i].reshape(comps, b, b, b)
    return grid


def test_projection_reduces_divergence():
    """One cold V-cycle and the projection leave the velocity far closer to divergence-free
    than the forced field the solve started from (rhs / h^2 with no expansion). A wrong
    gradient sign, restriction scale or prolongation weight fails this."""
    sc = scenes.small(steps=6, collider=False, burst=False)
    sc.sim.configure({"flame": {"expansionRate": 0}})
    sim = oracle.OracleSimulation.create(sc.grid(), 64)
    for frame, g in sc.frames():
        sim.apply_grid(g)
        sim.step(frame)
    st = sim.read_state(velocity=True, dtype=np.float64)
    coords = sim.brick_coords(st["ids"])
    b = round(st["velocity"].shape[2] ** (1 / 3))
    h = 0.1
    v = _dense(coords, st["velocity"], b)
    before = _dense(coords, sim.rhs()[:, None, :], b)[0] / h**2
    div = np.full(v.shape[1:], np.nan)
    div[1:, 1:, 1:] = ((v[0, 1:, 1:, 1:] - v[0, 1:, 1:, :-1]) + (v[1, 1:, 1:, 1:] - v[1, 1:, :-1, 1:])
                       + (v[2, 1:, 1:, 1:] - v[2, :-1, 1:, 1:])) / h
    # cells two or more away from the sparse boundary (the ambient ring is open)
    inner = np.isfinite(div)
    for axis in range(3):
        for shift in (-2, -1, 1, 2):
            inner &= np.isfinite(np.roll(div, shift, axis))
    assert inner.sum() > 500
    rms_after = np.sqrt(np.mean(div[inner] ** 2))
    rms_before = np.sqrt(np.mean(before[inner] ** 2))
    assert rms_before > 0
    assert rms_after < 0.5 * rms_before


def test_fp32_restatement_tracks_fp64():
    """The float twin stays within tolerance of the FP64 restatement over 20 campfire
    steps (tables identical, fields within 1e-4 relative L2)."""
    sc32, sc64 = scenes.campfire(steps=20), scenes.campfire(steps=20)
    a, b = _run(sc32, 32), _run(sc64, 64)
    ta, tb = a.read_tables(), b.read_tables()
    assert np.array_equal(ta["pages"], tb["pages"])
    sa, sb = a.read_state(pressure=True), b.read_state(pressure=True, dtype=np.float64)
    for k in ("fields", "velocity", "pressure"):
        x, y = sa[k].astype(np.float64), sb[k]
        assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(y), 1e-30), k


def test_storage_rounding_matches_library_conversions():
    """The mixed-precision multigrid's storage rounding equals numpy's binary16 and
    torch's bfloat16 conversions (both round to nearest even), subnormals included."""
    import torch
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.normal(0, 1, 3000) * 10.0 ** rng.integers(-9, 5, 3000),
                        np.array([0.0, -0.0, 1.0, 65504.0, 65519.0, 6.1e-5, 5.96e-8, 2.98e-8, 1.5e-8])]).astype(np.float32)
    f16 = np.array([oracle.round16(3, v) for v in x], np.float32)
    np.testing.assert_array_equal(f16, x.astype(np.float16).astype(np.float32))
    bf = np.array([oracle.round16(2, v) for v in x], np.float32)
    np.testing.assert_array_equal(bf, torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def test_defect_correction_converges_and_projects():
    """Defect correction in FP64 reaches a 1e-10 relative residual, and its projection
    leaves the velocity far closer to divergence-free than one V-cycle does."""
    def divergence_rms(mode):
        sc = scenes.small(steps=6, collider=False, burst=False)
        sc.sim.configure({"flame": {"expansionRate": 0}})
        sim = oracle.OracleSimulation.create(sc.grid(), 64)
        if mode:
            sim.set_pressure(1, 1e-10, 60)
        for frame, g in sc.frames():
            sim.apply_grid(g)
            sim.step(frame)
        report = sim.pressure_report()
        st = sim.read_state(velocity=True, dtype=np.float64)
        coords = sim.brick_coords(st["ids"])
        b = round(st["velocity"].shape[2] ** (1 / 3))
        v = _dense(coords, st["velocity"], b)
        div = np.full(v.shape[1:], np.nan)
        div[1:, 1:, 1:] = ((v[0, 1:, 1:, 1:] - v[0, 1:, 1:, :-1]) + (v[1, 1:, 1:, 1:] - v[1, 1:, :-1, 1:])
                           + (v[2, 1:, 1:, 1:] - v[2, :-1, 1:, 1:])) / 0.1
        inner = np.isfinite(div)
        for axis in range(3):
            for shift in (-2, -1, 1, 2):
                inner &= np.isfinite(np.roll(div, shift, axis))
        return np.sqrt(np.mean(div[inner] ** 2)), report
    one, _ = divergence_rms(False)
    converged, (cycles, residual) = divergence_rms(True)
    assert residual <= 1e-10 and 1 < cycles < 60, (cycles, residual)
    assert converged < 0.1 * one, (converged, one)


@pytest.mark.parametrize("mode", [2, 3])
def test_half_precision_levels_converge(mode):
    """V-cycle levels stored in bfloat16 / binary16 still drive the FP32 outer loop to a
    1e-6 relative residual."""
    sc = scenes.campfire(steps=12)
    sim = oracle.OracleSimulation.create(sc.grid(), 32)
    sim.set_pressure(mode, 1e-6, 40)
    worst = 0.0
    for frame, g in sc.frames():
        sim.apply_grid(g)
        sim.step(frame)
        cycles, residual = sim.pressure_report()
        worst = max(worst, residual)
    assert worst <= 1e-6, worst


def test_compensated_solution_passes_the_fp32_floor():
    """Defect correction keeps x as a pair of FP32 values hi + lo. With a single FP32 x
    these scenes' relative residuals level off between 1e-6 and 1e-4 (the unit roundoff
    times the operator's condition number); with the pair, FP32 arithmetic drives them
    below 1e-8. A dropped low part would stall at the single-FP32 level."""
    for seed, kw in [(1, {}), (2, {"collider": False})]:
        sc = scenes.small(seed=seed, steps=6, brick=8, divisor=2, **kw)
        sim = oracle.OracleSimulation.create(sc.grid(), 32)
        sim.set_pressure(1, 1e-8, 40)
        worst = 0.0
        for frame, g in sc.frames():
            sim.apply_grid(g)
            sim.step(frame)
            cycles, residual = sim.pressure_report()
            worst = max(worst, residual)
        assert worst <= 1e-8, (seed, worst)
