"""Pins the CPU restatement's solver stages with facts it does not restate.

Each test sets a state the mathematics answers for (a linear or random field under a
uniform velocity, a constant or linear buoyant field, a force at hand-picked points, a
linear or quadratic velocity field, a burning fuel cell), runs one step of the FP64
restatement, and compares the stage's output with that answer:

* scalar transport (fluid-scalar-transport.wgsl:6-198): a linear field under uniform
  velocity is carried exactly; a random field matches a textbook first-order
  semi-Lagrangian step and the bounded MacCormack step (Selle et al. 2008, with the
  reference's fall-back to the first-order value), both computed with scipy;
* velocity transport (fluid-velocity-transport.wgsl:28-179): a face component that varies
  linearly along the flow is carried exactly, on the right face positions;
* buoyancy and its face averaging (fluid-forces.wgsl:79-120): dt (buoyancy heat -
  smokeWeight smoke) on each vertical face, at the face's own height;
* wind, radial and vortex forces (force-fields.wgsl:60-103) at points whose values follow
  by hand;
* curl (fluid-forces.wgsl:47-75) of a linear velocity field, and vorticity confinement on
  a vortex whose vorticity grows linearly across it;
* the fuel burn (fluid-field-evolution.wgsl:41-51): remaining fuel times exp(-burnRate dt)
  above the ignition heat, untouched below;
* slot grants (fluid-allocation.wgsl:179-230, fluid-pool-growth.wgsl:8-21) against a
  brute-force model of the free stack;
* the burst noise (fluid-sources.wgsl:40-66) against a separate evaluation of its
  definition, and its gradients at lattice points.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle
from firepro import scenes
from firepro.abi import FRESH_SLOT
from firepro.library import FireSimulation

H = 0.05          # field cell (m)
HV = 0.1          # velocity cell (m): velocityDivisor 2
B, BV = 16, 8     # brick edges in field and velocity cells
DT = 1 / 60
# The restatement takes its parameters as the device receives them, in single precision.
f32 = lambda x: float(np.float32(x))  # noqa: E731
H32, HV32, DT32 = f32(H), f32(HV), f32(DT)
QUIET = {"motion": {"buoyancy": 0, "smokeWeight": 0, "vorticity": 0, "damping": 0},
         "smoke": {"dissipation": 0}, "flame": {"cooling": 0}, "fuel": {"enabled": False}}


def _merge(a, b):
    out = {k: dict(v) for k, v in a.items()}
    for k, v in b.items():
        out.setdefault(k, {}).update(v)
    return out


class Harness:
    """One FP64 restatement around a still, empty emitter (it only keeps bricks active):
    step once to allocate, write a state, step again and read the stages."""

    def __init__(self, settings=None, forces=(), maccormack=True, radius=1.0):
        self.lib = FireSimulation({"voxelSize": H, "brickSize": B, "velocityDivisor": 2, "seed": 1,
                                   "scalarMacCormack": maccormack})
        self.lib.configure(_merge(QUIET, settings or {}))
        self.lib.add_emitter({"position": (0.0, 1.6, 0.0), "shape": {"radius": radius}})
        for f in forces:
            self.lib.add_force(f)
        o = self.lib.options
        grid = {"voxelSize": o["voxelSize"], "velocityDivisor": 2, "brickSize": B, "smokeDivisor": 1,
                "scalarMacCormack": maccormack}
        self.sim = oracle.OracleSimulation.create(grid, 64)
        self.sim.apply_grid(self.lib.grid_settings())
        self.frame = self.lib.build_frame()
        self.sim.step(self.frame)
        st = self.sim.read_state(dtype=np.float64)
        self.ids = st["ids"]
        self.coords = self.sim.brick_coords(self.ids).astype(np.int64)
        assert len(self.ids) >= 8

    def field_centers(self):
        """World centers of every field cell, [brick, cell, 3] (x fastest)."""
        l = np.arange(B ** 3)
        local = np.stack([l % B, (l // B) % B, l // (B * B)], 1)
        return ((self.coords[:, None, :] * B + local[None]) + 0.5) * H

    def velocity_cells(self):
        """Integer velocity-cell coordinates, [brick, cell, 3]."""
        l = np.arange(BV ** 3)
        local = np.stack([l % BV, (l // BV) % BV, l // (BV * BV)], 1)
        return self.coords[:, None, :] * BV + local[None]

    def step(self, fields=None, velocity=None, dt=DT):
        n = len(self.ids)
        f = np.zeros((n, B ** 3, 4), np.float32) if fields is None else fields
        v = np.zeros((n, 3, BV ** 3), np.float32) if velocity is None else velocity
        self.sim.write_state(f.astype(np.float32), v.astype(np.float32))
        self.sim.step(self.frame, dt)
        out = self.sim.read_state(velocity=True, expansion=True, dtype=np.float64)
        assert np.array_equal(out["ids"], self.ids)
        return out

    def interior(self, cells, edge, margin):
        """Mask over [brick, cell] of cells at least `margin` cells (of a grid `edge` per
        brick) inside the active bricks and above the ground."""
        lo = self.coords.min(0)
        dims = (self.coords.max(0) - lo + 1) * edge
        active = np.zeros(dims[::-1], bool)
        c = cells - lo * edge
        active[c[..., 2], c[..., 1], c[..., 0]] = True
        inner = ndimage.binary_erosion(active, np.ones((3, 3, 3), bool), iterations=margin)
        return inner[c[..., 2], c[..., 1], c[..., 0]] & (cells[..., 1] >= margin)


def _field_cells(h):
    l = np.arange(B ** 3)
    local = np.stack([l % B, (l // B) % B, l // (B * B)], 1)
    return h.coords[:, None, :] * B + local[None]


def _dense(h, values, cells, edge):
    """Per-cell values [brick, cell] onto a dense [z, y, x] array (zero where no brick is)."""
    lo = h.coords.min(0)
    dims = (h.coords.max(0) - lo + 1) * edge
    grid = np.zeros(dims[::-1])
    c = cells - lo * edge
    grid[c[..., 2], c[..., 1], c[..., 0]] = values
    return grid, c


U = np.array([0.9, -0.6, 0.45], np.float32).astype(np.float64)  # m/s, as the state stores it


def _uniform_velocity(h):
    v = np.zeros((len(h.ids), 3, BV ** 3), np.float32)
    for a in range(3):
        v[:, a, :] = U[a]
    return v


@pytest.mark.parametrize("maccormack", [True, False])
def test_linear_scalar_field_carried_exactly(maccormack):
    """Trilinear backtraces reproduce a linear field: after one step under uniform
    velocity u, smoke and heat at x equal their values at x - u dt (both transports). A
    wrong backtrace sign, a half-cell offset or a missing dt would be off by about
    1e-3 here."""
    h = Harness(maccormack=maccormack)
    x = h.field_centers()
    smoke = lambda p: 0.5 + 0.05 * p[..., 0] + 0.04 * p[..., 1] - 0.03 * p[..., 2]  # noqa: E731
    heat = lambda p: 0.3 + 0.02 * p[..., 0] - 0.01 * p[..., 1] + 0.03 * p[..., 2]  # noqa: E731
    f = np.zeros((len(h.ids), B ** 3, 4))
    f[..., 0], f[..., 1] = smoke(x), heat(x)
    out = h.step(f.astype(np.float32), _uniform_velocity(h))
    inner = h.interior(_field_cells(h), B, 6)
    assert inner.sum() > 2000
    dep = x - U * DT
    assert np.abs(out["fields"][..., 0] - smoke(dep))[inner].max() < 2e-7
    assert np.abs(out["fields"][..., 1] - heat(dep))[inner].max() < 2e-7


def _trilinear(grid, coords):
    return ndimage.map_coordinates(grid, coords, order=1, mode="constant", cval=0.0)


@pytest.mark.parametrize("maccormack", [True, False])
def test_random_scalar_field_matches_textbook_step(maccormack):
    """A random smoke field under uniform velocity: first order is the trilinear sample at
    the departure point x - u dt; MacCormack adds half the round-trip error, phi_hat +
    (phi - phi_tilde) / 2 with phi_tilde the forward sample of phi_hat, and falls back to
    phi_hat where that leaves the range of the eight cells phi_hat was interpolated from.
    Both are computed here on the dense grid with scipy."""
    h = Harness(maccormack=maccormack)
    rng = np.random.default_rng(3)
    f = np.zeros((len(h.ids), B ** 3, 4), np.float32)
    f[..., 0] = rng.uniform(0.2, 0.8, f.shape[:2])
    out = h.step(f, _uniform_velocity(h))
    cells = _field_cells(h)
    phi, c = _dense(h, f[..., 0].astype(np.float64), cells, B)
    s = U * DT32 / H32  # the shift in cells per axis
    zyx = np.indices(phi.shape).astype(np.float64)
    back = zyx - s[::-1, None, None, None]
    hat = _trilinear(phi, back.reshape(3, -1)).reshape(phi.shape)
    if maccormack:
        tilde = _trilinear(hat, (zyx + s[::-1, None, None, None]).reshape(3, -1)).reshape(phi.shape)
        corr = hat + 0.5 * (phi - tilde)
        base = np.floor(back).astype(int)
        lo = np.full(phi.shape, np.inf)
        hi = np.full(phi.shape, -np.inf)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    idx = [np.clip(base[0] + dz, 0, phi.shape[0] - 1), np.clip(base[1] + dy, 0, phi.shape[1] - 1),
                           np.clip(base[2] + dx, 0, phi.shape[2] - 1)]
                    lo, hi = np.minimum(lo, phi[tuple(idx)]), np.maximum(hi, phi[tuple(idx)])
        outside = (corr < lo) | (corr > hi)
        want = np.where(outside, hat, corr)
        ambiguous = np.minimum(np.abs(corr - lo), np.abs(corr - hi)) < 1e-9
    else:
        want, ambiguous, outside = hat, np.zeros(phi.shape, bool), None
    got = out["fields"][..., 0]
    inner = h.interior(cells, B, 6)
    want_c = want[c[..., 2], c[..., 1], c[..., 0]]
    ok = inner & ~ambiguous[c[..., 2], c[..., 1], c[..., 0]]
    assert ok.sum() > 2000
    assert np.abs(got - want_c)[ok].max() < 1e-12
    if maccormack:
        # The limiter matters here: a sizable share of cells fall back to first order,
        # and the rest differ from first order.
        share = outside[c[..., 2], c[..., 1], c[..., 0]][ok].mean()
        assert 0.005 < share < 0.95, share
        assert np.abs(got - hat[c[..., 2], c[..., 1], c[..., 0]])[ok].max() > 1e-3
        # The recorded limiter decisions (the precision report's flip counts read them): the
        # limiter ran, and kept the corrected smoke exactly where it stays inside its donors.
        lim = h.sim.stage("limiter")
        evaluated = (lim & 16) != 0
        kept = (lim & 1) != 0
        assert evaluated[ok].all()
        assert np.array_equal(kept[ok], ~outside[c[..., 2], c[..., 1], c[..., 0]][ok])


def test_maccormack_makes_no_new_extrema():
    """Bounded MacCormack never leaves the range of its donors, so a sharp step in smoke
    (0.2 | 0.8) carried by uniform velocity stays inside [0.2, 0.8]; the unbounded
    correction would overshoot at the step."""
    h = Harness(maccormack=True)
    x = h.field_centers()
    f = np.zeros((len(h.ids), B ** 3, 4), np.float32)
    f[..., 0] = np.where(x[..., 0] + 0.5 * x[..., 2] > 0.1, 0.8, 0.2)
    out = h.step(f, _uniform_velocity(h))
    inner = h.interior(_field_cells(h), B, 6)
    s = out["fields"][..., 0][inner]
    lo, hi = f32(0.2), f32(0.8)  # the state is stored in single precision
    assert s.min() >= lo and s.max() <= hi
    assert ((s > lo + 1e-3) & (s < hi - 1e-3)).any(), "the step is smeared, not frozen"


def test_velocity_component_carried_along_the_flow():
    """u = (U, c x + e z, W): the vertical component varies linearly along the flow, so one
    step carries it exactly to c (x - U dt) + e (z - W dt) on each vertical face (whose
    x and z are its cell's center), in the advected and the corrected velocity alike,
    while the uniform components stay put. A face offset on the wrong axis moves x or z by
    half a cell."""
    h = Harness()
    cells = h.velocity_cells()
    xc, zc = (cells[..., 0] + 0.5) * HV, (cells[..., 2] + 0.5) * HV
    c, e = 0.8, -0.5
    v = np.zeros((len(h.ids), 3, BV ** 3))
    v[:, 0], v[:, 1], v[:, 2] = U[0], c * xc + e * zc, U[2]
    h.step(None, v.astype(np.float32))
    inner = h.interior(cells, BV, 3)
    assert inner.sum() > 300
    want = c * (xc - U[0] * DT) + e * (zc - U[2] * DT)
    for name in ("advected_velocity", "corrected_velocity"):
        got = h.sim.stage(name)
        assert np.abs(got[:, 1] - want)[inner].max() < 1e-6, name
        assert np.abs(got[:, 0] - U[0])[inner].max() < 1e-6, name
        assert np.abs(got[:, 2] - U[2])[inner].max() < 1e-6, name


def test_buoyancy_on_each_vertical_face():
    """Still air, heat rising linearly with height, constant smoke: each vertical face
    gains dt (buoyancy heat(face height) - smokeWeight smoke) in one step (the cell forces
    are averaged onto the face between them), horizontal faces gain nothing; the cell
    force itself is the same expression at the cell's center."""
    b, w = 1.6, 0.12
    h = Harness({"motion": {"buoyancy": b, "smokeWeight": w}})
    x = h.field_centers()
    f = np.zeros((len(h.ids), B ** 3, 4))
    f[..., 0] = 0.2
    f[..., 1] = 0.4 + 0.3 * x[..., 1]
    h.step(f.astype(np.float32), None)
    cells = h.velocity_cells()
    inner = h.interior(cells, BV, 2)
    assert inner.sum() > 300
    forced = h.sim.stage("forced_velocity")
    y_face = (cells[..., 1] + 1) * HV
    want = DT * (b * (0.4 + 0.3 * y_face) - w * 0.2)
    assert np.abs(forced[:, 1] - want)[inner].max() < 1e-7
    assert np.abs(forced[:, 0])[inner].max() < 1e-12 and np.abs(forced[:, 2])[inner].max() < 1e-12
    probe = cells[inner][:50]
    got = h.sim.force_probe(probe)
    yc = (probe[:, 1] + 0.5) * HV
    assert np.abs(got[:, 1] - (b * (0.4 + 0.3 * yc) - w * 0.2)).max() < 1e-7
    assert np.abs(got[:, [0, 2]]).max() < 1e-12


def _probe_force(force, offsets):
    """The cell force of `force` at velocity cells whose centers sit at `offsets` from the
    force's center (the center is placed relative to a fixed cell)."""
    p = np.array([3, 20, 2])
    center0 = (p + 0.5) * HV
    out = []
    for off in offsets:
        f = dict(force)
        if "center" in f:
            f["center"] = tuple(center0 - np.asarray(off, float))
        h = Harness(forces=[f])
        h.step()
        out.append(h.sim.force_probe([p])[0])
    return np.array(out)


def test_wind_radial_and_vortex_forces_at_hand_points():
    """Wind: strength along its direction. Radial (strength s, radius r): at distance r
    half of s outward, at 2r a fifth, inside r/4 the direction is scaled by the distance
    over r/4. Vortex about +y (strength s, lift L, inward I, radius r): at +x r, the swirl
    -s/2 along z, -I/2 along x and L/2 along y; at +z r (any height) the swirl is +s/2
    along x."""
    got = _probe_force({"type": "wind", "direction": (0, 0, 2), "strength": 2.5}, [(0, 0, 0)])
    np.testing.assert_allclose(got[0], (0, 0, 2.5), atol=1e-6)
    s, r = 3.0, 0.8
    got = _probe_force({"type": "radial", "center": (0, 0, 0), "strength": s, "radius": r},
                       [(r, 0, 0), (0, 0, -2 * r), (0.1 * r, 0, 0)])
    np.testing.assert_allclose(got[0], (s / 2, 0, 0), atol=1e-6)
    np.testing.assert_allclose(got[1], (0, 0, -s / 5), atol=1e-6)
    np.testing.assert_allclose(got[2], (0.4 * s / 1.01, 0, 0), atol=1e-5)  # the center is float32
    L, I = 1.5, 0.7
    got = _probe_force({"type": "vortex", "center": (0, 0, 0), "axis": (0, 1, 0), "strength": s, "radius": r,
                        "lift": L, "inward": I}, [(r, 0, 0), (0, 0.9, r)])
    np.testing.assert_allclose(got[0], (-I / 2, L / 2, -s / 2), atol=1e-6)
    np.testing.assert_allclose(got[1], (s / 2, L / 2, -I / 2), atol=1e-6)


def test_curl_of_linear_velocity():
    """The curl of u = A x (each component on its own faces) is the constant
    (A21 - A12, A02 - A20, A10 - A01), and its fourth channel the magnitude. A tiny step
    keeps the corrected velocity (the curl's input) within 1e-5 of u."""
    A = np.array([[0.1, 0.5, -0.3], [0.2, -0.2, 0.4], [-0.6, 0.3, 0.1]])
    h = Harness({"motion": {"vorticity": 1.0}})
    cells = h.velocity_cells().astype(np.float64)
    v = np.zeros((len(h.ids), 3, BV ** 3))
    for a in range(3):
        face = (cells + 0.5) * HV
        face[..., a] += 0.5 * HV
        v[:, a] = face @ A[a]
    h.step(None, v.astype(np.float32), dt=1e-5)
    curl = h.sim.stage("curl")
    inner = h.interior(h.velocity_cells(), BV, 3)
    assert inner.sum() > 300
    want = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.abs(curl[..., :3][inner] - want).max() < 1e-4
    assert np.abs(curl[..., 3][inner] - np.linalg.norm(want)).max() < 1e-4


def test_confinement_on_a_sheared_vortex():
    """u_y = c x + g x^2 / 2 (nothing else) is steady under its own transport and has
    vorticity (0, 0, c + g x) at cell centers, growing along +x. Confinement pushes along
    N x omega with N the unit gradient of |omega| (here +x, regularized by 1e-4 as the
    reference does): the cell force is (0, -eps h (c + g x) 2gh / (2gh + 1e-4), 0)."""
    c, g, eps = 2.0, 0.5, 3.2
    h = Harness({"motion": {"vorticity": eps}})
    cells = h.velocity_cells()
    xc = (cells[..., 0] + 0.5) * HV
    v = np.zeros((len(h.ids), 3, BV ** 3))
    v[:, 1] = c * xc + 0.5 * g * xc ** 2
    h.step(None, v.astype(np.float32))
    inner = h.interior(cells, BV, 3)
    assert inner.sum() > 300
    # Single-precision inputs: the quadratic is exact only to float32 rounding.
    curl = h.sim.stage("curl")
    np.testing.assert_allclose(curl[..., 2][inner], (c + g * xc)[inner], atol=2e-5)
    probe = cells[inner][::7][:60]
    got = h.sim.force_probe(probe)
    x = (probe[:, 0] + 0.5) * HV
    n = 2 * g * HV / (2 * g * HV + 1e-4)
    np.testing.assert_allclose(got[:, 1], -eps * HV * (c + g * x) * n, rtol=1e-4, atol=1e-6)
    assert np.abs(got[:, [0, 2]]).max() < 1e-4


def test_fuel_burns_by_exp_of_rate_above_ignition():
    """Fuel 0.8 in still air: where the heat is at or above the ignition heat 0.5 the
    fuel left after a step is 0.8 exp(-burnRate dt) and the flame lifetime starts at
    min(fuel, 1) less one step's decrement dt / lifespan; below it nothing burns."""
    rate, life = 4.0, 1.2
    h = Harness({"fuel": {"enabled": True, "ignitionHeat": 0.5, "burnRate": rate}})
    x = h.field_centers()
    f = np.zeros((len(h.ids), B ** 3, 4))
    hot = x[..., 0] < 0
    f[..., 1] = np.where(hot, 1.0, 0.2)
    f[..., 3] = 0.8
    out = h.step(f.astype(np.float32), None)
    inner = h.interior(_field_cells(h), B, 2)
    fuel, lifetime = out["fields"][..., 3], out["fields"][..., 2]
    f08 = float(np.float32(0.8))
    keep = math.exp(-float(np.float32(rate)) * float(np.float32(DT)))
    assert np.abs(fuel[hot & inner] - f08 * keep).max() < 1e-12
    assert np.abs(fuel[~hot & inner] - f08).max() < 1e-12
    assert np.abs(lifetime[hot & inner] - (f08 - float(np.float32(DT)) / float(np.float32(life)))).max() < 1e-6
    assert np.abs(lifetime[~hot & inner]).max() == 0.0


class FreeStack:
    """Brute-force model of the slot allocator: a stack of free slots with `free` live
    entries. Per step: a pool that grows by k slots puts the new slots, highest first,
    below the live entries; freed slots are pushed at free + rank in brick-id order; then
    the granted bricks, in brick-id order, pop from the top (the brick of rank r takes
    entry free - 1 - r)."""

    def __init__(self, stack, free):
        self.stack, self.free = list(stack), free

    def grow(self, old, new):
        self.stack = list(range(new - 1, old - 1, -1)) + self.stack
        self.free += new - old

    def step(self, freed, granted):
        live = self.stack[: self.free] + list(freed)
        self.stack[: len(live)] = live
        top = len(live)
        out = {b: live[top - 1 - r] for r, b in enumerate(granted)}
        self.free = top - len(granted)
        return out


@pytest.mark.parametrize("name,budget", [("explosions", None), ("small", 4096 * 3)])
def test_slot_grants_follow_the_free_stack_model(name, budget):
    """Every step's pages and free stack equal the brute-force model's, through pool
    growth; with a small voxel budget the failures are the requests the available slots
    (free minus those held back by the budget) cannot serve."""
    sc = scenes.SCENES[name](steps=40 if name == "explosions" else 10)
    if budget:
        sc.sim.options["grid"]["maxVoxels"] = budget
    sim = oracle.OracleSimulation.create(sc.grid(), 32)
    prev = None
    model = None
    cells = sc.grid()["brickSize"] ** 3
    for k, (frame, g) in enumerate(sc.frames()):
        sim.apply_grid(g)
        sim.step(frame)
        t = sim.read_tables()
        side = sim.stats().pool_side
        slot_of = lambda p: ((p >> 16) * side + ((p >> 8) & 255)) * side + (p & 255)  # noqa: E731
        if prev is None:
            # The first step's stack starts full: slots - 1 down to 0.
            model = FreeStack(list(range(t["pool_slots"] - 1, -1, -1)), t["pool_slots"])
            prev_pages = np.full(t["brick_count"], -1)
            prev_slots = t["pool_slots"]
        else:
            prev_pages, prev_slots = prev
        if t["pool_slots"] > prev_slots:
            model.grow(prev_slots, t["pool_slots"])
        n = min(len(prev_pages), t["brick_count"])
        pages = t["pages"]
        fresh = sorted(int(i) for i, pl in t["zeroed"] if int(pl) & FRESH_SLOT)
        freed = [int(slot_of(prev_pages[i])) for i in range(n) if prev_pages[i] >= 0 and pages[i] < 0]
        granted = model.step(freed, fresh)
        for b, slot in granted.items():
            assert slot_of(int(pages[b])) == slot, (k, b)
        assert t["free_count"] == model.free, k
        assert np.array_equal(t["free_slots"][: model.free], model.stack[: model.free]), k
        if budget:
            held = max(0, t["pool_slots"] - int(budget // cells))
            assert t["failures"] >= 0
            if t["failures"]:
                assert model.free <= held, (k, model.free, held)
        prev = (pages.copy(), t["pool_slots"])
    if budget:
        assert sim.stats().pool_limited == 1


def _burst_perlin(seed_bits, p):
    """The burst noise's definition (fluid-sources.wgsl:40-66), evaluated separately:
    a 32-bit hash of the lattice corner and the seed picks one of 12 edge gradients (4
    repeated), dotted with the offset, blended by the quintic fade."""
    m = 0xFFFFFFFF
    cell = [math.floor(c) for c in p]
    f = [c - math.floor(c) for c in p]

    def grad(cx, cy, cz, ox, oy, oz):
        hsh = ((cx & m) * 374761393 + (cy & m) * 668265263 + (cz & m) * 2246822519 + seed_bits * 3266489917) & m
        hsh = ((hsh ^ (hsh >> 13)) * 1274126177) & m
        hsh = (hsh ^ (hsh >> 16)) & 15
        a = ox if hsh < 8 else oy
        b = oy if hsh < 4 else (ox if hsh in (12, 14) else oz)
        return (a if hsh & 1 == 0 else -a) + (b if hsh & 2 == 0 else -b)

    fade = [t * t * t * (t * (t * 6 - 15) + 10) for t in f]
    corner = {}
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                corner[dx, dy, dz] = grad(cell[0] + dx, cell[1] + dy, cell[2] + dz, f[0] - dx, f[1] - dy, f[2] - dz)
    lerp = lambda a, b, t: a + (b - a) * t  # noqa: E731
    x00 = lerp(corner[0, 0, 0], corner[1, 0, 0], fade[0])
    x10 = lerp(corner[0, 1, 0], corner[1, 1, 0], fade[0])
    x01 = lerp(corner[0, 0, 1], corner[1, 0, 1], fade[0])
    x11 = lerp(corner[0, 1, 1], corner[1, 1, 1], fade[0])
    return lerp(lerp(x00, x10, fade[1]), lerp(x01, x11, fade[1]), fade[2])


def test_burst_perlin_matches_its_definition_off_lattice():
    """At random points (negative coordinates included) the restatement equals a separate
    evaluation of the definition, for several seeds; a changed hash multiplier, gradient
    choice or fade polynomial fails."""
    rng = np.random.default_rng(17)
    for seed in (0, 1, 0x3F800000, 0x40490FDB):
        for p in rng.uniform(-20, 20, (60, 3)):
            assert oracle.burst_perlin(seed, *p) == pytest.approx(_burst_perlin(seed, p), abs=1e-12)


def test_burst_perlin_gradients_are_cube_edges():
    """Near a lattice point the noise is g . (p - corner) with g one of the 12 cube-edge
    vectors (two components of +-1, one 0): its finite-difference gradient there is such a
    vector, and every edge direction turns up."""
    seen = set()
    eps = 1e-5
    for seed in (1, 7):
        for c in np.random.default_rng(seed).integers(-30, 30, (80, 3)):
            g = [(oracle.burst_perlin(seed, *(c + eps * e)) - oracle.burst_perlin(seed, *(c - eps * e))) / (2 * eps)
                 for e in np.eye(3)]
            r = np.round(g).astype(int)
            assert np.abs(np.asarray(g) - r).max() < 1e-6
            assert sorted(np.abs(r).tolist()) == [0, 1, 1], r
            seen.add(tuple(r))
    assert len(seen) == 12, seen
