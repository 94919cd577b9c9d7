"""Pins the render restatement (oracle/render.hpp) to what the mathematics fixes.

The reference holds no images or known-answer vectors for its renderer, and its WebGPU
path cannot run here, so each part is checked against closed forms, invariants and
independent computations (numpy/scipy on dense arrays):
  * fexp2 against exp2, within 2 ulp;
  * the blackbody table: peak-normalized rows, 1800 K at log2 radiance 0, monotone
    radiance, the colour order of hot and cold bodies;
  * the per-pixel jitter: uniform on [0, 1), one value per 1/60 s;
  * an empty volume renders nothing;
  * trilinear, quadratic and cubic B-spline reconstruction reproduce linear fields;
  * the render field's smoke is the separable [1,2,1]/4 filter of the fields (scipy);
  * occupancy flags equal their definition evaluated on the dense render field;
  * a homogeneous slab's opacity is 1 - exp(-sigma L);
  * incident light lies between full and the transmittance of the densest path.
"""
import numpy as np
import pytest
from scipy import ndimage, stats


def _mods():
    from firepro import abi, scenes
    from oracle import OracleSimulation
    import oracle as orc
    return abi, scenes, OracleSimulation, orc


def _run(scene):
    abi, scenes, OracleSimulation, _ = _mods()
    sim = OracleSimulation.create(scene.grid(), 32)
    for frame, g in scene.frames():
        sim.apply_grid(g)
        sim.step(frame)
    return sim


def _dense(sim, ids, values, B):
    """Scatter per-brick cell values (n, B^3, ...) into a dense array over the bricks'
    bounding box; returns (array, origin in cells, slotted mask)."""
    coords = sim.brick_coords(ids)
    lo, hi = coords.min(0), coords.max(0) + 1
    shape = tuple((hi - lo) * B)
    extra = values.shape[2:]
    out = np.zeros(shape[::-1] + extra, values.dtype)  # z, y, x
    mask = np.zeros(shape[::-1], bool)
    for i, c in enumerate(coords):
        o = (c - lo) * B
        block = values[i].reshape((B, B, B) + extra)
        out[o[2]:o[2] + B, o[1]:o[1] + B, o[0]:o[0] + B] = block
        mask[o[2]:o[2] + B, o[1]:o[1] + B, o[0]:o[0] + B] = True
    return out, lo * B, mask


def test_fexp2_within_2ulp():
    _, _, OracleSimulation, orc = _mods()
    sim = OracleSimulation(0.05, 2, 8)
    xs = np.concatenate([np.linspace(-24, 10, 4001), np.array([-24, -1, -0.5, 0, 0.5, 1, 10])]).astype(np.float32)
    got = np.array([sim.probe(3, [x])[0] for x in xs], np.float64)
    want = np.exp2(xs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - want) / ulp) <= 2.0
    assert np.all(got[xs == np.round(xs)] == want[xs == np.round(xs)])  # powers of two exact


def test_blackbody_table_invariants():
    _, _, _, orc = _mods()
    t = orc.blackbody().astype(np.float64)
    kelvin = 500 + 9500 * np.arange(512) / 511
    lit = t[:, 3] > -32
    assert lit.sum() > 500
    assert np.allclose(t[lit, :3].max(1), 1.0, atol=1e-7)
    assert np.all(np.diff(t[lit, 3]) > 0), "radiance grows with temperature"
    # log2 radiance relative to 1800 K: zero there (linear between the table's nodes).
    assert abs(np.interp(1800.0, kelvin, t[:, 3])) < 0.01
    warm = (kelvin > 1000) & (kelvin < 3000)
    assert np.all(t[warm, 0] == 1.0) and np.all(t[warm, 1] > t[warm, 2]), "warm bodies are red-orange"
    assert t[-1, 2] == 1.0 and t[-1, 0] < t[-1, 1] < 1.0, "10000 K is blue-white"
    near_white = t[np.argmin(np.abs(kelvin - 6500)), :3]
    assert near_white.min() > 0.75, near_white


def test_jitter_uniform_and_per_frame():
    _, _, OracleSimulation, _ = _mods()
    sim = OracleSimulation(0.05, 2, 8)
    vals = np.array([sim.probe(2, [x, y, 0.0])[0] for y in range(48) for x in range(48)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert stats.kstest(vals, "uniform").statistic < 0.04
    a = [sim.probe(2, [5, 7, t])[0] for t in (0.0, 0.4 / 60, 1.0 / 60, 1.5 / 60)]
    assert a[0] == a[1] and a[2] == a[3] and a[0] != a[2], "one value per 1/60 s of time"


def test_empty_volume_renders_nothing():
    abi, scenes, OracleSimulation, _ = _mods()
    sim = OracleSimulation(0.05, 2, 16)
    frame = abi.FrameData()
    sim.step(frame)
    img = sim.render(abi.camera((0, 1, 4), (0, 1, 0), 45, 24, 16))
    assert not img.any()


def _linear_state(sim, B, a=1.0, slope=(0.01, -0.02, 0.015)):
    st = sim.read_state(velocity=False)
    ids = st["ids"]
    coords = sim.brick_coords(ids)
    f = np.zeros_like(st["fields"])
    z, y, x = np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")
    for i, c in enumerate(coords):
        cx, cy, cz = c[0] * B + x, c[1] * B + y, c[2] * B + z
        f[i, :, 0] = (a + slope[0] * cx + slope[1] * cy + slope[2] * cz).reshape(-1)
    sim.write_state(f, None)
    return ids, coords


def test_reconstruction_reproduces_linear_fields():
    """Trilinear and the B-splines reproduce linear functions; the [1,2,1]/4 smoothing
    keeps them; so reconstruction at interior points is the linear function itself."""
    abi, scenes, _, _ = _mods()
    B = 16
    sim = _run(scenes.campfire(steps=40))
    a, slope = 1.0, np.array([0.01, -0.02, 0.015])
    ids, coords = _linear_state(sim, B, a, slope)
    sim.render(abi.camera((0, 1, 4), (0, 1, 0), 45, 4, 4), image=False)
    # Listed bricks: slotted and inside the box, where the render field lives.
    listed = sim.brick_coords(sim.read_render()["bricks"])
    slotted = {tuple(c) for c in listed}
    coords = listed
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(400):
        c = coords[rng.integers(len(coords))]
        p = c * B + rng.uniform(0, B, 3)
        lo, hi = np.floor(p - 4).astype(int), np.floor(p + 4).astype(int)
        ok = all(tuple(b) in slotted for b in {tuple(v // B) for v in
                 [np.array([x, y, z]) for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]})
        if not ok:
            continue
        want = a + slope @ (p - 0.5)
        for mode in (0, 1, 2):
            got = sim.probe(0, [p[0], p[1], p[2], mode])[0]
            assert abs(got - want) <= 2e-6 * abs(want) + 1e-6, (mode, p, got, want)
        checked += 1
    assert checked > 20


def test_smoothing_is_separable_121_filter_and_occupancy_definition():
    abi, scenes, _, _ = _mods()
    B = 8
    sim = _run(scenes.small(steps=6, brick=B, divisor=2))
    st = sim.read_state(velocity=False)
    ids = st["ids"]
    rng = np.random.default_rng(5)
    f = st["fields"].copy()
    f[:, :, 0] = rng.uniform(0, 1, f.shape[:2]) * (rng.uniform(0, 1, (len(ids), 1)) > 0.3)
    f[:, :, 2] = rng.uniform(0, 0.001, f.shape[:2]) * (rng.uniform(0, 1, (len(ids), 1)) > 0.8)
    sim.write_state(f, None)
    sim.render(abi.camera((0, 1, 4), (0, 1, 0), 45, 4, 4), image=False)
    rt = sim.read_render()
    # Dense original fields over the listed bricks' neighbourhood.
    dense, origin, mask = _dense(sim, ids, f, B)
    k = np.array([1.0, 2.0, 1.0]) / 4
    smooth = dense[..., 0].astype(np.float64)
    for axis in range(3):
        smooth = ndimage.correlate1d(smooth, k, axis=axis, mode="constant")
    listed = rt["bricks"]
    lcoords = sim.brick_coords(listed)
    box = rt["boxes"][0]
    box_lo, box_hi = box[:3] * B, (box[:3] + box[4:7]) * B - 1
    E = B + 1
    checked = 0
    for i, c in enumerate(lcoords):
        tex = rt["field"][i].reshape(E, E, E, 4)
        for lz in range(E):
            for ly in range(E):
                for lx in range(E):
                    cell = c * B + np.array([lx, ly, lz])
                    d = cell - origin
                    if np.any(cell - 1 < box_lo) or np.any(cell + 1 > box_hi):
                        continue  # clamped to the box's edge
                    if np.any(d < 1) or np.any(d + 1 >= np.array(dense.shape[2::-1])):
                        continue
                    want = smooth[d[2], d[1], d[0]]
                    assert abs(tex[lz, ly, lx, 0] - want) <= 1e-6 * max(1.0, abs(want)), (cell, want)
                    np.testing.assert_array_equal(tex[lz, ly, lx, 1:], dense[d[2], d[1], d[0], 1:])
                    checked += 1
    assert checked > 1000
    # Occupancy: per 8^3 block, flame (bit 0) and anything visible (bit 1) among the
    # block's cells and two around, in bricks with slots, read from their own texels.
    R = np.zeros(mask.shape + (4,), np.float32)
    for i, c in enumerate(lcoords):
        o = c * B - origin
        R[o[2]:o[2] + B, o[1]:o[1] + B, o[0]:o[0] + B] = rt["field"][i].reshape(E, E, E, 4)[:B, :B, :B]
    glow = np.clip(R[..., 2] * 5.0, 0, 1)
    vis = (R[..., 0] + glow > 0.002)
    flame = glow > 0.002
    in_box = np.zeros(mask.shape, bool)
    for c in lcoords:
        o = c * B - origin
        in_box[o[2]:o[2] + B, o[1]:o[1] + B, o[0]:o[0] + B] = True
    for i, c in enumerate(lcoords):
        o = c * B - origin - 2
        sl = tuple(slice(max(0, o[a]), max(0, o[a] + 12)) for a in (2, 1, 0))
        want = int(flame[sl][in_box[sl]].any()) | (int(vis[sl][in_box[sl]].any()) << 1)
        assert rt["occupancy"][i][0] == want, (c, rt["occupancy"][i], want)


def test_homogeneous_slab_opacity():
    """Constant smoke, no flame: the march's opacity along a ray is 1 - exp(-sigma L), with
    L the ray's length inside the slotted bricks (edges blurred by the smoothing)."""
    abi, scenes, _, _ = _mods()
    B = 16
    sim = _run(scenes.campfire(steps=40))
    st = sim.read_state(velocity=False)
    rho, density = 0.08, 2.5
    f = np.zeros_like(st["fields"])
    f[:, :, 0] = rho
    sim.write_state(f, None)
    voxel = 0.05
    sim.render(abi.camera((0, 1, 4), (0, 1, 0), 45, 4, 4), image=False)
    coords = sim.brick_coords(sim.read_render()["bricks"])
    slotted = {tuple(c) for c in coords}
    center = (coords.min(0) + coords.max(0) + 1) / 2 * B * voxel
    eye = center + np.array([0.0, 0.0, 5.0])
    cam = abi.camera(eye, center, 20, 9, 9)
    img = sim.render(cam, abi.render_settings(smoke_density=density, sun_intensity=0.0))
    alpha = float(img[4, 4, 3])
    ts = np.arange(0, 10, 0.0005)
    d = (center - eye) / np.linalg.norm(center - eye)
    pts = eye[None] + ts[:, None] * d[None]
    inside = np.array([tuple(b) in slotted for b in np.floor(pts / (B * voxel)).astype(int)])
    length = inside.sum() * 0.0005
    assert length > 0.3
    tau = -np.log1p(-alpha)
    want = density * rho * length
    assert abs(tau - want) < 0.1 * want, (tau, want)


def test_incident_light_bounds():
    abi, scenes, _, _ = _mods()
    B = 16
    sim = _run(scenes.campfire(steps=40))
    st = sim.read_state(velocity=False)
    rho, shadow = 0.1, 3.5
    f = np.zeros_like(st["fields"])
    f[:, :, 0] = rho
    sim.write_state(f, None)
    cam = abi.camera((0, 1, 4), (0, 1, 0), 45, 4, 4)
    sun = np.array([0.5, 1.0, 0.3])
    sim.render(cam, abi.render_settings(shadow_density=shadow, fire_intensity=0.0, sun_direction=tuple(sun)),
               image=False)
    rt = sim.read_render()
    light = rt["light"]
    assert np.all(light[..., :3] == 0.0), "no fire, no glow"
    w = light[..., 3]
    assert w.min() > 0 and w.max() <= 1.0
    # Densest path: rho everywhere to the box's exit along the sun.
    box = rt["boxes"][0]
    voxel, texel = 0.05, 0.05 * 4
    s = sun / np.linalg.norm(sun)
    lo, hi = box[:3] * B * voxel, (box[:3] + box[4:7]) * B * voxel
    LE = B // 4 + 1
    coords = sim.brick_coords(rt["bricks"])
    for i, c in enumerate(coords):
        for t in range(LE ** 3):
            loc = np.array([t % LE, (t // LE) % LE, t // (LE * LE)])
            p = (c * (B // 4) + loc + 0.5) * texel
            exit_d = np.min(np.where(s > 0, (hi - p) / s, (lo - p) / s))
            assert w[i, t] >= np.exp(-shadow * rho * max(exit_d, 0)) * (1 - 1e-5) - 1e-7
    # Without a sun, no shadowing at all.
    sim.render(cam, abi.render_settings(fire_intensity=0.0, sun_intensity=0.0), image=False)
    assert np.all(sim.read_render()["light"][..., 3] == 1.0)


def _lit_campfire(steps=24):
    from firepro import scenes
    from oracle import OracleSimulation
    sc = scenes.campfire(steps=steps)
    sim = OracleSimulation.create(sc.grid(), 32)
    for frame, g in sc.frames():
        sim.apply_grid(g)
        sim.step(frame)
    return sim


def test_scene_lights_invariants():
    """Scene-light totals (the restatement of collectLights): the flames' total luminance
    and their centroid do not depend on where the anchors stand; a far anchor and the second
    of two coincident anchors get nothing; height above an anchor counts a quarter; the
    totals scale exactly with the flame brightness; nothing is lit before the first step."""
    from firepro import abi, scenes
    from oracle import OracleSimulation
    rs = abi.render_settings()
    fresh = OracleSimulation.create(scenes.campfire(steps=1).grid(), 32)
    assert not fresh.scene_lights([(0, 0.2, 0)], rs).any()
    sim = _lit_campfire()
    one = sim.scene_lights([(0.0, 0.2, 0.0)], rs)
    total = float(one[0, 0])
    assert total > 0
    moved = sim.scene_lights([(0.3, 1.0, -0.2)], rs)
    np.testing.assert_allclose(moved[0, 0], total, rtol=1e-6)
    np.testing.assert_allclose(moved[0, 4:7], one[0, 4:7], rtol=1e-6)
    # The centroid anchor + offset / luminance is the same from both anchors.
    c1 = np.array([0.0, 0.2, 0.0]) + one[0, 1:4] / one[0, 0]
    c2 = np.array([0.3, 1.0, -0.2]) + moved[0, 1:4] / moved[0, 0]
    np.testing.assert_allclose(c1, c2, atol=1e-5)
    split = sim.scene_lights([(0.0, 0.2, 0.0), (40.0, 0.0, 40.0), (0.0, 0.2, 0.0)], rs)
    assert not split[1].any() and not split[2].any(), "far and duplicate anchors get nothing"
    np.testing.assert_array_equal(split[0], one[0])
    # Flames rise from their anchor: height above an anchor counts a quarter. The lit bricks
    # are the bottom layer (centers at y = 0.4 m). An anchor at y = -1 m is 1.4 m below them,
    # 0.35 m by the rule; one at y = 1.2 m is 0.8 m above them either way. Plain distance
    # would give them to the upper anchor; the rule gives them to the lower.
    below_above = sim.scene_lights([(0.0, -1.0, 0.0), (0.0, 1.2, 0.0)], rs)
    np.testing.assert_allclose(below_above[0, 0], total, rtol=1e-6)
    assert below_above[1, 0] == 0
    # Brightness doubles every sample's emission: the totals double exactly.
    brighter = sim.scene_lights([(0.0, 0.2, 0.0)], abi.render_settings(fire_intensity=2 * rs.fire_intensity))
    np.testing.assert_array_equal(brighter[0, [0, 4, 5, 6]], 2 * one[0, [0, 4, 5, 6]])


def test_scene_light_placement():
    """SceneLights' host part: a light at anchor + offset / luminance, coloured by the
    normalized rgb, as bright as the luminance times the sample cube's volume (capped at
    400), reaching to where it falls to 0.01; an explosion's light finishes once it has been
    read ten times and is below 1 % of its peak; without clustered lighting the four lights go to the
    anchors brightest at the viewer."""
    from firepro.lights import LightAnchor, SceneLights
    sl = SceneLights()
    anchors = [LightAnchor("a", (0, 0, 0)), LightAnchor("b", (10, 0, 0), transient=True)]
    sums = np.zeros((2, 8), np.float32)
    sums[0, :7] = [8.0, 8.0, 16.0, 0.0, 2.0, 1.0, 0.5]
    sums[1, :7] = [40000.0, 0, 0, 0, 1, 1, 1]
    sl.apply(sums, anchors, 1.0, brick_cells=16, voxel_size=0.05)
    a, b = sl.measured["a"], sl.measured["b"]
    assert a.position == [1.0, 2.0, 0.0] and a.color == [1.0, 0.5, 0.25]
    sample = 16 / 4 * 0.05
    assert abs(a.intensity - 8.0 * sample ** 3 * 2) < 1e-12
    assert b.intensity == 400.0
    lights = sl.lights
    assert len(lights) == 2 and abs(lights[0].distance - (lights[0].intensity / 0.01) ** 0.5) < 1e-12
    # Viewer at the far anchor: it lights the viewer most, so it comes first.
    sl.viewer = [10, 0, 0]
    sl.place()
    assert sl.lights[0].position == [10.0, 0.0, 0.0]
    faded = sums.copy()
    faded[1, 0] = 1.0
    for _ in range(8):  # readings 2 to 9
        sl.apply(faded, anchors, 1.0, 16, 0.05)
    assert not sl.finished("b")
    sl.apply(faded, anchors, 1.0, 16, 0.05)
    assert sl.finished("b") and not sl.finished("a")
    sl.apply(sums[:1], anchors[:1], 1.0, 16, 0.05)
    assert "b" not in sl.measured


def test_coarse_smoke_render_field_is_the_121_filter_of_the_smoke_grid():
    """With smokeDivisor 2 the render smoke is the separable [1,2,1]/4 filter of the coarse
    density (scipy) on the smoke lattice, the render field holds the fine fields unfiltered,
    and every block whose own smoothed smoke passes the cutoff is marked visible."""
    abi, scenes, _, _ = _mods()
    B, sd = 8, 2
    n = B // sd
    sim = _run(scenes.small(steps=8, brick=B, divisor=2, smoke=sd))
    sim.render(abi.camera((0, 1, 4), (0, 1, 0), 45, 4, 4), image=False)
    rt = sim.read_render()
    sm = sim.read_smoke()
    st = sim.read_state(velocity=False)
    dense, origin, _ = _dense(sim, sm["ids"], sm["smoke"][..., None], n)
    dense = dense[..., 0].astype(np.float64)
    k = np.array([1.0, 2.0, 1.0]) / 4
    smooth = dense
    for axis in range(3):
        smooth = ndimage.correlate1d(smooth, k, axis=axis, mode="constant")
    fine = {int(i): f for i, f in zip(st["ids"], st["fields"])}
    lcoords = sim.brick_coords(rt["bricks"])
    box = rt["boxes"][0]
    box_lo, box_hi = box[:3] * n, (box[:3] + box[4:7]) * n - 1
    ES, E = n + 1, B + 1
    checked = 0
    for i, c in enumerate(lcoords):
        tex = rt["smoke"][i].reshape(ES, ES, ES)
        field = rt["field"][i].reshape(E, E, E, 4)
        np.testing.assert_array_equal(field[:B, :B, :B].reshape(-1, 4), fine[int(rt["bricks"][i])])
        for lz in range(ES):
            for ly in range(ES):
                for lx in range(ES):
                    cell = c * n + np.array([lx, ly, lz])
                    d = cell - origin
                    if np.any(cell - 1 < box_lo) or np.any(cell + 1 > box_hi):
                        continue
                    if np.any(d < 1) or np.any(d + 1 >= np.array(dense.shape[::-1])):
                        continue
                    want = smooth[d[2], d[1], d[0]]
                    assert abs(tex[lz, ly, lx] - want) <= 1e-6 * max(1.0, abs(want)), (cell, want)
                    checked += 1
        own = tex[:n, :n, :n]
        for b in range(len(rt["occupancy"][i])):
            if own.max() > 0.002 and B == 8:
                assert rt["occupancy"][i][b] & 2, (c, own.max())
    assert checked > 200


def test_half_resolution_is_the_upsampled_half_size_march():
    """With half_resolution and no scene depth, each pixel resumes from the last depth slice
    of the half-size march, sampled bilinearly (clamped to the edge): the image is within a
    fraction of a percent of a full march rendered at half size and upsampled by numpy. The
    slices' steps also stop at the eight depths, so the two differ by that partition only;
    the full-resolution image differs from both far more at the edges of the flames."""
    abi, scenes, _, _ = _mods()
    sim = _run(scenes.campfire(steps=30))
    W, H = 64, 48
    eye, at = (-4.12, 1.96, 5.30), (0, 1.25, 0)
    half = sim.render(abi.camera(eye, at, 44, W, H), abi.render_settings(half_resolution=True))
    full = sim.render(abi.camera(eye, at, 44, W, H), abi.render_settings(half_resolution=False))
    small = sim.render(abi.camera(eye, at, 44, W // 2, H // 2), abi.render_settings(half_resolution=False))
    h, w = small.shape[:2]
    fx = (np.arange(W) + 0.5) / W * w - 0.5
    fy = (np.arange(H) + 0.5) / H * h - 0.5
    gx, gy = np.floor(fx), np.floor(fy)
    tx, ty = (fx - gx)[None, :, None], (fy - gy)[:, None, None]
    x0, x1 = np.clip(gx, 0, w - 1).astype(int), np.clip(gx + 1, 0, w - 1).astype(int)
    y0, y1 = np.clip(gy, 0, h - 1).astype(int), np.clip(gy + 1, 0, h - 1).astype(int)
    s = small.astype(np.float64)
    up = (s[y0][:, x0] * (1 - tx) * (1 - ty) + s[y0][:, x1] * tx * (1 - ty)
          + s[y1][:, x0] * (1 - tx) * ty + s[y1][:, x1] * tx * ty)
    scale = np.abs(half).max()
    assert half[..., 3].max() > 0.3
    assert np.abs(half - up).max() <= 0.005 * scale
    assert np.abs(half - full).max() > 10 * np.abs(half - up).max()
