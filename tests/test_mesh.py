"""Mesh emitters: the geometry playback builds, the seeded surface samples
(src/library/MeshSurface.ts) and the emission the oracle's rasterizer
(MeshSourceRasterizer.ts) injects, each pinned by closed forms or brute force."""
import math

import numpy as np
import pytest

from firepro import mesh
from firepro.library import FireSimulation


def test_geometry_areas_closed_form():
    """A 40-segment disk is the regular 40-gon (n/2 r^2 sin(2 pi/n)); a box has area
    2(wh + hd + dw); a 12 x 40 torus approaches 4 pi^2 R r; laying a shape flat (rotateX by
    -pi/2) keeps its area and turns the disk's normal to +y."""
    def area(g):
        t = g.triangles()
        return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1).sum()
    disk = mesh.circle(0.8, 40)
    assert area(disk) == pytest.approx(0.5 * 40 * 0.64 * math.sin(2 * math.pi / 40), rel=1e-6)
    assert area(mesh.box(0.4, 0.1, 0.3)) == pytest.approx(2 * (0.4 * 0.1 + 0.1 * 0.3 + 0.3 * 0.4), rel=1e-6)
    assert area(mesh.torus(1.0, 0.4, 48, 160)) == pytest.approx(4 * math.pi ** 2 * 0.4, rel=2e-3)
    flat = mesh.document_geometry("disk", 0.8, (1, 1, 1))
    assert area(flat) == pytest.approx(area(disk), rel=1e-6)
    t = flat.triangles()
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    assert np.all(n[:, 1] > 0) and np.allclose(n[:, [0, 2]], 0, atol=1e-6)
    # Box faces point outward: each triangle's normal agrees with its centroid's direction.
    b = mesh.box(0.4, 0.1, 0.3).triangles()
    nb = np.cross(b[:, 1] - b[:, 0], b[:, 2] - b[:, 0])
    assert np.all((nb * b.mean(axis=1)).sum(axis=1) > 0)


def test_lcg_known_answer():
    """state' = 1664525 state + 1013904223 mod 2^32 (Numerical Recipes' generator)."""
    g = mesh.lcg(1)
    s1 = (1664525 + 1013904223) % 2 ** 32
    s2 = (s1 * 1664525 + 1013904223) % 2 ** 32
    assert next(g) == s1 / 2 ** 32 and next(g) == s2 / 2 ** 32


def test_surface_samples_on_the_offset_surface():
    """Samples of a flat disk lie 0.75 of a cell above it and inside its rim, their count is
    min(8192, max(256, ceil(8 area / cell^2))), their weights sum to one, and they are spread
    by area: about half fall inside radius r / sqrt(2)."""
    r, cell = 0.8, 0.05
    s = mesh.MeshSurface(mesh.document_geometry("disk", r, (1, 1, 1)), cell, 7)
    area = 0.5 * 40 * r * r * math.sin(2 * math.pi / 40)
    assert len(s.samples) == min(8192, max(256, math.ceil(area / cell ** 2 * 8)))
    assert s.samples[:, 3].sum() == pytest.approx(1.0)
    np.testing.assert_allclose(s.samples[:, 1], 0.75 * cell, atol=1e-6)
    rho = np.hypot(s.samples[:, 0], s.samples[:, 2])
    assert rho.max() <= r
    assert abs((rho < r / math.sqrt(2)).mean() - 0.5) < 0.02
    small = mesh.MeshSurface(mesh.box(0.01, 0.01, 0.01), cell, 1)
    assert len(small.samples) == 256


def test_transform_places_samples():
    """compose(position, Euler XYZ) is a rigid motion: distances between samples are kept
    and a quarter turn about y takes +x to -z."""
    m = mesh.compose((1.0, 2.0, 3.0), (0.0, math.pi / 2, 0.0))
    p = mesh.transform_samples(np.array([[1.0, 0.0, 0.0, 0.5]]), m)
    np.testing.assert_allclose(p[0], [1.0, 2.0, 2.0, 0.5], atol=1e-12)
    s = mesh.MeshSurface(mesh.box(0.4, 0.1, 0.3), 0.05, 3).samples
    q = mesh.transform_samples(s, mesh.compose((0.3, 0.35, 0.0), (0.2, 0.5, 0.0)))
    d0 = np.linalg.norm(s[1:, :3] - s[:-1, :3], axis=1)
    d1 = np.linalg.norm(q[1:, :3] - q[:-1, :3], axis=1)
    np.testing.assert_allclose(d0, d1, rtol=1e-12, atol=1e-15)


def test_mesh_emission_conserves_coverage():
    """One step of a heat-only mesh emitter on the oracle: every covered cell receives
    heat rate x min(1, coverage) x dt, decayed over half a step, where coverage is the
    brute-force trilinear splat of the samples' weight x area in units of the cell face
    area. The total heat of the step is that sum over the covered cells."""
    import oracle
    cell = 0.05
    sim = FireSimulation({"voxelSize": cell, "brickSize": 8, "velocityDivisor": 2, "seed": 5,
                          "motion": {"buoyancy": 0, "smokeWeight": 0, "vorticity": 0}})
    sim.add_emitter({"shape": {"type": "mesh", "geometry": mesh.box(0.3, 0.2, 0.25)},
                         "position": (0.1, 0.6, -0.05), "rotation": (0.3, 0.1, 0.0),
                         "emission": {"heatRate": 3.0}})
    frame = sim.build_frame()
    src = frame.sources[0]
    cov = {}
    for x, y, z, w in src.samples:
        g = np.array([x, y, z]) / cell - 0.5
        base = np.floor(g).astype(int)
        f = g - base
        for k in range(8):
            o = np.array([k & 1, (k >> 1) & 1, k >> 2])
            t = np.prod(np.where(o == 1, f, 1 - f))
            key = tuple(base + o)
            cov[key] = cov.get(key, 0.0) + w * src.surface_area * t / cell ** 2
    total_mask = sum(min(1.0, v) for v in cov.values() if v > 0)
    assert max(cov.values()) > 0.2
    orc = oracle.OracleSimulation.create({"voxelSize": cell, "velocityDivisor": 2, "brickSize": 8,
                                          "smokeDivisor": 1, "scalarMacCormack": True})
    orc.apply_grid(sim.grid_settings())
    orc.step(frame)
    heat = orc.read_state()["fields"][..., 1].astype(np.float64).sum()
    dt = 1 / 60
    half = math.exp(-frame.cooling * dt * 0.5)
    assert heat == pytest.approx(3.0 * dt * half * total_mask, rel=1e-5)


def _inside_count(mask, lo, spacing, inside):
    """The cells of the box whose centers `inside` (world coordinates) accepts, and the mask's."""
    nz, ny, nx = mask.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    centers = (np.stack([x, y, z], -1) + np.array(lo) + 0.5) * spacing
    return inside(centers), mask.astype(bool)


@pytest.mark.parametrize("rotation", [(0.0, 0.0, 0.0), (0.3, 0.2, 0.1)])
def test_voxelized_box_is_its_cell_centers(rotation):
    """MeshCollider.voxelize of a closed box marks exactly the cells whose centers lie inside
    the placed box (a cell of margin around the bounds stays empty)."""
    pos, cell = (0.1, 0.5, -0.05), 0.05
    m = mesh.compose(pos, rotation)
    mask, lo, size = mesh.voxelize(mesh.box(0.5, 0.3, 0.4), m, cell)
    assert mask.shape == (size[2], size[1], size[0])
    r = m[:3, :3]

    def inside(c):
        local = (c - np.array(pos)) @ r  # world to the box's axes
        return (np.abs(local[..., 0]) < 0.25) & (np.abs(local[..., 1]) < 0.15) & (np.abs(local[..., 2]) < 0.2)

    want, got = _inside_count(mask, lo, cell, inside)
    # Cells whose centers sit within rounding of a face may go either way.
    local = ((np.stack(np.meshgrid(*[np.arange(n) for n in size[::-1]], indexing="ij")[::-1], -1)
              + np.array(lo) + 0.5) * cell - np.array(pos)) @ r
    margin = np.min(np.abs(np.abs(local) - np.array([0.25, 0.15, 0.2])), axis=-1) < 1e-9
    assert np.array_equal(want[~margin], got[~margin])
    assert got.sum() > 100 and not got[0].any() and not got[-1].any()


def test_voxelized_octahedron_and_closedness():
    """A closed octahedron voxelizes to the cells with |x| + |y| + |z| < r about its center;
    an open mesh and an inconsistently wound one are refused."""
    cell, r, c0 = 0.05, 0.31, np.array([0.02, 0.4, -0.03])
    mask, lo, _ = mesh.voxelize(mesh.octahedron(r), mesh.compose(c0, (0, 0, 0)), cell)
    want, got = _inside_count(mask, lo, cell, lambda c: np.abs(c - c0).sum(-1) < r)
    assert np.array_equal(want, got) and got.sum() > 100
    box = mesh.box(0.3, 0.2, 0.2)
    mesh.check_closed(box)
    with pytest.raises(ValueError, match="closed"):
        mesh.check_closed(mesh.Geometry(box.positions, box.indices[:-3]))
    flipped = box.indices.copy()
    flipped[:3] = flipped[:3][::-1]
    with pytest.raises(ValueError, match="consistently wound"):
        mesh.check_closed(mesh.Geometry(box.positions, flipped))


def test_mesh_collider_solids_on_the_oracle():
    """The oracle draws a mesh collider's occupancy into the solids: with a mesh box over
    the emitter, cells inside the box get no heat, cells beside it do."""
    import oracle
    from firepro import scenes
    sc = scenes.meshes(steps=3, colliders=True)
    orc = oracle.OracleSimulation.create(sc.grid())
    for frame, g in sc.frames():
        orc.apply_grid(g)
        orc.step(frame)
    frame = sc.sim.build_frame()
    assert [c.shape for c in frame.colliders] == ["mesh", "mesh"]
    st = orc.read_state(velocity=False)
    B = 8
    coords = orc.brick_coords(st["ids"])
    heat = {}
    for c, f in zip(coords, st["fields"]):
        cells = np.stack(np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")[::-1], -1).reshape(-1, 3)
        for q, v in zip(cells + c * B, f):
            heat[tuple(q)] = v
    solid = 0
    for col in frame.colliders:
        mask, lo, size = col.occupancy
        for z, y, x in zip(*np.nonzero(mask)):
            q = (x + lo[0], y + lo[1], z + lo[2])
            if q in heat:
                assert not heat[q].any(), q
                solid += 1
    hot = [q for q, v in heat.items() if v[1] > 0]
    assert solid > 50 and len(hot) > 100
