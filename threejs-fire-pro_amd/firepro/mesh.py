"""Mesh surfaces for mesh emitters: the three shapes a preset's emitter may take besides a
sphere (box, disk, torus, as examples/playback/main.ts:58-70 builds them with three.js
geometry), their seeded, area-weighted surface samples (src/library/MeshSurface.ts), and the
rigid transform that places them in the world (handles.ts:160-183).

Positions are stored in single precision, as three.js buffer attributes hold them; the
sampling and the transform run in double precision. The geometry generators follow the
public three.js layouts (vertex and triangle order): CircleGeometry's fan around a center
vertex, BoxGeometry's six planes (+x, -x, +y, -y, +z, -z), TorusGeometry's tube rings.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np


@dataclass
class Geometry:
    positions: np.ndarray                 # (n, 3) float32
    indices: Optional[np.ndarray] = None  # (3 t,) triangle corners, or None for a soup

    def triangles(self) -> np.ndarray:
        """(t, 3, 3) corner positions in double precision."""
        idx = self.indices if self.indices is not None else np.arange(len(self.positions))
        if len(idx) % 3:
            raise ValueError("Mesh emitter geometry must contain complete triangles.")
        return self.positions.astype(np.float64)[np.asarray(idx).reshape(-1, 3)]


def circle(radius: float, segments: int = 32) -> Geometry:
    """CircleGeometry: a center vertex, then segments + 1 rim vertices; triangles (i, i+1, 0)."""
    pos = [(0.0, 0.0, 0.0)]
    for s in range(segments + 1):
        a = s / segments * (2 * math.pi)
        pos.append((radius * math.cos(a), radius * math.sin(a), 0.0))
    idx = []
    for i in range(1, segments + 1):
        idx += [i, i + 1, 0]
    return Geometry(np.array(pos, np.float32), np.array(idx, np.int64))


def box(width: float, height: float, depth: float) -> Geometry:
    """BoxGeometry with one segment per side: six planes of four vertices, two triangles
    each, in the order +x, -x, +y, -y, +z, -z."""
    pos, idx = [], []

    def plane(u, v, w, udir, vdir, pw, ph, pd):
        base = len(pos)
        for iy in range(2):
            y = iy * ph - ph / 2
            for ix in range(2):
                x = ix * pw - pw / 2
                vec = [0.0, 0.0, 0.0]
                vec[u] = x * udir
                vec[v] = y * vdir
                vec[w] = pd / 2
                pos.append(tuple(vec))
        a, b, c, d = base, base + 2, base + 3, base + 1
        idx.extend([a, b, d, b, c, d])

    plane(2, 1, 0, -1, -1, depth, height, width)
    plane(2, 1, 0, 1, -1, depth, height, -width)
    plane(0, 2, 1, 1, 1, width, depth, height)
    plane(0, 2, 1, 1, -1, width, depth, -height)
    plane(0, 1, 2, 1, -1, width, height, depth)
    plane(0, 1, 2, -1, -1, width, height, -depth)
    return Geometry(np.array(pos, np.float32), np.array(idx, np.int64))


def torus(radius: float, tube: float, radial: int = 12, tubular: int = 48) -> Geometry:
    """TorusGeometry: (radial + 1) rings of (tubular + 1) vertices; two triangles per quad."""
    pos = []
    for j in range(radial + 1):
        for i in range(tubular + 1):
            u = i / tubular * (2 * math.pi)
            v = j / radial * (2 * math.pi)
            pos.append(((radius + tube * math.cos(v)) * math.cos(u),
                        (radius + tube * math.cos(v)) * math.sin(u), tube * math.sin(v)))
    idx = []
    for j in range(1, radial + 1):
        for i in range(1, tubular + 1):
            a = (tubular + 1) * j + i - 1
            b = (tubular + 1) * (j - 1) + i - 1
            c = (tubular + 1) * (j - 1) + i
            d = (tubular + 1) * j + i
            idx += [a, b, d, b, c, d]
    return Geometry(np.array(pos, np.float32), np.array(idx, np.int64))


def rotate_x(g: Geometry, theta: float) -> Geometry:
    """BufferGeometry.rotateX: y' = c y - s z, z' = s y + c z, stored back in single precision."""
    c, s = math.cos(theta), math.sin(theta)
    p = g.positions.astype(np.float64)
    out = np.empty_like(p)
    out[:, 0] = p[:, 0]
    out[:, 1] = c * p[:, 1] + (-s) * p[:, 2]
    out[:, 2] = s * p[:, 1] + c * p[:, 2]
    return Geometry(out.astype(np.float32), g.indices)


def lcg(seed: int):
    """MeshSurface's generator (MeshSurface.ts:76-80): state = state * 1664525 + 1013904223
    mod 2^32, returned over 2^32."""
    state = seed & 0xFFFFFFFF
    while True:
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        yield state / 4294967296.0


class MeshSurface:
    """MeshSurface (MeshSurface.ts:15-97): the surface area of the nondegenerate triangles and
    min(8192, max(256, ceil(8 area / cell^2))) samples, the i-th in the triangle holding
    (i + 1/2) / count of the cumulative area, at a seeded uniform point of it, moved 0.75 of
    a cell along the triangle's normal; each weighs 1 / count. Local coordinates."""

    def __init__(self, geometry: Geometry, cell_size: float, seed: int):
        tri = geometry.triangles()
        if not np.all(np.isfinite(tri)):
            raise ValueError("Mesh emitter geometry contains nonfinite positions.")
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        normal = np.cross(b - a, c - a)
        areas = np.sqrt((normal * normal).sum(axis=1)) * 0.5
        cumulative = np.empty(len(tri))
        area = 0.0
        for t in range(len(tri)):  # running sum in triangle order, degenerate ones skipped
            if areas[t] > 1e-12:
                area += float(areas[t])
            cumulative[t] = area
        if not area > 0:
            raise ValueError("Mesh emitter geometry has no nondegenerate triangles.")
        self.surface_area = area
        count = min(8192, max(256, math.ceil(area / cell_size ** 2 * 8)))
        rnd = lcg(seed)
        samples = np.empty((count, 4))
        t = 0
        for i in range(count):
            target = (i + 0.5) / count * area
            while cumulative[t] < target:
                t += 1
            n = normal[t]
            length = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
            n = n * (1 / (length or 1))
            root = math.sqrt(next(rnd))
            along = next(rnd)
            p = a[t] * (1 - root)
            p = p + b[t] * (root * (1 - along))
            p = p + c[t] * (root * along)
            p = p + n * (cell_size * 0.75)
            samples[i, :3] = p
        samples[:, 3] = 1 / count
        self.samples = samples


def compose(position: Sequence[float], rotation: Sequence[float]) -> np.ndarray:
    """Matrix4.compose of a position, an Euler rotation in XYZ order (through its quaternion)
    and unit scale: the 4x4 world matrix, row-major."""
    hx, hy, hz = (r / 2 for r in rotation)
    c1, c2, c3 = math.cos(hx), math.cos(hy), math.cos(hz)
    s1, s2, s3 = math.sin(hx), math.sin(hy), math.sin(hz)
    x = s1 * c2 * c3 + c1 * s2 * s3
    y = c1 * s2 * c3 - s1 * c2 * s3
    z = c1 * c2 * s3 + s1 * s2 * c3
    w = c1 * c2 * c3 - s1 * s2 * s3
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    return np.array([
        [1 - (yy + zz), xy - wz, xz + wy, position[0]],
        [xy + wz, 1 - (xx + zz), yz - wx, position[1]],
        [xz - wy, yz + wx, 1 - (xx + yy), position[2]],
        [0.0, 0.0, 0.0, 1.0]])


def transform_samples(samples: np.ndarray, matrix: np.ndarray) -> np.ndarray:
    """Vector3.applyMatrix4 of each sample position (the weight column is kept)."""
    out = samples.copy()
    p = samples[:, :3]
    for r in range(3):
        out[:, r] = ((matrix[r, 0] * p[:, 0] + matrix[r, 1] * p[:, 1]) + matrix[r, 2] * p[:, 2]) + matrix[r, 3]
    return out


def transform_direction(direction: Sequence[float], matrix: np.ndarray) -> Tuple[float, float, float]:
    """Vector3.transformDirection: the rotated direction, normalized."""
    d = [sum(matrix[r, k] * direction[k] for k in range(3)) for r in range(3)]
    length = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    inv = 1 / (length or 1)
    return (d[0] * inv, d[1] * inv, d[2] * inv)


def document_geometry(shape: str, radius: float, size: Sequence[float]) -> Geometry:
    """The mesh playback builds for a document emitter (examples/playback/main.ts:61-67): a
    box of `size`, or a disk (40 segments) or torus (tube 0.4 radius, 12 by 40 segments) of
    `radius` laid flat, facing up."""
    if shape == "box":
        return box(*size)
    if shape == "disk":
        return rotate_x(circle(radius, 40), -math.pi / 2)
    if shape == "torus":
        return rotate_x(torus(radius, radius * 0.4, 12, 40), -math.pi / 2)
    raise ValueError(f"no mesh for emitter shape {shape}.")


def check_closed(g: Geometry) -> None:
    """MeshCollider's constructor (MeshCollider.ts:15-78): every edge of a nondegenerate
    triangle is shared by exactly two triangles that run it in opposite directions."""
    tri = g.triangles()
    edges = {}
    count = 0
    for t in tri:
        a, b, c = t
        n = np.cross(b - a, c - a)
        if float(n @ n) == 0:
            continue
        count += 1
        keys = [tuple(float(v) for v in p) for p in t]
        for e in range(3):
            u, v = keys[e], keys[(e + 1) % 3]
            forward = u < v
            key = (u, v) if forward else (v, u)
            cnt, wind = edges.get(key, (0, 0))
            edges[key] = (cnt + 1, wind + (1 if forward else -1))
    if not count or any(c != 2 or w != 0 for c, w in edges.values()):
        raise ValueError("Mesh colliders require closed, consistently wound manifold surfaces. "
                         "Weld matching seam positions and close open edges.")


def voxelize(g: Geometry, matrix: np.ndarray, spacing: float) -> Tuple[np.ndarray, Tuple[int, int, int],
                                                                        Tuple[int, int, int]]:
    """Collider.meshMask and MeshCollider.voxelize (handles.ts:298-325, MeshCollider.ts:80-171):
    the cells of the world lattice (cells `spacing` wide from the origin) whose centers the
    closed mesh placed by `matrix` contains, over the mesh's bounds with a cell of margin.
    Each triangle is rasterized over the cell centers of the x-y plane with a top-left tie
    rule; its depth there is a crossing of the z column, signed by its winding; a cell is
    inside where the crossings below its center sum to nonzero. Returns the occupancy
    (nz, ny, nx), the box's first cell and its size (nx, ny, nz)."""
    p = g.positions.astype(np.float64)
    # Box3.applyMatrix4 of the geometry's bounding box: the bounds of its eight corners.
    blo, bhi = p.min(axis=0), p.max(axis=0)
    corners = np.array([[(blo, bhi)[k >> a & 1][a] for a in range(3)] for k in range(8)])
    world = transform_samples(np.c_[corners, np.zeros(8)], matrix)[:, :3]
    lo = tuple(int(math.floor(world[:, a].min() / spacing)) - 1 for a in range(3))
    size = tuple(int(math.ceil(world[:, a].max() / spacing)) + 1 - lo[a] for a in range(3))
    nx, ny, nz = size
    tri = g.triangles()
    verts = transform_samples(np.c_[tri.reshape(-1, 3), np.zeros(len(tri) * 3)], matrix)[:, :3]
    # The voxelizer's box coordinates: cells from lo (toBox of handles.ts:312-318).
    cellv = np.empty_like(verts)
    dom = [size[a] * spacing for a in range(3)]
    shifted = verts - np.array([(lo[0] + size[0] / 2) * spacing, lo[1] * spacing, (lo[2] + size[2] / 2) * spacing])
    cellv[:, 0] = (shifted[:, 0] + dom[0] / 2) / spacing
    cellv[:, 1] = shifted[:, 1] / spacing
    cellv[:, 2] = (shifted[:, 2] + dom[2] / 2) / spacing
    cellv = cellv.reshape(-1, 3, 3)

    def edge(a, b, x, y):
        return (a[0] - x) * (b[1] - y) - (a[1] - y) * (b[0] - x)

    def owns(a, b):
        return b[1] > a[1] or (b[1] == a[1] and b[0] < a[0])

    columns = [[] for _ in range(nx * ny)]
    for a, b, c in cellv:
        area = edge(a, b, c[0], c[1])
        if area == 0:
            continue
        sign = 1 if area > 0 else -1
        if area < 0:
            b, c = c, b
            area = -area
        x0 = max(0, math.ceil(min(a[0], b[0], c[0]) - 0.5))
        x1 = min(nx - 1, math.floor(max(a[0], b[0], c[0]) - 0.5))
        y0 = max(0, math.ceil(min(a[1], b[1], c[1]) - 0.5))
        y1 = min(ny - 1, math.floor(max(a[1], b[1], c[1]) - 0.5))
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                v = (edge(b, c, x + 0.5, y + 0.5), edge(c, a, x + 0.5, y + 0.5), edge(a, b, x + 0.5, y + 0.5))
                pairs = ((b, c), (c, a), (a, b))
                if any(val < 0 or (val == 0 and not owns(*pairs[i])) for i, val in enumerate(v)):
                    continue
                columns[y * nx + x].append(((v[0] * a[2] + v[1] * b[2] + v[2] * c[2]) / area, sign))
    mask = np.zeros((nz, ny, nx), np.uint32)
    for col, crossings in enumerate(columns):
        crossings.sort(key=lambda e: e[0])
        winding, cursor = 0, 0
        for z in range(nz):
            while cursor < len(crossings) and crossings[cursor][0] <= z + 0.5:
                winding += crossings[cursor][1]
                cursor += 1
            if winding != 0:
                mask[z, col // nx, col % nx] = 1
    return mask, lo, size


def octahedron(radius: float) -> Geometry:
    """A closed octahedron of the given vertex radius, outward winding: the points with
    |x| + |y| + |z| <= radius."""
    r = radius
    pos = [(r, 0, 0), (-r, 0, 0), (0, r, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    idx = []
    for sx, x in ((1, 0), (-1, 1)):
        for sy, y in ((1, 2), (-1, 3)):
            for sz, z in ((1, 4), (-1, 5)):
                idx += [x, y, z] if sx * sy * sz > 0 else [x, z, y]
    return Geometry(np.array(pos, np.float32), np.array(idx, np.int64))
