"""ctypes mirror of include/firepro.h and the loader of the HIP library.

The structures follow the reference's plain-data frame (src/engine/types.ts:235-311):
`SimulationFrame`, `SimulationSource`, `SimulationForce`, `SimulationCollider`. The loader
fails loudly when libfirepro.so is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libfirepro.so")

FP_OK, FP_EINVAL, FP_EDEVICE, FP_ESTATE, FP_ENOMEM = 0, -1, -2, -3, -4
SHAPES = {"sphere": 0, "box": 1, "mesh": 2}
FORCE_TYPES = {"wind": 0, "turbulence": 1, "vortex": 2, "radial": 3}
MAX_FLAME_KNOTS = 8
BRICK_RECORD_WORDS = 28
FRESH_ENTRY = 0x2000000
FRESH_SLOT = 0x80000000

Vec3 = C.c_double * 3


class Grid(C.Structure):
    _fields_ = [
        ("voxel_size", C.c_double),
        ("velocity_divisor", C.c_int32),
        ("brick_size", C.c_int32),
        ("smoke_divisor", C.c_int32),
        ("scalar_maccormack", C.c_int32),
    ]


class Settings(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "buoyancy", "smoke_weight", "velocity_damping", "cooling", "dissipation", "vorticity",
        "seed")]


class Flame(C.Structure):
    _fields_ = [
        ("lifespan", C.c_double), ("smoke_rate", C.c_double), ("heat_rate", C.c_double),
        ("expansion_rate", C.c_double),
        ("knot_count", C.c_int32 * 3),
        ("knots", ((C.c_double * 2) * MAX_FLAME_KNOTS) * 3),
    ]


class Fuel(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("ignition_heat", C.c_double), ("burn_rate", C.c_double)]


class Source(C.Structure):
    _fields_ = [
        ("id", C.c_uint64), ("shape", C.c_int32),
        ("position", Vec3), ("size", Vec3),
        ("flame", C.c_double), ("heat_rate", C.c_double), ("smoke_rate", C.c_double),
        ("fuel_rate", C.c_double),
        ("velocity", Vec3),
        ("velocity_response", C.c_double), ("outward_speed", C.c_double),
        ("has_variation", C.c_int32),
        ("variation_period", C.c_double), ("variation_strength", C.c_double),
        ("samples", C.POINTER(C.c_double)), ("sample_count", C.c_int32),
        ("surface_area", C.c_double),
    ]


class Force(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("vector", Vec3), ("center", Vec3),
        ("strength", C.c_double), ("scale", C.c_double), ("lift", C.c_double),
        ("inward", C.c_double),
        ("targeted", C.c_int32), ("source_count", C.c_int32),
        ("sources", C.POINTER(Source)),
    ]


class Collider(C.Structure):
    _fields_ = [("shape", C.c_int32), ("position", Vec3), ("size", Vec3),
                ("axes", (C.c_double * 3) * 3),
                ("occupancy", C.POINTER(C.c_uint32)), ("lo", C.c_int32 * 3), ("cells", C.c_int32 * 3)]


class Frame(C.Structure):
    _fields_ = [
        ("simulation", Settings), ("combustion", Flame),
        ("has_fuel", C.c_int32), ("fuel", Fuel),
        ("source_count", C.c_int32), ("sources", C.POINTER(Source)),
        ("force_count", C.c_int32), ("forces", C.POINTER(Force)),
        ("collider_count", C.c_int32), ("colliders", C.POINTER(Collider)),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("steps", C.c_int64), ("active_bricks", C.c_int64),
        ("active_voxels", C.c_int64), ("pool_slots", C.c_int64), ("used_slots", C.c_int64),
        ("pool_limited", C.c_int32), ("tile_capacity", C.c_int32), ("tile_count", C.c_int32),
        ("pool_side", C.c_int32), ("memory_bytes", C.c_int64), ("mesh_memory_bytes", C.c_int64),
        ("force_memory_bytes", C.c_int64),
    ]


class RenderBricks(C.Structure):
    _fields_ = [
        ("version", C.c_int32), ("boxes", C.POINTER(C.c_int32)), ("tiles", C.POINTER(C.c_int32)),
        ("box_capacity", C.c_int32), ("tile_capacity", C.c_int32), ("box_count", C.c_int32),
        ("tile_count", C.c_int32), ("pages_device", C.c_void_p), ("active_device", C.c_void_p),
        ("zeroed_device", C.c_void_p), ("brick_count", C.c_int32), ("active_count", C.c_int32),
        ("zeroed_count", C.c_int32),
    ]


SOLVER_FIELDS = {None: 0, "pressure": 1, "divergence": 2}


class Tables(C.Structure):
    _fields_ = [
        ("pages", C.POINTER(C.c_int32)), ("entries", C.POINTER(C.c_int32)),
        ("records", C.POINTER(C.c_int32)), ("free_slots", C.POINTER(C.c_uint32)),
        ("active", C.POINTER(C.c_int32)), ("zeroed", C.POINTER(C.c_int32)),
        ("brick_count", C.c_int32), ("free_count", C.c_int32), ("failures", C.c_int32),
        ("active_count", C.c_int32), ("zeroed_count", C.c_int32), ("pool_slots", C.c_int32),
    ]


class Camera(C.Structure):
    _fields_ = [("inverse_view_projection", C.c_double * 16), ("position", Vec3),
                ("width", C.c_int32), ("height", C.c_int32)]


class RenderSettings(C.Structure):
    _fields_ = [
        ("fire_intensity", C.c_double), ("flame_opacity", C.c_double),
        ("smoke_density", C.c_double), ("scattering", C.c_double),
        ("shadow_density", C.c_double), ("temperature", C.c_double), ("soot_glow", C.c_double),
        ("smoke_color", Vec3), ("fire_color", Vec3),
        ("sun_direction", Vec3), ("sun_intensity", C.c_double),
        ("sun_color", Vec3), ("ambient", Vec3),
        ("filter", C.c_int32), ("ray_steps", C.c_int32), ("lighting_divisor", C.c_int32),
        ("half_resolution", C.c_int32),
    ]


class RenderTables(C.Structure):
    _fields_ = [
        ("boxes", C.POINTER(C.c_int32)), ("page_table", C.POINTER(C.c_int32)),
        ("bricks", C.POINTER(C.c_int32)), ("field", C.POINTER(C.c_float)),
        ("light", C.POINTER(C.c_float)), ("occupancy", C.POINTER(C.c_uint8)),
        ("blackbody", C.POINTER(C.c_float)),
        ("box_capacity", C.c_int32), ("page_capacity", C.c_int32), ("brick_capacity", C.c_int32),
        ("box_count", C.c_int32), ("page_count", C.c_int32), ("brick_count", C.c_int32),
        ("light_work", C.c_int32),
        ("smoke", C.POINTER(C.c_float)),
    ]


FILTERS = ("trilinear", "quadratic", "cubic")


def srgb_to_linear(c: float) -> float:
    """IEC 61966-2-1 sRGB decoding (three.js Color's hex input is sRGB)."""
    return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4


def hex_color(h: int):
    return tuple(srgb_to_linear(((h >> s) & 255) / 255.0) for s in (16, 8, 0))


# Rendering defaults (src/library/options.ts:340-365): flame brightness, opacity,
# temperature, soot glow and color; smoke density, scattering, shadow density and color;
# the cubic filter at 256 ray steps with lighting every fourth cell. The sun is the
# editor's (editor/runtime.ts:99): white, intensity 3.
RENDER_DEFAULTS = {
    "fire_intensity": 1.4, "flame_opacity": 0.55, "temperature": 3200.0, "soot_glow": 1.0,
    "fire_color": hex_color(0xFFFFFF), "smoke_density": 2.5, "scattering": 0.38,
    "shadow_density": 3.5, "smoke_color": hex_color(0x25292E),
    "sun_direction": (0.5, 1.0, 0.3), "sun_intensity": 3.0, "sun_color": (1.0, 1.0, 1.0),
    "ambient": (0.0, 0.0, 0.0), "filter": "cubic", "ray_steps": 256, "lighting_divisor": 4,
    "half_resolution": True,  # options.ts:364
}


def render_settings(**overrides) -> RenderSettings:
    o = dict(RENDER_DEFAULTS)
    for k, v in overrides.items():
        if k not in o:
            raise ValueError(f"unknown render setting {k}")
        o[k] = v
    rs = RenderSettings()
    for k in ("fire_intensity", "flame_opacity", "smoke_density", "scattering", "shadow_density",
              "temperature", "soot_glow", "sun_intensity"):
        setattr(rs, k, float(o[k]))
    for k in ("smoke_color", "fire_color", "sun_direction", "sun_color", "ambient"):
        setattr(rs, k, Vec3(*[float(x) for x in o[k]]))
    f = o["filter"]
    rs.filter = FILTERS.index(f) if isinstance(f, str) else int(f)
    rs.ray_steps = int(o["ray_steps"])
    rs.lighting_divisor = int(o["lighting_divisor"])
    rs.half_resolution = int(bool(o["half_resolution"]))
    return rs


_hip_runtime = None


def hip_runtime():
    """The HIP runtime libfirepro.so runs on, through ctypes (device buffers, events)."""
    global _hip_runtime
    if _hip_runtime is not None:
        return _hip_runtime
    load()
    with open("/proc/self/maps") as f:
        paths = sorted({line.split()[-1] for line in f if "libamdhip64" in line})
    if not paths:
        raise FireProError("libfirepro.so has no HIP runtime loaded")
    hip = C.CDLL(paths[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    _hip_runtime = hip
    return hip


class StreamTimer:
    """Device intervals on one stream: mark() records a HIP event there; elapsed() waits for
    the events and returns the milliseconds between consecutive pairs (begin, end) of
    marks, each the device time from the work before `begin` finishing to the work before
    `end` finishing. Events are reused across rounds."""

    def __init__(self, stream: int):
        self._hip = hip_runtime()
        self.stream = C.c_void_p(stream)
        self._free, self._marks = [], []

    def mark(self) -> None:
        if self._free:
            e = self._free.pop()
        else:
            e = C.c_void_p()
            if self._hip.hipEventCreate(C.byref(e)) != 0:
                raise FireProError("hipEventCreate failed")
        if self._hip.hipEventRecord(e, self.stream) != 0:
            raise FireProError("hipEventRecord failed")
        self._marks.append(e)

    def elapsed(self) -> list:
        out, ms = [], C.c_float()
        if self._marks and self._hip.hipEventSynchronize(self._marks[-1]) != 0:
            raise FireProError("hipEventSynchronize failed")
        for a, b in zip(self._marks[0::2], self._marks[1::2]):
            if self._hip.hipEventElapsedTime(C.byref(ms), a, b) != 0:
                raise FireProError("hipEventElapsedTime failed")
            out.append(float(ms.value))
        self._free.extend(self._marks)
        self._marks = []
        return out

    def close(self) -> None:
        for e in self._free + self._marks:
            self._hip.hipEventDestroy(e)
        self._free, self._marks = [], []


class DeviceBuffer:
    """Device memory from the HIP runtime libfirepro.so uses (for device-output entry
    points and probes), released on close()."""

    def __init__(self, nbytes: int):
        self._hip = hip = hip_runtime()
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        if hip.hipMalloc(C.byref(self.ptr), self.nbytes) != 0:
            raise FireProError(f"hipMalloc of {nbytes} bytes failed")

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        if self._hip.hipMemcpy(self.ptr, other.ptr, int(nbytes), 3) != 0:  # device to device
            raise FireProError("hipMemcpy failed")

    def read(self, nbytes: Optional[int] = None) -> bytes:
        """The first `nbytes` (default all) as host bytes, after the device has finished."""
        self.synchronize()
        n = self.nbytes if nbytes is None else int(nbytes)
        out = C.create_string_buffer(n)
        if self._hip.hipMemcpy(out, self.ptr, n, 2) != 0:  # device to host
            raise FireProError("hipMemcpy failed")
        return out.raw

    def synchronize(self) -> None:
        if self._hip.hipDeviceSynchronize() != 0:
            raise FireProError("hipDeviceSynchronize failed")

    def close(self) -> None:
        if self.ptr.value:
            self._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def read_render_tables(call, brick_size: int, lighting_divisor: int, smoke_divisor: int = 1) -> dict:
    """Run a read-render entry point (`call(RenderTables) -> status`) twice: once for the
    counts, once into arrays of those sizes."""
    import numpy as np
    t = RenderTables()
    call(t)
    E, LE, nO = brick_size + 1, brick_size // lighting_divisor + 1, (brick_size // 8) ** 3
    out = {
        "boxes": np.zeros((max(1, t.box_count), 8), np.int32),
        "page_table": np.zeros(max(1, t.page_count), np.int32),
        "bricks": np.zeros(max(1, t.brick_count), np.int32),
        "field": np.zeros((max(1, t.brick_count), E ** 3, 4), np.float32),
        "light": np.zeros((max(1, t.brick_count), LE ** 3, 4), np.float32),
        "occupancy": np.zeros((max(1, t.brick_count), nO), np.uint8),
        "blackbody": np.zeros((512, 4), np.float32),
    }
    if smoke_divisor > 1:
        out["smoke"] = np.zeros((max(1, t.brick_count), (brick_size // smoke_divisor + 1) ** 3), np.float32)
    u = RenderTables()
    for k, a in out.items():
        ct = {np.int32: C.c_int32, np.float32: C.c_float, np.uint8: C.c_uint8}[a.dtype.type]
        setattr(u, k, a.ctypes.data_as(C.POINTER(ct)))
    u.box_capacity, u.page_capacity, u.brick_capacity = t.box_count, t.page_count, t.brick_count
    call(u)
    out["boxes"] = out["boxes"][:u.box_count]
    out["page_table"] = out["page_table"][:u.page_count]
    for k in ("bricks", "field", "light", "occupancy", "smoke"):
        if k in out:
            out[k] = out[k][:u.brick_count]
    out["light_work"] = u.light_work
    return out


def camera(position, target, fov_degrees: float, width: int, height: int, near: float = 0.05,
           far: float = 250.0, up=(0.0, 1.0, 0.0)) -> Camera:
    """A three.js PerspectiveCamera at `position` looking at `target`, with WebGPU's
    [0, 1] clip depth: inverse_view_projection = matrixWorld x projectionMatrixInverse."""
    import numpy as np
    eye, at, upv = (np.asarray(v, dtype=np.float64) for v in (position, target, up))
    z = eye - at
    z /= np.linalg.norm(z)
    x = np.cross(upv, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    world = np.eye(4)
    world[:3, 0], world[:3, 1], world[:3, 2], world[:3, 3] = x, y, z, eye
    top = near * np.tan(np.radians(0.5 * fov_degrees))
    right = top * (width / height)
    proj = np.zeros((4, 4))
    proj[0, 0] = near / right
    proj[1, 1] = near / top
    proj[2, 2] = -far / (far - near)
    proj[2, 3] = -far * near / (far - near)
    proj[3, 2] = -1.0
    inv = world @ np.linalg.inv(proj)
    cam = Camera()
    cam.inverse_view_projection = (C.c_double * 16)(*inv.T.reshape(-1))  # column-major
    cam.position = Vec3(*eye)
    cam.width, cam.height = int(width), int(height)
    return cam


# ---- plain-data frame (SimulationFrame, src/engine/types.ts:235-311) --------------------

DEFAULT_RAMPS = {  # DEFAULT_FLAME_SETTINGS.ramps (src/engine/flame.ts:14-38)
    "smoke": [(0, 0), (0.2, 1), (0.5, 0), (1, 0)],
    "heat": [(0, 0), (0.2, 1), (1, 1)],
    "expansion": [(0, 0), (0.5, 0), (0.75, 1), (1, 1)],
}


@dataclass
class SourceData:
    id: int
    shape: str = "sphere"
    position: Sequence[float] = (0.0, 0.0, 0.0)
    size: Sequence[float] = (0.3, 0.3, 0.3)
    flame: float = 0.0
    heat_rate: float = 0.0
    smoke_rate: float = 0.0
    fuel_rate: float = 0.0
    velocity: Sequence[float] = (0.0, 0.0, 0.0)
    velocity_response: float = 0.0
    outward_speed: float = 0.0
    variation: Optional[Sequence[float]] = None  # (period, strength)
    # Mesh sources: world-space surface samples, rows (x, y, z, weight), and the surface area.
    samples: Optional[np.ndarray] = None
    surface_area: float = 0.0


@dataclass
class ForceData:
    type: str
    vector: Sequence[float] = (0.0, 1.0, 0.0)
    center: Sequence[float] = (0.0, 0.0, 0.0)
    strength: float = 0.0
    scale: float = 1.0
    lift: float = 0.0
    inward: float = 0.0
    sources: Optional[List[SourceData]] = None


@dataclass
class ColliderData:
    shape: str
    position: Sequence[float]
    size: Sequence[float]
    axes: Sequence[Sequence[float]] = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    # Mesh colliders: (occupancy (nz, ny, nx) uint32, lo cell, (nx, ny, nz)) of a box of field cells.
    occupancy: Optional[tuple] = None


@dataclass
class FrameData:
    buoyancy: float = 1.6
    smoke_weight: float = 0.12
    velocity_damping: float = 0.1
    cooling: float = 0.62
    dissipation: float = 0.35
    vorticity: float = 3.2
    seed: float = 1
    lifespan: float = 1.2
    flame_smoke_rate: float = 2.0
    flame_heat_rate: float = 4.0
    expansion_rate: float = 1.0
    ramps: dict = field(default_factory=lambda: dict(DEFAULT_RAMPS))
    fuel: Optional[Sequence] = None  # (enabled, ignition_heat, burn_rate)
    sources: List[SourceData] = field(default_factory=list)
    forces: List[ForceData] = field(default_factory=list)
    colliders: List[ColliderData] = field(default_factory=list)


def _source(s: SourceData, keep: Optional[list] = None) -> Source:
    out = Source()
    out.id = s.id
    out.shape = SHAPES[s.shape]
    out.position = Vec3(*s.position)
    out.size = Vec3(*s.size)
    out.flame, out.heat_rate, out.smoke_rate, out.fuel_rate = (
        s.flame, s.heat_rate, s.smoke_rate, s.fuel_rate)
    out.velocity = Vec3(*s.velocity)
    out.velocity_response, out.outward_speed = s.velocity_response, s.outward_speed
    if s.variation is not None:
        out.has_variation = 1
        out.variation_period, out.variation_strength = s.variation
    if s.samples is not None:
        arr = np.ascontiguousarray(s.samples, np.float64).reshape(-1, 4)
        if keep is not None:
            keep.append(arr)
        out.samples = arr.ctypes.data_as(C.POINTER(C.c_double))
        out.sample_count = len(arr)
        out.surface_area = s.surface_area
    return out


class PackedFrame:
    """An fp_frame and the arrays it points into, kept alive together."""

    def __init__(self, f: FrameData):
        fr = Frame()
        fr.simulation = Settings(f.buoyancy, f.smoke_weight, f.velocity_damping, f.cooling,
                                 f.dissipation, f.vorticity, f.seed)
        fl = Flame()
        fl.lifespan, fl.smoke_rate, fl.heat_rate, fl.expansion_rate = (
            f.lifespan, f.flame_smoke_rate, f.flame_heat_rate, f.expansion_rate)
        for ch, name in enumerate(("smoke", "heat", "expansion")):
            knots = f.ramps[name]
            fl.knot_count[ch] = len(knots)
            for i, (x, y) in enumerate(knots[:MAX_FLAME_KNOTS]):
                fl.knots[ch][i][0] = x
                fl.knots[ch][i][1] = y
        fr.combustion = fl
        if f.fuel is not None:
            fr.has_fuel = 1
            fr.fuel = Fuel(int(bool(f.fuel[0])), f.fuel[1], f.fuel[2])
        self.samples = []  # the mesh samples the sources point into
        self.sources = (Source * max(1, len(f.sources)))(*[_source(s, self.samples) for s in f.sources])
        fr.source_count = len(f.sources)
        fr.sources = self.sources
        self.force_sources = []
        forces = []
        for fd in f.forces:
            fo = Force()
            fo.type = FORCE_TYPES[fd.type]
            fo.vector = Vec3(*fd.vector)
            fo.center = Vec3(*fd.center)
            fo.strength, fo.scale, fo.lift, fo.inward = fd.strength, fd.scale, fd.lift, fd.inward
            if fd.sources is not None:
                fo.targeted = 1
                arr = (Source * max(1, len(fd.sources)))(*[_source(s, self.samples) for s in fd.sources])
                self.force_sources.append(arr)
                fo.source_count = len(fd.sources)
                fo.sources = arr
            forces.append(fo)
        self.forces = (Force * max(1, len(forces)))(*forces)
        fr.force_count = len(forces)
        fr.forces = self.forces
        cols = []
        for cd in f.colliders:
            co = Collider()
            co.shape = SHAPES[cd.shape]
            co.position = Vec3(*cd.position)
            co.size = Vec3(*cd.size)
            for k in range(3):
                for a in range(3):
                    co.axes[k][a] = cd.axes[k][a]
            if cd.occupancy is not None:
                mask, lo, cells = cd.occupancy
                arr = np.ascontiguousarray(mask, np.uint32).reshape(-1)
                if arr.size != cells[0] * cells[1] * cells[2]:
                    raise ValueError("Mesh collider occupancy does not match its size.")
                self.samples.append(arr)
                co.occupancy = arr.ctypes.data_as(C.POINTER(C.c_uint32))
                for a in range(3):
                    co.lo[a] = int(lo[a])
                    co.cells[a] = int(cells[a])
            cols.append(co)
        self.colliders = (Collider * max(1, len(cols)))(*cols)
        fr.collider_count = len(cols)
        fr.colliders = self.colliders
        self.frame = fr

    def pointer(self):
        return C.byref(self.frame)


# ---- the HIP library ------------------------------------------------------------------

_lib = None


# Pressure policies (include/firepro.h FP_PRESSURE_*).
PRESSURE_MODES = {"vcycle": 0, "fp32": 1, "bf16": 2, "fp16": 3}


class FireProError(RuntimeError):
    pass


def load(path: str = LIBRARY):
    """Load libfirepro.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FireProError(
            f"{path} is missing: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    P = C.POINTER
    sigs = {
        "fp_last_error": (C.c_char_p, []),
        "fp_version": (C.c_char_p, []),
        "fp_sim_create": (C.c_int, [P(Grid), P(C.c_void_p)]),
        "fp_sim_step": (C.c_int, [C.c_void_p, P(Frame), C.c_double]),
        "fp_sim_reset": (C.c_int, [C.c_void_p]),
        "fp_sim_destroy": (None, [C.c_void_p]),
        "fp_sim_set_cutoff": (C.c_int, [C.c_void_p, C.c_double]),
        "fp_sim_set_budget": (C.c_int, [C.c_void_p, C.c_double]),
        "fp_sim_set_ground": (C.c_int, [C.c_void_p, C.c_int32]),
        "fp_sim_stats": (C.c_int, [C.c_void_p, P(Stats)]),
        "fp_sim_read_bricks": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_int32)]),
        "fp_sim_read_tables": (C.c_int, [C.c_void_p, P(Tables)]),
        "fp_sim_read_state": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_float),
                                        P(C.c_float), P(C.c_float), C.c_int32]),
        "fp_sim_write_state": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float)]),
        "fp_sim_read_smoke": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_float), C.c_int32]),
        "fp_sim_show_solver_field": (C.c_int, [C.c_void_p, C.c_int32]),
        "fp_sim_read_solver_field": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_float), C.c_int32]),
        "fp_sim_render_bricks": (C.c_int, [C.c_void_p, P(RenderBricks)]),
        "fp_sim_stream": (C.c_void_p, [C.c_void_p]),
        "fp_sim_profile": (C.c_int, [C.c_void_p, C.c_int32]),
        "fp_sim_profile_read": (C.c_int, [C.c_void_p, P(C.c_char_p), P(C.c_double),
                                          P(C.c_int32)]),
        "fp_batch_create": (C.c_int, [P(C.c_void_p)]),
        "fp_batch_destroy": (None, [C.c_void_p]),
        "fp_batch_step": (C.c_int, [C.c_void_p, P(C.c_void_p), P(Frame), C.c_int32, C.c_double]),
        "fp_batch_last_ms": (C.c_int, [C.c_void_p, P(C.c_double)]),
        "fp_batch_phases": (C.c_int, [C.c_void_p, P(C.c_double)]),
        "fp_batch_use_graphs": (C.c_int, [C.c_void_p, C.c_int32]),
        "fp_sim_set_pressure": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32]),
        "fp_sim_pressure_report": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_double)]),
        "fp_sim_render": (C.c_int, [C.c_void_p, P(Camera), P(RenderSettings), P(C.c_float)]),
        "fp_sim_scene_lights": (C.c_int, [C.c_void_p, P(RenderSettings), P(C.c_double), C.c_int32, P(C.c_float)]),
        "fp_sim_scene_lights_noalloc": (C.c_int, [C.c_void_p, P(RenderSettings), P(C.c_double), C.c_int32,
                                                  P(C.c_float)]),
        "fp_sim_render_device": (C.c_int, [C.c_void_p, P(Camera), P(RenderSettings), C.c_void_p]),
        "fp_sim_read_render": (C.c_int, [C.c_void_p, P(RenderTables)]),
        "fp_sim_render_stats": (C.c_int, [C.c_void_p, P(C.c_uint64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, lib=None):
    if status != FP_OK:
        lib = lib or _lib
        msg = lib.fp_last_error().decode() if lib is not None else ""
        if status == FP_EINVAL:
            raise ValueError(msg)
        raise FireProError(f"firepro status {status}: {msg}")
