"""ORACLE — test infrastructure only: the CPU restatement (liboracle.so) behind the same
interface as firepro.engine.FluidSimulation. Only tests, __graft_entry__.smoke() and
bench.py's CPU-baseline leg may import it, as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "threejs-fire-pro_amd"))
from firepro import abi  # noqa: E402

LIBRARY = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIBRARY


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        build()
    lib = C.CDLL(LIBRARY)
    P = C.POINTER
    v = C.c_void_p
    sigs = {
        "orc_last_error": (C.c_char_p, []),
        "orc_create": (C.c_int, [P(abi.Grid), C.c_int, P(v)]),
        "orc_destroy": (None, [v]),
        "orc_step": (C.c_int, [v, P(abi.Frame), C.c_double]),
        "orc_reset": (C.c_int, [v]),
        "orc_stats": (C.c_int, [v, P(abi.Stats)]),
        "orc_tables": (C.c_int, [v, P(abi.Tables)]),
        "orc_state": (C.c_int, [v, P(C.c_int32), P(C.c_float), P(C.c_float), P(C.c_float),
                                P(C.c_float), C.c_int]),
        "orc_state64": (C.c_int, [v, P(C.c_int32), P(C.c_double), P(C.c_double),
                                  P(C.c_double), P(C.c_double), C.c_int]),
        "orc_write_state": (C.c_int, [v, P(C.c_float), P(C.c_float)]),
        "orc_smoke": (C.c_int, [v, P(C.c_int32), P(C.c_float), C.c_int]),
        "orc_set_cutoff": (C.c_int, [v, C.c_double]),
        "orc_set_budget": (C.c_int, [v, C.c_double]),
        "orc_set_ground": (C.c_int, [v, C.c_int]),
        "orc_noise": (C.c_int, [v, P(C.c_float)]),
        "orc_content_tiles": (C.c_int, [v, P(C.c_int32), P(C.c_int)]),
        "orc_write_activity": (C.c_int, [v, P(C.c_uint32), C.c_int, P(C.c_int32), C.c_int]),
        "orc_read_activity": (C.c_int, [v, P(C.c_uint32), C.c_int]),
        "orc_fexp": (C.c_float, [C.c_float]),
        "orc_round16": (C.c_float, [C.c_int, C.c_float]),
        "orc_tile_tables": (C.c_int, [v, P(C.c_int32), P(C.c_int32), C.c_int, C.c_int]),
        "orc_brick_coords": (C.c_int, [v, P(C.c_int32), C.c_int, P(C.c_int32)]),
        "orc_rhs": (C.c_int, [v, P(C.c_double), C.c_int]),
        "orc_render": (C.c_int, [v, P(abi.Camera), P(abi.RenderSettings), P(C.c_float)]),
        "orc_scene_lights": (C.c_int, [v, P(abi.RenderSettings), P(C.c_double), C.c_int, P(C.c_float)]),
        "orc_read_render": (C.c_int, [v, P(abi.RenderTables)]),
        "orc_render_counts": (C.c_int, [v, P(C.c_uint64)]),
        "orc_render_probe": (C.c_int, [v, P(abi.Camera), P(abi.RenderSettings), C.c_int,
                                       P(C.c_float), P(C.c_float)]),
        "orc_blackbody": (None, [P(C.c_float)]),
        "orc_set_pressure": (C.c_int, [v, C.c_int, C.c_double, C.c_int]),
        "orc_pressure_report": (C.c_int, [v, P(C.c_int), P(C.c_double)]),
        "orc_ramp_integral": (C.c_double, [P(C.c_double), P(C.c_int), C.c_int, C.c_double, C.c_double]),
        "orc_evaluate_flame": (None, [P(C.c_double), P(C.c_int)] + [C.c_double] * 6 + [P(C.c_double)]),
        "orc_emitter_weight": (C.c_double, [P(C.c_double), P(C.c_double), C.c_int, P(C.c_double)]),
        "orc_burst_perlin": (C.c_double, [C.c_uint32, C.c_double, C.c_double, C.c_double]),
        "orc_time_steps": (C.c_double, [v, P(abi.Frame), C.c_double, C.c_int]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_stage": (C.c_int, [v, C.c_int, P(C.c_double), C.c_int]),
        "orc_force_probe": (C.c_int, [v, P(C.c_int32), C.c_int, P(C.c_double)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status, lib):
    if status != 0:
        msg = lib.orc_last_error().decode()
        if status == abi.FP_EINVAL:
            raise ValueError(msg)
        raise RuntimeError(msg)


class OracleSimulation:
    """CPU restatement of one simulation; bits = 32 (float twin) or 64 (FP64 reference)."""

    def __init__(self, voxel_size, velocity_divisor=1, brick_size=16, smoke_divisor=1,
                 scalar_maccormack=True, bits=32):
        self._lib = load()
        g = abi.Grid(voxel_size, velocity_divisor, brick_size, smoke_divisor,
                     int(bool(scalar_maccormack)))
        h = C.c_void_p()
        _check(self._lib.orc_create(C.byref(g), bits, C.byref(h)), self._lib)
        self._h = h
        self.bits = bits
        self.brick_size = brick_size
        self.brick_cells = brick_size ** 3
        self.velocity_cells = (brick_size // velocity_divisor) ** 3
        self.smoke_divisor = smoke_divisor
        self.smoke_cells = (brick_size // smoke_divisor) ** 3

    @classmethod
    def create(cls, grid: dict, bits=32):
        return cls(grid["voxelSize"], grid.get("velocityDivisor", 1), grid.get("brickSize", 16),
                   grid.get("smokeDivisor", 1), grid.get("scalarMacCormack", True), bits)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def apply_grid(self, g: dict):
        _check(self._lib.orc_set_cutoff(self._h, g["cutoff"]), self._lib)
        _check(self._lib.orc_set_budget(self._h, g["maxVoxels"]), self._lib)
        _check(self._lib.orc_set_ground(self._h, int(bool(g["ground"]))), self._lib)

    def step(self, frame, dt=1 / 60):
        packed = frame if isinstance(frame, abi.PackedFrame) else abi.PackedFrame(frame)
        _check(self._lib.orc_step(self._h, packed.pointer(), dt), self._lib)

    def time_steps(self, frame, dt, steps) -> float:
        packed = frame if isinstance(frame, abi.PackedFrame) else abi.PackedFrame(frame)
        return self._lib.orc_time_steps(self._h, packed.pointer(), dt, steps)

    def reset(self):
        _check(self._lib.orc_reset(self._h), self._lib)

    def stats(self):
        s = abi.Stats()
        _check(self._lib.orc_stats(self._h, C.byref(s)), self._lib)
        return s

    def read_tables(self) -> dict:
        s = self.stats()
        n = s.tile_capacity * 64
        out = {
            "pages": np.zeros(n, np.int32), "entries": np.zeros(n, np.int32),
            "records": np.zeros(n * 28, np.int32),
            "free_slots": np.zeros(s.pool_slots, np.uint32),
            "active": np.zeros(n, np.int32), "zeroed": np.zeros(4 * n, np.int32),
        }
        t = abi.Tables()
        for k, v in out.items():
            setattr(t, k, v.ctypes.data_as(C.POINTER(C.c_uint32 if k == "free_slots" else C.c_int32)))
        _check(self._lib.orc_tables(self._h, C.byref(t)), self._lib)
        out["active"] = out["active"][:t.active_count]
        out["zeroed"] = out["zeroed"][:2 * t.zeroed_count].reshape(-1, 2)
        out["records"] = out["records"].reshape(-1, 28)
        out.update(free_count=t.free_count, failures=t.failures, pool_slots=t.pool_slots,
                   brick_count=t.brick_count)
        return out

    def read_state(self, fields=True, velocity=True, pressure=False, expansion=False, dtype=np.float32):
        s = self.stats()
        n = max(1, s.active_bricks)
        ids = np.zeros(n, np.int32)
        f = np.zeros((n, self.brick_cells, 4), dtype) if fields else None
        v = np.zeros((n, 3, self.velocity_cells), dtype) if velocity else None
        p = np.zeros((n, self.velocity_cells), dtype) if pressure else None
        e = np.zeros((n, self.brick_cells), dtype) if expansion else None
        ct = C.c_float if dtype == np.float32 else C.c_double
        ptr = lambda a: a.ctypes.data_as(C.POINTER(ct)) if a is not None else None
        fn = self._lib.orc_state if dtype == np.float32 else self._lib.orc_state64
        count = fn(self._h, ids.ctypes.data_as(C.POINTER(C.c_int32)), ptr(f), ptr(v), ptr(p), ptr(e), n)
        out = {"ids": ids[:count]}
        for k, a in (("fields", f), ("velocity", v), ("pressure", p), ("expansion", e)):
            if a is not None:
                out[k] = a[:count]
        return out

    def read_smoke(self) -> dict:
        """The coarse smoke density of each active brick (smokeDivisor 2 or 4)."""
        n = max(1, self.stats().active_bricks)
        ids = np.zeros(n, np.int32)
        out = np.zeros((n, self.smoke_cells), np.float32)
        count = self._lib.orc_smoke(self._h, ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                    out.ctypes.data_as(C.POINTER(C.c_float)), n)
        if count < 0:
            _check(count, self._lib)
        return {"ids": ids[:count], "smoke": out[:count]}

    def write_state(self, fields, velocity):
        fa = np.ascontiguousarray(fields, np.float32) if fields is not None else None
        va = np.ascontiguousarray(velocity, np.float32) if velocity is not None else None
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None
        _check(self._lib.orc_write_state(self._h, ptr(fa), ptr(va)), self._lib)

    def noise(self) -> np.ndarray:
        out = np.zeros(64 * 64 * 64 * 4, np.float32)
        self._lib.orc_noise(self._h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out.reshape(64, 64, 64, 4)

    def content_tiles(self) -> np.ndarray:
        n = C.c_int(0)
        self._lib.orc_content_tiles(self._h, None, C.byref(n))
        out = np.zeros((max(1, n.value), 4), np.int32)
        m = C.c_int(n.value)
        self._lib.orc_content_tiles(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(m))
        return out[:n.value]


    def tile_tables(self):
        s = self.stats()
        cap = s.tile_capacity
        rec = np.zeros(cap * 32, np.int32)
        hsh = np.zeros(4 * 65536, np.int32)
        n = self._lib.orc_tile_tables(self._h, rec.ctypes.data_as(C.POINTER(C.c_int32)),
                                      hsh.ctypes.data_as(C.POINTER(C.c_int32)), cap, 65536)
        assert n > 0
        return rec.reshape(cap, 32), hsh[:4 * n].reshape(n, 4)

    def rhs(self) -> np.ndarray:
        n = max(1, self.stats().active_bricks)
        out = np.zeros((n, self.velocity_cells))
        c = self._lib.orc_rhs(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out[:c]

    def set_pressure(self, mode: int, tolerance: float = 1e-6, max_cycles: int = 20) -> None:
        _check(self._lib.orc_set_pressure(self._h, mode, tolerance, max_cycles), self._lib)

    def pressure_report(self):
        cycles, res = C.c_int(), C.c_double()
        self._lib.orc_pressure_report(self._h, C.byref(cycles), C.byref(res))
        return cycles.value, res.value

    def render(self, camera, settings=None, image=True):
        rs = settings if settings is not None else abi.render_settings()
        img = np.zeros((camera.height, camera.width, 4), np.float32)
        _check(self._lib.orc_render(self._h, C.byref(camera), C.byref(rs),
                                    img.ctypes.data_as(C.POINTER(C.c_float)) if image else None), self._lib)
        self._render = (camera, rs)
        return img

    def render_counts(self) -> np.ndarray:
        """The latest image's sample counts: the half-resolution slices' positions visited and
        samples shaded, then the march's."""
        out = (C.c_uint64 * 4)()
        self._lib.orc_render_counts(self._h, out)
        return np.array(out[:], np.uint64)

    def scene_lights(self, anchors, settings=None) -> np.ndarray:
        """Per anchor the 8 totals of SceneLights (luminance, weighted offset, rgb, 0)."""
        rs = settings if settings is not None else abi.render_settings()
        a = np.ascontiguousarray(np.asarray(anchors, np.float64).reshape(-1, 3))
        out = np.zeros((len(a), 8), np.float32)
        _check(self._lib.orc_scene_lights(self._h, C.byref(rs), a.ctypes.data_as(C.POINTER(C.c_double)), len(a),
                                          out.ctypes.data_as(C.POINTER(C.c_float))), self._lib)
        return out

    def read_render(self) -> dict:
        def call(t):
            if self._lib.orc_read_render(self._h, C.byref(t)) != 0:
                raise RuntimeError("render tables need a single-precision oracle")
        return abi.read_render_tables(call, self.brick_size, self._render[1].lighting_divisor,
                                      self.smoke_divisor)

    def probe(self, what: int, values) -> np.ndarray:
        """A single function of the render restatement on the latest render's state:
        0 reconstruct(cell x, y, z, mode) in box 0; 1 incidentLight(world x, y, z) in box 0;
        2 jitter(pixel x, y, time); 3 fexp2(x)."""
        cam, rs = getattr(self, "_render", (abi.camera((0, 0, 1), (0, 0, 0), 45, 1, 1),
                                            abi.render_settings()))
        inp = (C.c_float * 4)(*[float(v) for v in values] + [0.0] * (4 - len(values)))
        out = (C.c_float * 4)()
        r = self._lib.orc_render_probe(self._h, C.byref(cam), C.byref(rs), what, inp, out)
        if r != 0:
            raise RuntimeError(f"probe {what} failed")
        return np.array(out[:], np.float32)

    STAGES = {"advected_velocity": (0, 3, "v"), "corrected_velocity": (1, 3, "v"),
              "forced_velocity": (2, 3, "v"), "curl": (3, 4, "v"), "predictor": (4, 4, "f"),
              "limiter": (5, 1, "f")}

    def stage(self, name: str) -> np.ndarray:
        """An intermediate array of the last step per active brick (FP64): velocities as
        [brick, 3, nV], the curl as [brick, nV, 4], the scalar predictor as [brick, nF, 4];
        the limiter decisions as [brick, nF] masks (uint8)."""
        which, comps, grid = self.STAGES[name]
        st = self.stats()
        n = st.active_bricks
        cells = self.velocity_cells if grid == "v" else self.brick_cells
        out = np.zeros(max(n, 1) * comps * cells, np.float64)
        got = self._lib.orc_stage(self._h, which, out.ctypes.data_as(C.POINTER(C.c_double)), max(n, 1))
        _check(0 if got >= 0 else got, self._lib)
        out = out[: got * comps * cells]
        if comps == 1:
            return out.reshape(got, cells).astype(np.uint8)
        return out.reshape(got, 3, cells) if comps == 3 else out.reshape(got, cells, comps)

    def force_probe(self, cells) -> np.ndarray:
        """The cell-centred force (confinement, external forces, buoyancy) at velocity cells
        `cells` (integer x, y, z rows) on the last step's fields and curl (FP64)."""
        c = np.ascontiguousarray(np.asarray(cells, np.int32).reshape(-1, 3))
        out = np.zeros((len(c), 3), np.float64)
        self._lib.orc_force_probe(self._h, c.ctypes.data_as(C.POINTER(C.c_int32)), len(c),
                                  out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def brick_coords(self, ids) -> np.ndarray:
        ids = np.ascontiguousarray(ids, np.int32)
        out = np.zeros((len(ids), 3), np.int32)
        self._lib.orc_brick_coords(self._h, ids.ctypes.data_as(C.POINTER(C.c_int32)), len(ids),
                                   out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out


def _ramps(ramps):
    knots = np.zeros(3 * 8 * 2)
    counts = (C.c_int * 3)()
    for ch, name in enumerate(("smoke", "heat", "expansion")):
        counts[ch] = len(ramps[name])
        for i, (x, y) in enumerate(ramps[name]):
            knots[(ch * 8 + i) * 2] = x
            knots[(ch * 8 + i) * 2 + 1] = y
    return knots, counts


def ramp_integral(ramps, ch, lo, hi):
    knots, counts = _ramps(ramps)
    return load().orc_ramp_integral(knots.ctypes.data_as(C.POINTER(C.c_double)), counts, ch, lo, hi)


def evaluate_flame(ramps, dt, lifespan, smoke, heat, expansion, before):
    knots, counts = _ramps(ramps)
    out = np.zeros(4)
    load().orc_evaluate_flame(knots.ctypes.data_as(C.POINTER(C.c_double)), counts, dt, lifespan, smoke,
                              heat, expansion, before, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def emitter_weight(pos, size, box, w):
    arr = lambda v: (C.c_double * 3)(*v)
    return load().orc_emitter_weight(arr(pos), arr(size), int(box), arr(w))


def burst_perlin(seed_bits, x, y, z):
    return load().orc_burst_perlin(seed_bits, x, y, z)


def fexp(x: float) -> float:
    return load().orc_fexp(x)


def blackbody() -> np.ndarray:
    out = np.zeros((512, 4), np.float32)
    load().orc_blackbody(out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def round16(mode: int, x: float) -> float:
    """The multigrid storage rounding of the restatement: 2 bfloat16, 3 binary16."""
    return load().orc_round16(mode, float(x))


def set_threads(n: int = 0) -> int:
    """Host threads for the restatement's per-brick passes and image rows (0 keeps the
    current count); returns the count in effect. Results do not depend on it."""
    return load().orc_set_threads(int(n))
