"""FluidSimulation: the engine-layer object backed by the HIP library.

Mirrors the members of src/engine/FluidSimulation.ts that src/library/FireSimulation.ts
calls (create, step, reset, dispose, cutoff/voxelBudget/ground, stats getters, readBricks),
plus the parity surface (integer tables and cell state of the active bricks).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import abi


class FluidSimulation:
    """One sparse-brick fire/smoke simulation resident on the GPU."""

    def __init__(self, voxel_size: float, velocity_divisor: int = 1, brick_size: int = 16,
                 smoke_divisor: int = 1, scalar_maccormack: bool = True):
        self._lib = abi.load()
        grid = abi.Grid(voxel_size, velocity_divisor, brick_size, smoke_divisor,
                        int(bool(scalar_maccormack)))
        handle = C.c_void_p()
        abi.check(self._lib.fp_sim_create(C.byref(grid), C.byref(handle)), self._lib)
        self._handle = handle
        self.voxel_size = voxel_size
        self.velocity_divisor = velocity_divisor
        self.brick_size = brick_size
        self.brick_cells = brick_size ** 3
        self.velocity_cells = (brick_size // velocity_divisor) ** 3
        self.smoke_divisor = smoke_divisor
        self.smoke_cells = (brick_size // smoke_divisor) ** 3

    # FluidSimulation.create (FluidSimulation.ts:419-475) is the constructor above.
    @classmethod
    def create(cls, grid: dict) -> "FluidSimulation":
        return cls(grid["voxelSize"], grid.get("velocityDivisor", 1), grid.get("brickSize", 16),
                   grid.get("smokeDivisor", 1), grid.get("scalarMacCormack", True))

    def _h(self):
        if self._handle is None:
            raise abi.FireProError("This FluidSimulation has been disposed.")
        return self._handle

    def step(self, frame: "abi.FrameData | abi.PackedFrame", dt: float = 1 / 60) -> None:
        packed = frame if isinstance(frame, abi.PackedFrame) else abi.PackedFrame(frame)
        abi.check(self._lib.fp_sim_step(self._h(), packed.pointer(), dt), self._lib)

    def reset(self) -> None:
        abi.check(self._lib.fp_sim_reset(self._h()), self._lib)

    def dispose(self) -> None:
        if self._handle is not None:
            self._lib.fp_sim_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

    # Setters FireSimulation.step writes before each step (FireSimulation.ts:457-459).
    @property
    def cutoff(self):
        return self._cutoff

    @cutoff.setter
    def cutoff(self, value: float):
        abi.check(self._lib.fp_sim_set_cutoff(self._h(), value), self._lib)
        self._cutoff = value

    @property
    def voxel_budget(self):
        return self._budget

    @voxel_budget.setter
    def voxel_budget(self, value: float):
        abi.check(self._lib.fp_sim_set_budget(self._h(), value), self._lib)
        self._budget = value

    @property
    def ground(self):
        return self._ground

    @ground.setter
    def ground(self, value: bool):
        abi.check(self._lib.fp_sim_set_ground(self._h(), int(bool(value))), self._lib)
        self._ground = bool(value)

    def stats(self) -> abi.Stats:
        s = abi.Stats()
        abi.check(self._lib.fp_sim_stats(self._h(), C.byref(s)), self._lib)
        return s

    @property
    def active_brick_count(self) -> int:
        return self.stats().active_bricks

    @property
    def pool_slots(self) -> int:
        return self.stats().pool_slots

    @property
    def pool_limited(self) -> bool:
        return bool(self.stats().pool_limited)

    @property
    def time(self) -> float:
        return self.stats().time

    @property
    def memory_bytes(self) -> int:
        return self.stats().memory_bytes

    @property
    def mesh_memory_bytes(self) -> int:
        """meshMemoryBytes (FluidSimulation.ts:849-851)."""
        return self.stats().mesh_memory_bytes

    @property
    def force_memory_bytes(self) -> int:
        """forceMemoryBytes (FluidSimulation.ts:846-848)."""
        return self.stats().force_memory_bytes

    def show_solver_field(self, field: Optional[str]) -> None:
        """showSolverField (FluidSimulation.ts:1542-1549): 'pressure', 'divergence' or None."""
        if field not in abi.SOLVER_FIELDS:
            raise ValueError("solver field must be 'pressure', 'divergence' or None.")
        abi.check(self._lib.fp_sim_show_solver_field(self._h(), abi.SOLVER_FIELDS[field]), self._lib)

    def read_solver_field(self) -> dict:
        """The shown solver field of the latest step per active brick: {ids, values[brick, cell]}."""
        n = max(1, self.stats().active_bricks)
        ids = np.zeros(n, np.int32)
        out = np.zeros((n, self.velocity_cells), np.float32)
        count = self._lib.fp_sim_read_solver_field(self._h(), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   out.ctypes.data_as(C.POINTER(C.c_float)), n)
        if count < 0:
            abi.check(count, self._lib)
        return {"ids": ids[:count], "values": out[:count]}

    @property
    def render_bricks(self) -> dict:
        """renderBricks (FluidSimulation.ts:781-795): version, boxes [(lo, hi)], the tile words
        [slot, 4], and the device addresses of the pages, active and zeroed tables."""
        rb = abi.RenderBricks()
        abi.check(self._lib.fp_sim_render_bricks(self._h(), C.byref(rb)), self._lib)
        boxes = np.zeros((max(1, rb.box_count), 6), np.int32)
        tiles = np.zeros((max(1, rb.tile_count), 4), np.int32)
        rb.boxes = boxes.ctypes.data_as(C.POINTER(C.c_int32))
        rb.tiles = tiles.ctypes.data_as(C.POINTER(C.c_int32))
        rb.box_capacity, rb.tile_capacity = len(boxes), len(tiles)
        abi.check(self._lib.fp_sim_render_bricks(self._h(), C.byref(rb)), self._lib)
        return {"version": rb.version,
                "boxes": [(tuple(b[:3]), tuple(b[3:])) for b in boxes[:rb.box_count]],
                "tiles": tiles[:rb.tile_count], "pages": rb.pages_device, "active": rb.active_device,
                "zeroed": rb.zeroed_device, "brick_count": rb.brick_count, "active_count": rb.active_count,
                "zeroed_count": rb.zeroed_count}

    def read_bricks(self):
        """readBricks (FluidSimulation.ts:1749-1779): each slotted brick's (min, max) corners."""
        s = self.stats()
        cap = C.c_int32(s.tile_capacity * 64)
        mins = np.zeros((cap.value, 3), dtype=np.float64)
        abi.check(self._lib.fp_sim_read_bricks(
            self._h(), mins.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cap)), self._lib)
        size = self.brick_size * self.voxel_size
        return [(tuple(m), tuple(m + size)) for m in mins[:cap.value]]

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
        abi.check(self._lib.fp_sim_read_tables(self._h(), C.byref(t)), self._lib)
        out["active"] = out["active"][:t.active_count]
        out["zeroed"] = out["zeroed"][:2 * t.zeroed_count].reshape(-1, 2)
        out["records"] = out["records"].reshape(-1, 28)
        out.update(free_count=t.free_count, failures=t.failures, pool_slots=t.pool_slots,
                   brick_count=t.brick_count)
        return out

    def read_state(self, fields=True, velocity=True, pressure=False, expansion=False) -> dict:
        s = self.stats()
        n = max(1, s.active_bricks)
        ids = np.zeros(n, np.int32)
        f = np.zeros((n, self.brick_cells, 4), np.float32) if fields else None
        v = np.zeros((n, 3, self.velocity_cells), np.float32) if velocity else None
        p = np.zeros((n, self.velocity_cells), np.float32) if pressure else None
        e = np.zeros((n, self.brick_cells), np.float32) if expansion else None
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None
        count = self._lib.fp_sim_read_state(self._h(), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                            ptr(f), ptr(v), ptr(p), ptr(e), n)
        if count < 0:
            abi.check(count, self._lib)
        out = {"ids": ids[:count]}
        for k, a in (("fields", f), ("velocity", v), ("pressure", p), ("expansion", e)):
            if a is not None:
                out[k] = a[:count]
        return out

    def read_smoke(self) -> dict:
        """The coarse smoke grid's density per active brick (smokeDivisor 2 or 4)."""
        n = max(1, self.stats().active_bricks)
        ids = np.zeros(n, np.int32)
        out = np.zeros((n, self.smoke_cells), np.float32)
        count = self._lib.fp_sim_read_smoke(self._h(), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                            out.ctypes.data_as(C.POINTER(C.c_float)), n)
        if count < 0:
            abi.check(count, self._lib)
        return {"ids": ids[:count], "smoke": out[:count]}

    def write_state(self, fields: Optional[np.ndarray], velocity: Optional[np.ndarray]) -> None:
        ptr = lambda a: (np.ascontiguousarray(a, np.float32).ctypes.data_as(C.POINTER(C.c_float))
                         if a is not None else None)
        fa = np.ascontiguousarray(fields, np.float32) if fields is not None else None
        va = np.ascontiguousarray(velocity, np.float32) if velocity is not None else None
        abi.check(self._lib.fp_sim_write_state(self._h(), ptr(fa), ptr(va)), self._lib)

    # ---- pressure policy ----
    def set_pressure(self, mode, tolerance: float = 1e-6, max_cycles: int = 20) -> None:
        """'vcycle' (the reference's one V-cycle) or defect correction with V-cycle levels
        stored in 'fp32', 'bf16' or 'fp16' (FP32 outer loop, FP64 norms and coarsest level)."""
        m = abi.PRESSURE_MODES[mode] if isinstance(mode, str) else int(mode)
        abi.check(self._lib.fp_sim_set_pressure(self._h(), m, tolerance, max_cycles), self._lib)

    def pressure_report(self):
        """(V-cycles of the latest step, its final relative residual or -1)."""
        cycles, res = C.c_int32(), C.c_double()
        abi.check(self._lib.fp_sim_pressure_report(self._h(), C.byref(cycles), C.byref(res)), self._lib)
        return cycles.value, res.value

    # ---- render (SceneVolumeRenderer.prepare + the fragment march) ----
    def render(self, camera: abi.Camera, settings: Optional[abi.RenderSettings] = None) -> np.ndarray:
        """The premultiplied (radiance rgb, 1 - transmittance) image, (height, width, 4)."""
        rs = settings if settings is not None else abi.render_settings()
        img = np.zeros((camera.height, camera.width, 4), np.float32)
        abi.check(self._lib.fp_sim_render(self._h(), C.byref(camera), C.byref(rs),
                                          img.ctypes.data_as(C.POINTER(C.c_float))), self._lib)
        self._lighting_divisor = rs.lighting_divisor
        return img

    def scene_lights(self, anchors, settings: Optional[abi.RenderSettings] = None) -> np.ndarray:
        """SceneLights' totals of the latest step (fp_sim_scene_lights): per anchor (x, y, z)
        the flames' luminance, their luminance-weighted offset from it, and their rgb."""
        rs = settings if settings is not None else abi.render_settings()
        a = np.ascontiguousarray(np.asarray(anchors, np.float64).reshape(-1, 3))
        out = np.zeros((len(a), 8), np.float32)
        abi.check(self._lib.fp_sim_scene_lights(self._h(), C.byref(rs), a.ctypes.data_as(C.POINTER(C.c_double)),
                                                len(a), out.ctypes.data_as(C.POINTER(C.c_float))), self._lib)
        return out

    def scene_lights_noalloc(self, anchors, settings: Optional[abi.RenderSettings] = None) -> np.ndarray:
        """The same totals without allocating device memory (fp_sim_scene_lights_noalloc):
        at most 64 anchors over at most 65536 active bricks."""
        rs = settings if settings is not None else abi.render_settings()
        a = np.ascontiguousarray(np.asarray(anchors, np.float64).reshape(-1, 3))
        out = np.zeros((len(a), 8), np.float32)
        abi.check(self._lib.fp_sim_scene_lights_noalloc(
            self._h(), C.byref(rs), a.ctypes.data_as(C.POINTER(C.c_double)), len(a),
            out.ctypes.data_as(C.POINTER(C.c_float))), self._lib)
        return out

    def render_device(self, camera: abi.Camera, settings: abi.RenderSettings, pointer: int) -> None:
        """Render into device memory at `pointer` (width * height float4), enqueued on the
        simulation's stream."""
        abi.check(self._lib.fp_sim_render_device(self._h(), C.byref(camera), C.byref(settings),
                                                 C.c_void_p(pointer)), self._lib)
        self._lighting_divisor = settings.lighting_divisor

    def render_stats(self) -> dict:
        """The latest render's sample counts: positions looked up in the occupancy and
        samples shaded, by the half-resolution slices and by the full-resolution march."""
        out = (C.c_uint64 * 4)()
        abi.check(self._lib.fp_sim_render_stats(self._h(), out), self._lib)
        return {"slice_visits": out[0], "slice_shades": out[1], "march_visits": out[2], "march_shades": out[3]}

    def read_render(self) -> dict:
        ld = getattr(self, "_lighting_divisor", 4)
        return abi.read_render_tables(
            lambda t: abi.check(self._lib.fp_sim_read_render(self._h(), C.byref(t)), self._lib),
            self.brick_size, ld, self.smoke_divisor)

    def stream(self) -> int:
        return self._lib.fp_sim_stream(self._h())

    def profile(self, enable) -> None:
        """0 off; 1 (True) every stage and host phase; 2 the dominant stage only."""
        abi.check(self._lib.fp_sim_profile(self._h(), int(enable)), self._lib)

    def profile_read(self) -> dict:
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        count = C.c_int32(64)
        abi.check(self._lib.fp_sim_profile_read(self._h(), names, ms, C.byref(count)), self._lib)
        return {names[i].decode(): ms[i] for i in range(count.value)}


class Batch:
    """Steps many simulations of one brick layout together: one launch per stage over all
    of their bricks (fp_batch_step). Any set of simulations may be stepped on each call."""

    def __init__(self):
        self._lib = abi.load()
        h = C.c_void_p()
        abi.check(self._lib.fp_batch_create(C.byref(h)), self._lib)
        self._handle = h

    def step(self, sims, frames, dt: float = 1 / 60) -> None:
        n = len(sims)
        if n != len(frames):
            raise ValueError("one frame per simulation")
        packed = [f if isinstance(f, abi.PackedFrame) else abi.PackedFrame(f) for f in frames]
        handles = (C.c_void_p * max(1, n))(*[s._h().value for s in sims])
        arr = (abi.Frame * max(1, n))()
        for i, p in enumerate(packed):
            arr[i] = p.frame
        abi.check(self._lib.fp_batch_step(self._handle, handles, arr, n, dt), self._lib)
        self._keep = packed  # the frames' arrays stay alive through the call

    @property
    def last_ms(self) -> float:
        ms = C.c_double()
        abi.check(self._lib.fp_batch_last_ms(self._handle, C.byref(ms)), self._lib)
        return ms.value

    def use_graphs(self, enable: bool) -> None:
        """Replay captured HIP graphs (the default) or launch every kernel each step."""
        abi.check(self._lib.fp_batch_use_graphs(self._handle, int(bool(enable))), self._lib)

    @property
    def phases(self) -> dict:
        ms = (C.c_double * 5)()
        abi.check(self._lib.fp_batch_phases(self._handle, ms), self._lib)
        return dict(zip(("host", "pack", "launch", "wait", "readback"), ms[:]))

    def dispose(self) -> None:
        if self._handle is not None:
            self._lib.fp_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

