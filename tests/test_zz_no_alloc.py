"""The allocation-free scene lights (fp_sim_scene_lights_noalloc), called through the C
ABI, against fp_sim_scene_lights on the same simulation, bit for bit: two calls in a row,
the second with fewer anchors over fewer bricks, and sizes beyond the fixed scratch
refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from lockstep import apply, make_engines

pytestmark = pytest.mark.gpu


def _noalloc(sim, anchors, rs):
    from firepro import abi
    lib = abi.load()
    a = np.ascontiguousarray(np.asarray(anchors, np.float64).reshape(-1, 3))
    out = np.full((len(a), 8), np.nan, np.float32)
    status = lib.fp_sim_scene_lights_noalloc(sim._h(), C.byref(rs), a.ctypes.data_as(C.POINTER(C.c_double)),
                                             len(a), out.ctypes.data_as(C.POINTER(C.c_float)))
    return status, out


def test_noalloc_scene_lights_equal_the_allocating_entry_point():
    from firepro import abi, scenes
    rs = abi.render_settings()
    four = scenes.four_fires(steps=20)
    fires = [tuple(s.position) for s in next(four.frames(1))[0].sources]
    cases = [(four, fires + [(0.0, 6.0, 0.0), (9.0, 0.0, 9.0)]),
             (scenes.small(steps=6, brick=8, divisor=2), [(0.0, 0.4, 0.0), (0.6, 0.4, -0.3)])]
    bricks = []
    for sc, anchors in cases:
        gpu, cpu = make_engines(sc)
        try:
            for frame, g in sc.frames():
                apply(gpu, cpu, g)
                gpu.step(frame)
                cpu.step(frame)
            bricks.append(gpu.stats().active_bricks)
            print(f"case noalloc scene lights: {len(anchors)} anchors, {bricks[-1]} bricks", flush=True)
            status, got = _noalloc(gpu, anchors, rs)
            assert status == abi.FP_OK, abi.load().fp_last_error()
            want = gpu.scene_lights(anchors, rs)
            assert np.array_equal(got, want), (got, want)
            assert np.array_equal(got, cpu.scene_lights(anchors, rs))
            assert got[:, 0].max() > 0, "flames are lit"
            status, _ = _noalloc(gpu, [(0.0, 0.0, 0.0)] * 65, rs)
            assert status == abi.FP_EINVAL
            assert b"0 to 64 anchors" in abi.load().fp_last_error()
        finally:
            gpu.dispose()
    assert bricks[1] < bricks[0], bricks
