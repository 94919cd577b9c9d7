"""HIP render path against the FP32 restatement, bit for bit.

The render preparation (page tables, smoothed render field, block occupancy, incident
light on VolumeLighting.encode's schedule) and the camera march evaluate the same float
expression trees as the restatement, so images and tables compare with equality.
"""
import numpy as np
import pytest

from lockstep import apply, make_engines

pytestmark = pytest.mark.gpu

RENDER_KEYS = ("boxes", "page_table", "bricks", "field", "light", "occupancy", "blackbody")


def _abi():
    from firepro import abi
    return abi


def render_mismatches(gpu, cpu, cam, rs, label):
    img_g = gpu.render(cam, rs)
    img_c = cpu.render(cam, rs)
    out = []
    ta, tb = gpu.read_render(), cpu.read_render()
    if ta["light_work"] != tb["light_work"]:
        out.append(f"{label}: light work {ta['light_work']} != {tb['light_work']}")
    for k in RENDER_KEYS + (("smoke",) if "smoke" in ta else ()):
        if ta[k].shape != tb[k].shape:
            out.append(f"{label}: {k} shape {ta[k].shape} != {tb[k].shape}")
        elif not np.array_equal(ta[k], tb[k]):
            d = np.abs(ta[k].astype(np.float64) - tb[k].astype(np.float64))
            out.append(f"{label}: {k} max |diff| {d.max():.3e} in {int((d > 0).sum())} of {d.size}")
    if not np.array_equal(img_g, img_c):
        d = np.abs(img_g.astype(np.float64) - img_c.astype(np.float64))
        out.append(f"{label}: image max |diff| {d.max():.3e} in {int((d > 0).sum())} of {d.size}")
    # The marches take the same samples: positions looked up and samples shaded, by the
    # half-resolution slices and by the full-resolution march.
    sg = gpu.render_stats()
    got = [sg["slice_visits"], sg["slice_shades"], sg["march_visits"], sg["march_shades"]]
    want = [int(x) for x in cpu.render_counts()]
    if got != want:
        out.append(f"{label}: sample counts {got} != {want}")
    elif img_c[..., 3].max() > 0:
        # With half resolution and no scene depth the full-resolution pass resumes from the
        # last slice, past which nothing is left to sample.
        assert want[1] + want[3] > 0 and want[0] >= want[1] and want[2] >= want[3]
    return out, img_g


def test_campfire_render_bit_exact():
    """The campfire after 60 steps, 128 x 128, cubic reconstruction, sun and shadows."""
    from firepro import scenes
    abi = _abi()
    sc = scenes.campfire(steps=60)
    gpu, cpu = make_engines(sc)
    try:
        for frame, g in sc.frames():
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        cam = abi.camera((-4.12, 1.96, 5.30), (0, 1.25, 0), 44, 128, 128)
        print("case campfire 60 steps 128x128", flush=True)
        msgs, img = render_mismatches(gpu, cpu, cam, abi.render_settings(), "campfire")
        assert msgs == []
        assert img[..., 3].max() > 0.1, "the fire is in view"
    finally:
        gpu.dispose()


def test_half_resolution_tile_order():
    """The slice march takes its tiles in the order the previous frame's costs give (the
    first render in id order): renders of one state and of the next steps stay equal to the
    restatement's, frame after frame."""
    from firepro import scenes
    abi = _abi()
    sc = scenes.campfire(steps=40)
    gpu, cpu = make_engines(sc)
    try:
        frames = list(sc.frames())
        for frame, g in frames[:38]:
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        cam = abi.camera((-4.12, 1.96, 5.30), (0, 1.25, 0), 44, 160, 128)
        rs = abi.render_settings(half_resolution=True)
        for k in range(2):
            print(f"case tile order, same state, render {k}", flush=True)
            msgs, img = render_mismatches(gpu, cpu, cam, rs, f"render {k}")
            assert msgs == []
        for frame, g in frames[38:]:
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            print("case tile order, next step", flush=True)
            msgs, img = render_mismatches(gpu, cpu, cam, rs, "next step")
            assert msgs == []
        assert img[..., 3].max() > 0.1, "the fire is in view"
    finally:
        gpu.dispose()


def test_render_schedule_every_step():
    """A render after every step: lighting alternates between whole builds and the slots
    given out in between (lightFresh); settings changes rebuild it."""
    from firepro import scenes
    abi = _abi()
    sc = scenes.small(steps=9, brick=8, divisor=2)
    gpu, cpu = make_engines(sc)
    cam = abi.camera((3.0, 1.5, 3.5), (0.0, 0.6, 0.0), 50, 48, 40)
    works = []
    try:
        for k, (frame, g) in enumerate(sc.frames()):
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
            rs = abi.render_settings(filter=("cubic", "quadratic", "trilinear")[k % 3])
            print(f"case small step {k}", flush=True)
            msgs, _ = render_mismatches(gpu, cpu, cam, rs, f"step {k}")
            assert msgs == []
            works.append(gpu.read_render()["light_work"])
        # Same step, new settings: only the light is rebuilt.
        msgs, _ = render_mismatches(gpu, cpu, cam, abi.render_settings(shadow_density=1.0), "settings")
        assert msgs == []
        assert gpu.read_render()["light_work"] == 1
    finally:
        gpu.dispose()
    assert 2 in works and 1 in works, works


@pytest.mark.parametrize("divisor,filt,sun,half", [(1, "cubic", 3.0, False), (2, "quadratic", 0.0, True),
                                                   (4, "trilinear", 1.0, False), (4, "cubic", 3.0, True)])
def test_render_options(divisor, filt, sun, half):
    from firepro import scenes
    abi = _abi()
    sc = scenes.small(steps=6, brick=16, divisor=2)
    gpu, cpu = make_engines(sc)
    try:
        for frame, g in sc.frames():
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        cam = abi.camera((2.5, 2.0, -3.0), (0.0, 0.5, 0.0), 40, 41, 33)
        rs = abi.render_settings(lighting_divisor=divisor, filter=filt, sun_intensity=sun,
                                 ambient=(0.05, 0.06, 0.08), ray_steps=64, half_resolution=half)
        print(f"case options divisor={divisor} filter={filt} sun={sun} half={half}", flush=True)
        msgs, _ = render_mismatches(gpu, cpu, cam, rs, "options")
        assert msgs == []
    finally:
        gpu.dispose()


def test_explosions_render_many_boxes():
    """Bursts far apart: several boxes, ray traversal from box to box."""
    from firepro import scenes
    abi = _abi()
    sc = scenes.explosions(steps=45)  # bursts at steps 28 and 42, in two clusters
    gpu, cpu = make_engines(sc)
    try:
        for frame, g in sc.frames():
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        cam = abi.camera((0.0, 10.0, 30.0), (0.0, 10.0, 0.0), 60, 64, 48)
        print("case explosions 45 steps", flush=True)
        msgs, img = render_mismatches(gpu, cpu, cam, abi.render_settings(), "explosions")
        assert msgs == []
        assert len(gpu.read_render()["boxes"]) >= 2
        assert img[..., 3].max() > 0
    finally:
        gpu.dispose()


def _hip_runtime():
    """The HIP runtime libfirepro.so itself uses, for device buffers in tests."""
    import ctypes as C
    from firepro import abi
    abi.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libfirepro.so did not load a HIP runtime"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def test_render_device_matches_host():
    """The device-output entry point writes the same image as the host one."""
    import ctypes as C
    from firepro import scenes
    from firepro.engine import FluidSimulation
    abi = _abi()
    sc = scenes.campfire(steps=20)
    sim = FluidSimulation.create(sc.grid())
    hip = _hip_runtime()
    ptr = C.c_void_p()
    try:
        for frame, g in sc.frames():
            sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
            sim.step(frame)
        cam = abi.camera((-4.12, 1.96, 5.30), (0, 1.25, 0), 44, 64, 48)
        rs = abi.render_settings()
        host = sim.render(cam, rs)
        nbytes = host.nbytes
        assert hip.hipMalloc(C.byref(ptr), nbytes) == 0
        sim.render_device(cam, rs, ptr.value)
        assert hip.hipDeviceSynchronize() == 0
        dev = np.zeros_like(host)
        assert hip.hipMemcpy(dev.ctypes.data, ptr, nbytes, 2) == 0  # device to host
        assert np.array_equal(dev, host)
    finally:
        if ptr.value:
            hip.hipFree(ptr)
        sim.dispose()


def test_render_rejects_bad_settings():
    from firepro import scenes
    from firepro.engine import FluidSimulation
    abi = _abi()
    sim = FluidSimulation.create(scenes.campfire(steps=1).grid())
    try:
        cam = abi.camera((0, 1, 4), (0, 1, 0), 45, 8, 8)
        with pytest.raises(ValueError, match="lightingDivisor"):
            sim.render(cam, abi.render_settings(lighting_divisor=3))
        with pytest.raises(ValueError, match="filter"):
            sim.render(cam, abi.render_settings(filter=5))
        bad = abi.camera((0, 1, 4), (0, 1, 0), 45, 8, 8)
        bad.width = 0
        with pytest.raises(ValueError, match="pixels"):
            sim.render(bad, abi.render_settings())
        # An empty simulation renders nothing.
        img = sim.render(cam, abi.render_settings())
        assert not img.any()
    finally:
        sim.dispose()


@pytest.mark.parametrize("brick,smoke,steps", [(16, 2, 40), (8, 4, 12), (16, 4, 12)])
def test_coarse_smoke_render_bit_exact(brick, smoke, steps):
    """Rendering with smoke on its own grid: the fields as they are, the smoke smoothed on
    its lattice, occupancy from glow and smoke donors, light and march reading the coarse
    smoke, bit-exact; the campfire (brick 16) or the small scene."""
    from firepro import scenes
    abi = _abi()
    sc = (scenes.campfire(steps=steps, smoke=smoke) if brick == 16 and smoke == 2
          else scenes.small(steps=steps, brick=brick, divisor=2, smoke=smoke))
    gpu, cpu = make_engines(sc)
    try:
        for frame, g in sc.frames():
            apply(gpu, cpu, g)
            gpu.step(frame)
            cpu.step(frame)
        cam = abi.camera((-3.0, 1.6, 4.0), (0, 0.8, 0), 44, 96, 96)
        print(f"case coarse render brick={brick} smoke={smoke}", flush=True)
        msgs, img = render_mismatches(gpu, cpu, cam, abi.render_settings(), "coarse")
        assert msgs == []
        assert img[..., 3].max() > 0.05, "the volume is in view"
        anchors = [(0.0, 0.3, 0.0), (0.6, 0.4, -0.2)]
        a, b = gpu.scene_lights(anchors), cpu.scene_lights(anchors)
        assert np.array_equal(a, b), (a, b)
    finally:
        gpu.dispose()


def test_scene_lights_bit_exact():
    """Scene-light totals (fp_sim_scene_lights) equal the restatement's, bit for bit: the
    campfire after 40 steps with an anchor at the fire, one above it and one far away, then
    the small scene's bursts with anchors at its sources."""
    from firepro import scenes
    abi = _abi()
    rs = abi.render_settings()
    for sc, anchors in [(scenes.campfire(steps=40), [(0.0, 0.2, 0.0), (0.0, 2.5, 0.0), (6.0, 0.0, 6.0)]),
                        (scenes.small(steps=8, brick=8, divisor=2), [(0.0, 0.4, 0.0), (0.6, 0.4, -0.3)])]:
        gpu, cpu = make_engines(sc)
        try:
            for frame, g in sc.frames():
                apply(gpu, cpu, g)
                gpu.step(frame)
                cpu.step(frame)
            print(f"case scene lights {len(anchors)} anchors", flush=True)
            a, b = gpu.scene_lights(anchors, rs), cpu.scene_lights(anchors, rs)
            assert np.array_equal(a, b), (a, b)
            assert a[:, 0].max() > 0, "flames are lit"
        finally:
            gpu.dispose()


def test_device_render_then_write_state_waits():
    """fp_sim_render_device returns with its kernels queued. An entry point that replaces
    the fields they read (fp_sim_write_state) waits for them first, so the render field the
    next render reuses (same step: no new smoothing) is the one made from the old fields:
    the sequence device render, write, render equals host render, write, render."""
    from firepro import scenes
    from firepro.engine import FluidSimulation
    abi = _abi()
    sc = scenes.campfire(steps=30)
    cam = abi.camera((-4.12, 1.96, 5.30), (0, 1.25, 0), 44, 192, 160)
    rs = abi.render_settings(half_resolution=False)
    out = []
    for device in (True, False):
        sim = FluidSimulation.create(sc.grid())
        buf = abi.DeviceBuffer(cam.width * cam.height * 16)
        try:
            for frame, g in sc.frames():
                sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
                sim.step(frame)
            state = sim.read_state()
            if device:
                sim.render_device(cam, rs, buf.ptr.value)
            else:
                first = sim.render(cam, rs)
            print(f"case device render then write_state, device={device}", flush=True)
            sim.write_state(np.ascontiguousarray(state["fields"][::-1]), None)
            second = sim.render(cam, rs)
            if device:
                first = np.frombuffer(buf.read(), np.float32).reshape(cam.height, cam.width, 4).copy()
            out.append((first, second))
        finally:
            buf.close()
            sim.dispose()
    (a1, a2), (b1, b2) = out
    assert np.array_equal(a1, b1)
    assert np.array_equal(a2, b2)
    assert a1[..., 3].max() > 0.1, "the fire is in view"
