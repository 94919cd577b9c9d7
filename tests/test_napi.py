"""The Node N-API addon (threejs-fire-pro_amd/napi) over the C ABI, driven from plain
JavaScript as Fire Pro's TypeScript host would drive its engine object
(src/library/FireSimulation.ts:203-209, :517).

On the CPU: the addon loads into Node and rejects invalid grids with the reference's
messages (FluidSimulation.ts:362-380). On the GPU: a scene's SimulationFrames, written as
the reference's plain objects, step through the addon and give the same active bricks,
pool slots, brick list and pressure, bit for bit, as the same frames through the C ABI
from Python.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "threejs-fire-pro_amd", "napi", "firepro.node")
MODULE = os.path.join(ROOT, "threejs-fire-pro_amd", "napi", "fluid_simulation.js")
RUNNER = os.path.join(ROOT, "tests", "napi", "run_scene.js")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed on this host")


def _node(code: str) -> subprocess.CompletedProcess:
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=60)


@needs_node
def test_addon_loads_and_validates_like_the_reference():
    if not os.path.exists(ADDON):
        pytest.skip("the addon is not built")
    code = f"""
const {{ FluidSimulation, version }} = require({json.dumps(MODULE)});
const out = {{ version: version(), errors: [] }};
for (const grid of [{{ voxelSize: 0 }}, {{ voxelSize: 0.05, brickSize: 12 }},
                    {{ voxelSize: 0.05, velocityDivisor: 3 }}, {{ voxelSize: 0.05, smokeDivisor: 3 }},
                    {{ voxelSize: 0.05, scalarMacCormack: 1 }}]) {{
  try {{ new FluidSimulation(grid); out.errors.push(null); }} catch (e) {{ out.errors.push(e.message); }}
}}
console.log(JSON.stringify(out));
"""
    r = _node(code)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert "gfx950" in out["version"]
    assert out["errors"] == [
        "voxelSize must be from 0.0001 to 16 meters.",
        "brickSize must be 8, 16 or 32 fine cells.",
        "velocityDivisor must be 1, 2 or 4.",
        "smokeDivisor must be 1, 2 or 4.",
        "scalarMacCormack must be a boolean.",
    ]


def _source(s) -> dict:
    out = {"id": str(s.id), "shape": s.shape, "position": list(s.position), "size": list(s.size),
           "flame": s.flame, "heatRate": s.heat_rate, "smokeRate": s.smoke_rate, "fuelRate": s.fuel_rate,
           "velocity": list(s.velocity), "velocityResponse": s.velocity_response,
           "outwardSpeed": s.outward_speed}
    if s.variation is not None:
        out["variation"] = {"period": s.variation[0], "strength": s.variation[1]}
    if s.samples is not None:
        out["samples"] = [{"position": [float(r[0]), float(r[1]), float(r[2])], "weight": float(r[3])}
                          for r in np.asarray(s.samples)]
        out["surfaceArea"] = s.surface_area
    return out


def reference_frame(f) -> dict:
    """A frame as the reference's SimulationFrame object (src/engine/types.ts:235-311)."""
    out = {"simulation": {"buoyancy": f.buoyancy, "smokeWeight": f.smoke_weight,
                          "velocityDamping": f.velocity_damping, "cooling": f.cooling,
                          "dissipation": f.dissipation, "vorticity": f.vorticity, "seed": f.seed},
           "combustion": {"lifespan": f.lifespan, "smokeRate": f.flame_smoke_rate, "heatRate": f.flame_heat_rate,
                          "expansionRate": f.expansion_rate,
                          "ramps": {k: [list(p) for p in v] for k, v in f.ramps.items()}},
           "sources": [_source(s) for s in f.sources],
           "forces": [], "colliders": []}
    if f.fuel is not None:
        out["fuel"] = {"enabled": bool(f.fuel[0]), "ignitionHeat": f.fuel[1], "burnRate": f.fuel[2]}
    for fc in f.forces:
        d = {"type": fc.type, "vector": list(fc.vector), "center": list(fc.center), "strength": fc.strength,
             "scale": fc.scale, "lift": fc.lift, "inward": fc.inward}
        if fc.sources is not None:
            d["sources"] = [_source(s) for s in fc.sources]
        out["forces"].append(d)
    for c in f.colliders:
        if c.occupancy is not None:
            mask, lo, cells = c.occupancy
            out["colliders"].append({"shape": "mesh", "occupancy": {
                "mask": [int(x) for x in np.asarray(mask).ravel()], "lo": [int(x) for x in lo],
                "size": [int(x) for x in cells]}})
        else:
            out["colliders"].append({"shape": c.shape, "position": list(c.position), "size": list(c.size),
                                     "axes": [list(a) for a in c.axes]})
    return out


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "meshes"])
def test_addon_steps_like_the_c_abi(name, tmp_path):
    from firepro import scenes
    from firepro.engine import FluidSimulation
    if not os.path.exists(ADDON):
        pytest.fail("the addon is not built")
    sc = scenes.SCENES[name](steps=8)
    grid = sc.grid()
    steps = [{"frame": reference_frame(f), "settings": g} for f, g in sc.frames()]
    path = tmp_path / "scene.json"
    path.write_text(json.dumps({"grid": grid, "steps": steps, "solverField": "pressure"}))
    print(f"case napi {name}", flush=True)
    r = subprocess.run([NODE, RUNNER, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    js = json.loads(r.stdout)
    sim = FluidSimulation.create(grid)
    sim.show_solver_field("pressure")  # as the script does: its view counts in memoryBytes
    try:
        for (frame, g), got in zip(sc.frames(), js["steps"]):
            sim.cutoff, sim.voxel_budget, sim.ground = g["cutoff"], g["maxVoxels"], g["ground"]
            sim.step(frame)
            st = sim.stats()
            assert got["activeBrickCount"] == st.active_bricks
            assert got["poolSlots"] == st.pool_slots
            assert got["poolLimited"] == bool(st.pool_limited)
        bricks = sim.read_bricks()
        assert [tuple(b["min"]) for b in js["bricks"]] == [b[0] for b in bricks]
        state = sim.read_state(pressure=True)
        assert js["solverIds"] == state["ids"].tolist()
        values = np.asarray(js["solverValues"], np.float64).astype(np.float32)
        assert np.array_equal(values, state["pressure"].ravel())
        assert js["memoryBytes"] == sim.stats().memory_bytes
        assert js["meshMemoryBytes"] == sim.mesh_memory_bytes
        assert js["forceMemoryBytes"] == sim.force_memory_bytes
        rb = sim.render_bricks
        assert [tuple(map(tuple, (b["lo"], b["hi"]))) for b in js["renderBricks"]["boxes"]] == rb["boxes"]
        assert js["invalidStep"] == "Simulation step must be in (0, 1/30]. Use fixed substeps."
    finally:
        sim.dispose()
