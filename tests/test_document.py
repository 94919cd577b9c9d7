"""Preset documents (editor/document.ts) and their headless playback
(examples/playback/main.ts), on the twelve presets the reference ships
(tests/golden/presets, copied by tests/golden/make_presets.py)."""
import copy
import glob
import json
import math
import os

import numpy as np
import pytest

from firepro import abi, document

PRESETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "presets")
TEXT_SHAPED = {"flaming-letters"}


def presets():
    out = {}
    for path in sorted(glob.glob(os.path.join(PRESETS, "*.json"))):
        with open(path) as f:
            out[os.path.basename(path)[:-5]] = json.load(f)
    return out


def test_every_preset_validates():
    docs = presets()
    assert len(docs) == 12
    for name, doc in docs.items():
        d = document.validate_document(doc)
        assert d["id"] == name
        sim = d["simulation"]
        assert sim["brickSize"] in (8, 16, 32) and sim["smokeDivisor"] in (1, 2, 4)
        for e in d["emitters"]:
            assert set(e["options"]) == {"active", "emission", "velocity"}
            assert set(e["burst"]) == {"radius", "charge", "outwardSpeed", "variation"}
        for f in d["forces"]:
            if "direction" in f["options"]:
                assert math.isclose(math.hypot(*f["options"]["direction"]), 1.0, rel_tol=1e-12)
    # The resolved defaults of a burst that sets only some of its charge (options.ts:556-563).
    fireball = document.validate_document(docs["fuel-fireball"])
    burst = [e for e in fireball["emitters"] if e["mode"] == "burst"][0]["burst"]
    assert burst["charge"]["fuel"] == 0.62 and burst["variation"] == {"period": 0.4, "strength": 0.55}


@pytest.mark.parametrize("mutate,message", [
    (lambda d: d.update(version=10), "Unsupported preset format"),
    (lambda d: d.update(extra=1), "Simulation.extra: unknown property"),
    (lambda d: d["scene"].update(recipe="volcano"), "unknown built-in scene"),
    (lambda d: d["simulation"].update(brickSize=64), "simulation.brickSize"),
    (lambda d: d["simulation"].update(smokeDivisor=3), "smokeDivisor must be 1, 2 or 4"),
    (lambda d: d["emitters"][0].update(id=d["id"]), "duplicate ID"),
    (lambda d: d["emitters"][0].update(shape="cone"), "Unknown emitter shape"),
    (lambda d: d["emitters"][0].update(shape="letters"), "Letter geometry requires"),
    (lambda d: d["emitters"][0].update(size=[1, 1, 100]), "emitter.size"),
    (lambda d: d["emitters"][0]["options"]["emission"].update(flame=2), "emission.flame"),
    (lambda d: d["forces"][0].update(targets=["nobody"]), "Force targets must be unique"),
    (lambda d: d["forces"][0]["options"].update(type="gravity"), "force"),
    (lambda d: d["colliders"].append({"id": "c", "name": "c", "shape": "cone", "size": [1, 1, 1],
                                      "radius": 1, "position": [0, 0, 0], "rotation": [0, 0, 0]}),
     "Invalid collider shape"),
])
def test_validation_errors(mutate, message):
    """Each check of validateDocument (editor/document.ts:173-270) rejects its case."""
    doc = copy.deepcopy(presets()["fire-tornado"])
    mutate(doc)
    with pytest.raises(ValueError, match=message):
        document.validate_document(doc)


def test_playback_builds_presets():
    """Every preset but the text one builds; bursts fire in the first frame only; emitters
    keep their document positions; the campfire's disk is a mesh source with its surface
    samples; letter emitters are refused."""
    for name, doc in presets().items():
        if name in TEXT_SHAPED:
            with pytest.raises(abi.FireProError, match="mesh surface"):
                document.Playback(doc)
            continue
        play = document.Playback(doc)
        d = play.document
        frames = list(play.frames(3))
        assert len(frames) == 3
        bursts = [e for e in d["emitters"] if e["mode"] == "burst"]
        steady = [e for e in d["emitters"] if e["mode"] == "continuous" and e["options"]["active"]]
        first, second = frames[0][0], frames[1][0]
        assert len(first.sources) == len(steady) + len(bursts)
        assert len(second.sources) == len(steady)
        for e, src in zip(steady, first.sources):
            assert tuple(src.position) == tuple(e["position"])
            if e["shape"] == "sphere":
                assert src.size == (e["radius"],) * 3 and src.samples is None
            else:
                assert src.shape == "mesh" and src.size == (d["simulation"]["voxelSize"],) * 3
                area = 0.5 * 40 * e["radius"] ** 2 * math.sin(2 * math.pi / 40)  # the 40-gon
                assert src.surface_area == pytest.approx(area, rel=1e-6)
                assert len(src.samples) == min(8192, max(256, math.ceil(area / 0.05 ** 2 * 8)))
        for e, src in zip(bursts, first.sources[len(steady):]):
            assert tuple(src.position) == tuple(e["position"])
            assert src.heat_rate == pytest.approx(e["burst"]["charge"]["heat"] * 60)
        assert len(first.forces) == sum(1 for f in d["forces"] if f["options"]["active"])
        assert frames[0][1]["maxVoxels"] == d["simulation"]["grid"]["maxVoxels"]


def test_euler_axes_xyz_order():
    """three.js's XYZ order: R = Rx Ry Rz, so a quarter turn about y takes +x to -z, and a
    combined rotation stays orthonormal with determinant 1."""
    ax = np.array(document.euler_axes([0, math.pi / 2, 0]))
    np.testing.assert_allclose(ax[0], [0, 0, -1], atol=1e-15)
    r = np.array(document.euler_axes([0.3, -1.1, 2.0])).T
    np.testing.assert_allclose(r.T @ r, np.eye(3), atol=1e-12)
    assert np.linalg.det(r) == pytest.approx(1.0)
    # Rz is applied first: rotating +x about z then about y differs from the other order.
    rz_then_ry = np.array(document.euler_axes([0, math.pi / 2, math.pi / 2])).T @ [1, 0, 0]
    np.testing.assert_allclose(rz_then_ry, [0, 1, 0], atol=1e-12)


def test_preset_runs_on_the_oracle():
    """A few steps of the fire tornado (three sphere fires, a vortex and turbulence, coarse
    smoke) on the CPU restatement: bricks come alive and the coarse grid holds smoke."""
    import oracle
    sc = document.preset_scene(presets()["fire-tornado"], 4)
    orc = oracle.OracleSimulation.create(sc.grid())
    for frame, g in sc.frames():
        orc.apply_grid(g)
        orc.step(frame)
    assert orc.stats().active_bricks > 0
    st = orc.read_state()
    assert st["fields"][..., 1].max() > 0  # heat
    assert orc.read_smoke()["smoke"].max() >= 0
