"""Fire Pro preset documents (format "fire-pro", version 11) and their headless playback.

`validate_document` is the strict current-format check of editor/document.ts:173-270: the
header, the allowed keys of every object, unique ids, the built-in scene recipes, numeric
ranges, emitter modes and shapes, and force targets. `Playback` builds the simulation the
way examples/playback/main.ts:37-95 does: the simulation options as they stand, each burst
emitter as an explosion fired at the start, each sphere emitter as a sphere source, the
forces with their target emitters, and the colliders. The editor's recipes (scene props,
the moving emitter's wander, the projectiles' launcher) belong to the editor, not to the
document, and are not run.

Box, disk and torus emitters emit from the surface of the mesh playback builds for them
(firepro.mesh). Letter emitters need the editor's font geometry, which is not built, and
`Playback` raises for them.
"""
from __future__ import annotations

import copy
import json
import math
from typing import Dict, Iterator, List, Optional, Tuple

import numpy as np

from . import abi, mesh
from .library import LIMITS, FireSimulation, _direction, _number, resolve_force, resolve_simulation

FORMAT, VERSION = "fire-pro", 11
# editor/document.ts:15-28
RECIPES = ("empty", "flaming-letters", "fireplace", "campfire", "projectiles", "flamethrower",
           "fire-tornado", "magical-fire", "delayed-ignition", "industrial-smoke", "fuel-fireball",
           "moving-emitter")
EMITTER_SIZE = (0.01, 16)  # DOCUMENT_LIMITS.emitter.size (editor/document.ts:150-152)
SHAPES = ("sphere", "box", "disk", "torus", "letters")


def _object(v, path: str, allowed) -> None:
    if not isinstance(v, dict):
        raise ValueError(f"{path}: expected an object.")
    for k in v:
        if k not in allowed:
            raise ValueError(f"{path}.{k}: unknown property.")


def _num(v, path: str, lo=-1e5, hi=1e5) -> None:
    if (not isinstance(v, (int, float)) or isinstance(v, bool) or not math.isfinite(v)
            or v < lo or v > hi):
        raise ValueError(f"{path}: expected a finite number from {lo} to {hi}.")


def _vec(v, path: str, lo=-1e5, hi=1e5) -> None:
    if not isinstance(v, list) or len(v) != 3:
        raise ValueError(f"{path}: expected X, Y, Z.")
    for i, x in enumerate(v):
        _num(x, f"{path}[{i}]", lo, hi)


def _text(v, path: str) -> None:
    if not isinstance(v, str) or not v.strip() or len(v) > 120:
        raise ValueError(f"{path}: use 1-120 characters.")


def _list(v, path: str) -> None:
    if not isinstance(v, list):
        raise ValueError(f"{path}: expected a list.")


def _bool(v, path: str) -> None:
    if not isinstance(v, bool):
        raise ValueError(f"{path}: expected true or false.")


def resolve_emitter_options(options: dict, radius: float) -> dict:
    """resolveEmitter (options.ts:497-555) of a document emitter's options over
    emitterDefaults('emitter'), as a sphere of `radius`; the shape is dropped again."""
    _object(options, "emitter.options", ["active", "emission", "velocity"])
    out = {"active": True, "emission": {"flame": 0, "heatRate": 0, "smokeRate": 0, "fuelRate": 0},
           "velocity": None}
    _number(radius, "shape.radius", *LIMITS["emitter"]["radius"])
    if "active" in options:
        if not isinstance(options["active"], bool):
            raise ValueError("active must be a boolean.")
        out["active"] = options["active"]
    if "emission" in options:
        _object(options["emission"], "emission", ["flame", "heatRate", "smokeRate", "fuelRate"])
        for name, v in options["emission"].items():
            limits = LIMITS["emitter"]["flame" if name == "flame" else "rate"]
            out["emission"][name] = _number(v, f"emission.{name}", *limits)
    if "velocity" in options:
        v = options["velocity"]
        if v is None:
            out["velocity"] = None
        else:
            _object(v, "velocity", ["direction", "speed"])
            out["velocity"] = {
                "direction": _direction(v.get("direction", [0, 1, 0]), "velocity.direction"),
                "speed": _number(v.get("speed", 0), "velocity.speed", *LIMITS["emitter"]["speed"])}
    return out


def resolve_explosion(options: dict) -> dict:
    """resolveExplosion (options.ts:556-601) over its defaults."""
    _object(options, "explosion", ["radius", "charge", "outwardSpeed", "variation"])
    lim = LIMITS["explosion"]
    out = {"radius": 0.5, "charge": {"flame": 1, "heat": 2, "smoke": 0, "fuel": 0}, "outwardSpeed": 8,
           "variation": {"period": 0.5, "strength": 0}}
    if "radius" in options:
        out["radius"] = _number(options["radius"], "radius", *lim["radius"])
    if "outwardSpeed" in options:
        out["outwardSpeed"] = _number(options["outwardSpeed"], "outwardSpeed", *lim["outwardSpeed"])
    if "variation" in options:
        _object(options["variation"], "explosion.variation", ["period", "strength"])
        for name in ("period", "strength"):
            if name in options["variation"]:
                out["variation"][name] = _number(options["variation"][name], f"variation.{name}", *lim[name])
    if "charge" in options:
        _object(options["charge"], "charge", ["flame", "heat", "smoke", "fuel"])
        for name, v in options["charge"].items():
            out["charge"][name] = _number(v, f"charge.{name}", *(lim["flame"] if name == "flame" else lim["amount"]))
    return out


def validate_document(document) -> dict:
    """validateDocument (editor/document.ts:173-270): a resolved copy of a current-format
    preset, or ValueError naming the first problem."""
    doc = copy.deepcopy(document)
    if not isinstance(doc, dict) or doc.get("format") != FORMAT or doc.get("version") != VERSION:
        raise ValueError("Unsupported preset format. Export a current Fire Pro preset.")
    _object(doc, "Simulation", ["format", "version", "id", "name", "scene", "camera", "simulation",
                                "emitters", "forces", "colliders"])
    ids = set()

    def identity(v, path):
        _text(v.get("id"), f"{path}.id")
        _text(v.get("name"), f"{path}.name")
        if v["id"] in ids:
            raise ValueError(f"{path}: duplicate ID {v['id']}.")
        ids.add(v["id"])

    identity(doc, "Simulation")
    _object(doc.get("scene"), "scene", ["recipe", "sky"])
    scene = doc["scene"]
    if scene.get("recipe") not in RECIPES:
        raise ValueError("scene.recipe: unknown built-in scene.")
    _bool(scene.get("sky"), "scene.sky")
    _object(doc.get("camera"), "camera", ["position", "target"])
    _vec(doc["camera"].get("position"), "camera.position")
    _vec(doc["camera"].get("target"), "camera.target")
    sim = doc.get("simulation")
    _num(sim.get("brickSize") if isinstance(sim, dict) else None, "simulation.brickSize", 8, 32)
    doc["simulation"] = resolve_simulation(doc["simulation"])
    for key in ("emitters", "forces", "colliders"):
        _list(doc.get(key), key)
    for e in doc["emitters"]:
        _object(e, "emitter", ["id", "name", "mode", "shape", "radius", "size", "position", "rotation",
                               "options", "burst"])
        identity(e, "emitter")
        _vec(e.get("position"), "emitter.position")
        _vec(e.get("rotation"), "emitter.rotation")
        if e.get("mode") not in ("continuous", "burst"):
            raise ValueError("Unknown emitter mode.")
        if e.get("shape") not in SHAPES:
            raise ValueError("Unknown emitter shape.")
        if e["shape"] == "letters" and scene["recipe"] != "flaming-letters":
            raise ValueError("Letter geometry requires the Flaming letters scene.")
        _num(e.get("radius"), "emitter.radius", *LIMITS["emitter"]["radius"])
        _vec(e.get("size"), "emitter.size", *EMITTER_SIZE)
        e["options"] = resolve_emitter_options(e.get("options"), e["radius"])
        e["burst"] = resolve_explosion(e.get("burst"))
    targets = {e["id"] for e in doc["emitters"] if e["mode"] != "burst"}
    for f in doc["forces"]:
        _object(f, "force", ["id", "name", "options", "targets"])
        identity(f, "force")
        f["options"] = resolve_force(f.get("options"))
        if f.get("targets") is not None:
            _list(f["targets"], "force.targets")
            if len(set(f["targets"])) != len(f["targets"]) or any(t not in targets for t in f["targets"]):
                raise ValueError("Force targets must be unique continuous emitters.")
        elif "targets" not in f:
            raise ValueError("force.targets: expected a list.")
    for c in doc["colliders"]:
        _object(c, "collider", ["id", "name", "shape", "size", "radius", "position", "rotation"])
        identity(c, "collider")
        _vec(c.get("position"), "collider.position")
        _vec(c.get("rotation"), "collider.rotation")
        if c.get("shape") not in ("sphere", "box"):
            raise ValueError("Invalid collider shape.")
        _vec(c.get("size"), "collider.size", *LIMITS["collider"]["size"])
        _num(c.get("radius"), "collider.radius", *LIMITS["collider"]["radius"])
    return doc


def load_document(path: str) -> dict:
    with open(path) as f:
        return validate_document(json.load(f))


def euler_axes(rotation) -> Tuple[Tuple[float, ...], ...]:
    """The world axes (columns) of an object rotated by Euler angles in radians, applied in
    three.js's default XYZ order: R = Rx(x) Ry(y) Rz(z)."""
    x, y, z = rotation
    rx = np.array([[1, 0, 0], [0, math.cos(x), -math.sin(x)], [0, math.sin(x), math.cos(x)]])
    ry = np.array([[math.cos(y), 0, math.sin(y)], [0, 1, 0], [-math.sin(y), 0, math.cos(y)]])
    rz = np.array([[math.cos(z), -math.sin(z), 0], [math.sin(z), math.cos(z), 0], [0, 0, 1]])
    r = rx @ ry @ rz
    return tuple(tuple(float(v) for v in r[:, k]) for k in range(3))


class Playback:
    """examples/playback/main.ts:37-95 without a renderer: one FireSimulation in world
    space with the document's emitters, forces and colliders; bursts fire at the start."""

    def __init__(self, document: dict):
        self.document = validate_document(document)
        d = self.document
        self.simulation = FireSimulation(d["simulation"])
        self.emitters: Dict[str, object] = {}
        self.explosions: List[Tuple[object, List[float]]] = []
        for source in d["emitters"]:
            if source["mode"] == "burst":
                x = self.simulation.add_explosion(source["burst"])
                self.explosions.append((x, list(source["position"])))
                continue
            if source["shape"] == "letters":
                raise abi.FireProError(
                    f"emitter {source['id']}: letters emitters emit from a mesh surface of text "
                    "geometry, which is not supported.")
            if source["shape"] == "sphere":
                shape = {"radius": source["radius"]}
            else:
                shape = {"type": "mesh", "geometry": mesh.document_geometry(
                    source["shape"], source["radius"], source["size"])}
            opts = dict(source["options"], shape=shape, position=list(source["position"]),
                        rotation=list(source["rotation"]))
            self.emitters[source["id"]] = self.simulation.add_emitter(opts)
        for force in d["forces"]:
            targets = None if force["targets"] is None else [self.emitters[t] for t in force["targets"]]
            self.simulation.add_force(force["options"], targets)
        for c in d["colliders"]:
            axes = euler_axes(c["rotation"])
            if c["shape"] == "sphere":
                self.simulation.add_collider("sphere", c["position"], radius=c["radius"], axes=axes)
            else:
                self.simulation.add_collider("box", c["position"], size=c["size"], axes=axes)

    def detonate(self) -> None:
        """Fire every burst emitter at its position (main.ts:97-99)."""
        for x, position in self.explosions:
            x.trigger(position)

    def grid(self) -> dict:
        o = self.simulation.options
        return {"voxelSize": o["voxelSize"], "velocityDivisor": o["velocityDivisor"],
                "brickSize": o["brickSize"], "smokeDivisor": o["smokeDivisor"],
                "scalarMacCormack": o["scalarMacCormack"]}

    def frames(self, steps: int) -> Iterator[Tuple[abi.FrameData, dict]]:
        """Each fixed step's frame and grid settings, bursts fired before the first."""
        self.detonate()
        for _ in range(steps):
            yield self.simulation.build_frame(), self.simulation.grid_settings()


def preset_scene(document: dict, steps: int):
    """A preset as a scenes.Scene: the playback's frames, bursts at step 0."""
    from .scenes import Scene
    play = Playback(document)

    def hook(k: int):
        if k == 0:
            play.detonate()

    return Scene(play.document["id"], play.simulation, steps, hook=hook,
                 camera={"position": tuple(play.document["camera"]["position"]),
                         "target": tuple(play.document["camera"]["target"]), "fov": 44,
                         "width": 1280, "height": 720})
