"""The library layer: FireSimulation and its handles, mirroring src/library.

`FireSimulation` resolves options like src/library/options.ts, keeps emitter, explosion,
force and collider handles like src/library/handles.ts, and advances in fixed 1/60 s steps
like src/library/FireSimulation.ts:434-519, assembling each step's plain-data frame and
handing it to the HIP-backed engine (engine.FluidSimulation). three.js transforms are
replaced by world positions and axes set on the handles.
"""
from __future__ import annotations

import copy
import math
from typing import List, Optional, Sequence

from . import abi

FIXED_DT = 1 / 60  # src/engine/types.ts:44

# options.ts:337-365
DEFAULTS = {
    "grid": {"maxVoxels": 8_000_000, "cutoff": 0.05, "ground": True},
    "flame": {"lifespan": 1.2, "heatRate": 4, "smokeRate": 2, "expansionRate": 1,
              "cooling": 0.62, "color": "#ffffff", "brightness": 1.4, "opacity": 0.55,
              "temperature": 3200, "sootGlow": 1},
    "smoke": {"dissipation": 0.35, "color": "#25292e", "density": 2.5, "scattering": 0.38,
              "shadowDensity": 3.5},
    "motion": {"buoyancy": 1.6, "smokeWeight": 0.12, "damping": 0.1, "vorticity": 3.2},
    "fuel": {"enabled": False, "ignitionHeat": 0.5, "burnRate": 4},
    "lighting": {"illuminateScene": True, "intensity": 1},
    "rendering": {"filter": "cubic", "raySteps": 256, "halfResolution": True,
                  "lightingDivisor": 4},
}
# options.ts:219-231
DEBUG_FIELDS = ("beauty", "lifetime", "heat", "smoke", "velocity", "flame", "fuel", "vorticity", "expansion",
                "pressure", "divergence")

# options.ts:367-405
LIMITS = {
    "simulation": {"voxelSize": (0.002, 4)},
    "grid": {"maxVoxels": (4096, 268_435_456), "cutoff": (0, 1)},
    "flame": {"lifespan": (0.01, 5), "heatRate": (0, 20), "smokeRate": (0, 20),
              "expansionRate": (0, 20), "cooling": (0, 20), "brightness": (0, 20),
              "opacity": (0, 20), "temperature": (800, 10000), "sootGlow": (0, 5)},
    "smoke": {"dissipation": (0, 20), "density": (0, 20), "scattering": (-0.9, 0.9),
              "shadowDensity": (0, 20)},
    "motion": {"buoyancy": (0, 20), "smokeWeight": (0, 20), "damping": (0, 20),
               "vorticity": (0, 20)},
    "fuel": {"ignitionHeat": (0, 100), "burnRate": (0, 100)},
    "lighting": {"intensity": (0, 20)},
    "rendering": {"raySteps": (8, 1024)},
    "emitter": {"radius": (0.01, 16), "flame": (0, 1), "rate": (0, 100), "speed": (0, 50)},
    "explosion": {"radius": (0.01, 16), "outwardSpeed": (0, 50), "period": (0.01, 32),
                  "strength": (0, 1), "flame": (0, 1), "amount": (0, 100)},
    "wind": {"strength": (0, 50)},
    "turbulence": {"strength": (0, 50), "scale": (0.01, 32)},
    "vortex": {"strength": (-50, 50), "radius": (0.01, 32), "lift": (-50, 50),
               "inward": (0, 50)},
    "radial": {"strength": (-50, 50), "radius": (0.01, 32)},
    "collider": {"radius": (0.01, 32), "size": (0.01, 64)},
}
GROUPS = ("grid", "flame", "smoke", "motion", "fuel", "lighting", "rendering")


def _number(value, label, lo=0.0, hi=1000.0):
    """options.ts:315-320"""
    if (not isinstance(value, (int, float)) or isinstance(value, bool) or not math.isfinite(value)
            or value < lo or value > hi):
        raise ValueError(f"{label} must be a finite number from {lo} to {hi}.")
    return value


def _keys(value, allowed, label):
    """options.ts:309-314"""
    if not isinstance(value, dict):
        raise ValueError(f"{label} must be an options object.")
    for k in value:
        if k not in allowed:
            raise ValueError(f"Unknown {label} option: {k}.")


def resolve_settings(patch: dict, previous: Optional[dict] = None) -> dict:
    """resolveSettings (options.ts:418-450)."""
    result = copy.deepcopy(previous or DEFAULTS)
    for group in GROUPS:
        block = patch.get(group)
        if block is None:
            continue
        target = result[group]
        _keys(block, list(target.keys()), group)
        for key, value in block.items():
            if value is None:
                continue
            label = f"{group}.{key}"
            limits = LIMITS.get(group, {}).get(key)
            if limits:
                target[key] = _number(value, label, *limits)
            elif key == "color":
                target[key] = value
            elif key == "lightingDivisor":
                if value not in (1, 2, 4):
                    raise ValueError(f"{label} must be 1, 2 or 4.")
                target[key] = value
            elif key == "filter":
                if value not in ("trilinear", "quadratic", "cubic"):
                    raise ValueError(f"{label} must be trilinear, quadratic, cubic.")
                target[key] = value
            else:
                if not isinstance(value, bool):
                    raise ValueError(f"{label} must be a boolean.")
                target[key] = value
    if int(result["rendering"]["raySteps"]) != result["rendering"]["raySteps"]:
        raise ValueError("rendering.raySteps must be an integer.")
    if int(result["grid"]["maxVoxels"]) != result["grid"]["maxVoxels"]:
        raise ValueError("grid.maxVoxels must be an integer.")
    return result


def resolve_simulation(options: Optional[dict] = None) -> dict:
    """resolveSimulation (options.ts:451-478)."""
    options = options or {}
    _keys(options, ["voxelSize", "brickSize", "velocityDivisor", "smokeDivisor",
                    "scalarMacCormack", "seed", *GROUPS], "simulation")
    seed = _number(options.get("seed", 1), "seed", 0, 0xFFFFFFFF)
    if int(seed) != seed:
        raise ValueError("seed must be an integer.")
    voxel = _number(options.get("voxelSize", 0.02), "voxelSize", *LIMITS["simulation"]["voxelSize"])
    brick = options.get("brickSize", 16)
    if brick not in (8, 16, 32):
        raise ValueError("brickSize must be 8, 16 or 32.")
    divisor = options.get("velocityDivisor", 2)
    if divisor not in (1, 2, 4):
        raise ValueError("velocityDivisor must be 1, 2 or 4.")
    smoke = options.get("smokeDivisor", 1)
    if smoke not in (1, 2, 4):
        raise ValueError("smokeDivisor must be 1, 2 or 4.")
    mac = options.get("scalarMacCormack", True)
    if not isinstance(mac, bool):
        raise ValueError("scalarMacCormack must be a boolean.")
    out = {"voxelSize": voxel, "brickSize": brick, "velocityDivisor": divisor,
           "smokeDivisor": smoke, "scalarMacCormack": mac, "seed": seed}
    out.update(resolve_settings(options))
    return out


def _vector(value, label, lo=-1e6, hi=1e6):
    """options.ts:306-310"""
    if not isinstance(value, (list, tuple)) or len(value) != 3:
        raise ValueError(f"{label} must contain three numbers.")
    return [_number(v, f"{label}[{i}]", lo, hi) for i, v in enumerate(value)]


def _direction(value, label):
    """options.ts:311-318: normalized, keeping vectors already of unit length."""
    v = _vector(value, label)
    mag = math.hypot(*v)
    if mag == 0:
        raise ValueError(f"{label} must be nonzero.")
    div = 1 if abs(mag - 1) < 1e-12 else mag
    return [0 if c == 0 else c / div for c in v]


def _boolean(value, label):
    if not isinstance(value, bool):
        raise ValueError(f"{label} must be a boolean.")
    return value


def resolve_force(options: dict) -> dict:
    """resolveForce (options.ts:602-662) for a new force."""
    if not isinstance(options, dict):
        raise ValueError("force must be an options object.")
    kind = options.get("type")
    allowed = {"wind": ["direction", "strength", "active"],
               "turbulence": ["strength", "scale", "active"],
               "vortex": ["center", "axis", "strength", "radius", "lift", "inward", "active"]}.get(
        kind, ["center", "strength", "radius", "active"])
    _keys(options, ["type", *allowed], "force")
    active = _boolean(options.get("active", True), "force.active")
    if kind == "wind":
        return {"type": kind, "active": active,
                "direction": _direction(options.get("direction"), "wind.direction"),
                "strength": _number(options.get("strength"), "wind.strength", *LIMITS["wind"]["strength"])}
    if kind == "turbulence":
        return {"type": kind, "active": active,
                "strength": _number(options.get("strength"), "turbulence.strength",
                                    *LIMITS["turbulence"]["strength"]),
                "scale": _number(options.get("scale"), "turbulence.scale", *LIMITS["turbulence"]["scale"])}
    if kind in ("vortex", "radial"):
        out = {"type": kind, "active": active,
               "center": _vector(options.get("center", [0, 0, 0]), "force.center"),
               "strength": _number(options.get("strength"), "force.strength", *LIMITS[kind]["strength"]),
               "radius": _number(options.get("radius", 1), "force.radius", *LIMITS[kind]["radius"])}
        if kind == "vortex":
            out.update(axis=_direction(options.get("axis", [0, 1, 0]), "vortex.axis"),
                       lift=_number(options.get("lift", 0), "vortex.lift", *LIMITS["vortex"]["lift"]),
                       inward=_number(options.get("inward", 0), "vortex.inward", *LIMITS["vortex"]["inward"]))
        return out
    raise ValueError("force.type must be wind, turbulence, vortex or radial.")


# emitterDefaults (options.ts:480-502)
EMITTER_KINDS = {
    "fire": {"radius": 0.38, "flame": 1, "heatRate": 1.5, "smokeRate": 0, "speed": 1.4},
    "jet": {"radius": 0.22, "flame": 1, "heatRate": 2.5, "smokeRate": 0, "speed": 10},
    "smoke": {"radius": 0.4, "flame": 0, "heatRate": 0.85, "smokeRate": 4, "speed": 1.2},
    "emitter": {"radius": 0.3, "flame": 0, "heatRate": 0, "smokeRate": 0, "speed": 0},
}


class Emitter:
    """A continuous emitter (handles.ts:81-198): a sphere, or a mesh surface given as
    `{"type": "mesh", "geometry": mesh.Geometry}` and placed by `position` and `rotation`
    (Euler XYZ, radians), sampled once at creation (MeshSurface.ts)."""

    def __init__(self, sim: "FireSimulation", hid: int, kind: str, options: dict):
        d = EMITTER_KINDS[kind]
        self.id = hid
        self.position = list(options.get("position", (0.0, 0.0, 0.0)))
        self.rotation = list(options.get("rotation", (0.0, 0.0, 0.0)))
        self.active = options.get("active", True)
        shape = options.get("shape", {})
        self.surface = None
        self._placed = None
        if shape.get("type", "sphere") == "mesh":
            from . import mesh
            if "radius" in shape:
                raise ValueError("Mesh source does not accept radius.")
            geometry = shape.get("geometry")
            if not isinstance(geometry, mesh.Geometry):
                raise ValueError("shape.geometry must be a mesh.Geometry.")
            cell = sim.options["voxelSize"]
            self.surface = mesh.MeshSurface(geometry, cell, (int(sim.options["seed"]) + hid) & 0xFFFFFFFF)
            self.radius = cell
        elif shape.get("type", "sphere") != "sphere":
            raise ValueError("shape.type must be sphere or mesh.")
        else:
            self.radius = _number(shape.get("radius", d["radius"]), "shape.radius",
                                  *LIMITS["emitter"]["radius"])
        em = options.get("emission", {})
        self.emission = {
            "flame": _number(em.get("flame", d["flame"]), "emission.flame", *LIMITS["emitter"]["flame"]),
            "heatRate": _number(em.get("heatRate", d["heatRate"]), "emission.heatRate", *LIMITS["emitter"]["rate"]),
            "smokeRate": _number(em.get("smokeRate", d["smokeRate"]), "emission.smokeRate", *LIMITS["emitter"]["rate"]),
            "fuelRate": _number(em.get("fuelRate", 0), "emission.fuelRate", *LIMITS["emitter"]["rate"]),
        }
        vel = options.get("velocity", "default")
        if vel is None or (vel == "default" and kind == "emitter"):
            self.velocity = None
        else:
            vel = {} if vel == "default" else vel
            direction = vel.get("direction", (0, 1, 0))
            mag = math.sqrt(sum(c * c for c in direction))
            if mag == 0:
                raise ValueError("velocity.direction must be nonzero.")
            div = 1 if abs(mag - 1) < 1e-12 else mag
            self.velocity = {"direction": [c / div for c in direction],
                             "speed": _number(vel.get("speed", d["speed"]), "velocity.speed",
                                              *LIMITS["emitter"]["speed"])}
        self.removed = False
        self._sim = sim

    def source(self, include_inactive=False) -> Optional[abi.SourceData]:
        """Emitter.source (handles.ts:147-197): the world transform turns the velocity
        direction (normalized again) and carries the mesh samples, which are transformed
        again only when the pose changes."""
        from . import mesh
        if not self.active and not include_inactive:
            return None
        matrix = mesh.compose(self.position, self.rotation)
        v = [0.0, 0.0, 0.0]
        if self.velocity:
            v = [c * self.velocity["speed"] for c in mesh.transform_direction(self.velocity["direction"], matrix)]
        samples, area = None, 0.0
        if self.surface is not None:
            pose = (tuple(self.position), tuple(self.rotation))
            if self._placed is None or self._placed[0] != pose:
                self._placed = (pose, mesh.transform_samples(self.surface.samples, matrix))
            samples, area = self._placed[1], self.surface.surface_area
        return abi.SourceData(
            id=self.id, shape="mesh" if self.surface is not None else "sphere",
            position=tuple(float(x) for x in matrix[:3, 3]), size=(self.radius,) * 3,
            flame=self.emission["flame"], heat_rate=self.emission["heatRate"],
            smoke_rate=self.emission["smokeRate"], fuel_rate=self.emission["fuelRate"],
            velocity=tuple(v), velocity_response=60 if self.velocity else 0, outward_speed=0,
            samples=samples, surface_area=area)

    def remove(self):
        self.removed = True
        self._sim._emitters = [e for e in self._sim._emitters if e is not self]


class Explosion:
    """A burst that injects for one step when triggered (handles.ts:200-238)."""

    def __init__(self, sim: "FireSimulation", hid: int, options: dict):
        self.id = hid
        self._sim = sim
        self.radius = _number(options.get("radius", 0.5), "radius", *LIMITS["explosion"]["radius"])
        charge = {"flame": 1, "heat": 2, "smoke": 0, "fuel": 0}
        charge.update(options.get("charge", {}))
        self.charge = charge
        self.outward_speed = _number(options.get("outwardSpeed", 8), "outwardSpeed",
                                     *LIMITS["explosion"]["outwardSpeed"])
        var = {"period": 0.5, "strength": 0}
        var.update(options.get("variation", {}))
        self.variation = var

    def trigger(self, world_position: Sequence[float]):
        self._sim._pending.append((self, tuple(world_position), self._sim._next_event))
        self._sim._next_event += 1
        return self


class Force:
    """handles.ts:240-300"""

    def __init__(self, options: dict, targets: Optional[List[Emitter]]):
        self.options = dict(options)
        self.targets = targets
        self.active = options.get("active", True)

    def frame(self) -> Optional[abi.ForceData]:
        o = self.options
        if not self.active:
            return None
        kind = o["type"]
        axis = o.get("direction") if kind == "wind" else o.get("axis", (0, 1, 0)) if kind == "vortex" else (0, 1, 0)
        mag = math.sqrt(sum(c * c for c in axis))
        axis = tuple(c / mag for c in axis)
        scale = o["scale"] if kind == "turbulence" else o.get("radius", 1)
        return abi.ForceData(
            type=kind, vector=axis, center=tuple(o.get("center", (0, 0, 0))),
            strength=o["strength"], scale=scale, lift=o.get("lift", 0) if kind == "vortex" else 0,
            inward=o.get("inward", 0) if kind == "vortex" else 0,
            sources=None if self.targets is None else [t.source(True) for t in self.targets])


class Collider:
    """Sphere, box or mesh collider (handles.ts:284-355); `axes` are the box's world axes.
    A mesh collider is a closed mesh.Geometry placed by `position` and `rotation` (Euler XYZ),
    voxelized into field cells again whenever its pose changes (MeshCollider.ts)."""

    def __init__(self, shape: str, position, radius=None, size=None, axes=None, geometry=None,
                 rotation=(0.0, 0.0, 0.0), spacing: float = 0.0):
        self.shape = shape
        self.position = list(position)
        self.rotation = list(rotation)
        self.geometry = geometry
        self.spacing = spacing
        self._mask = None
        if shape == "sphere":
            self.size = (radius,) * 3
        elif shape == "box":
            self.size = tuple(s * 0.5 for s in size)
        elif shape == "mesh":
            from . import mesh
            if not isinstance(geometry, mesh.Geometry):
                raise ValueError("Mesh collider object must be a mesh.Geometry.")
            mesh.check_closed(geometry)
            self.size = (0.0, 0.0, 0.0)
        else:
            raise ValueError("collider.shape.type must be sphere, box or mesh.")
        self.axes = axes or ((1, 0, 0), (0, 1, 0), (0, 0, 1))

    def frame(self) -> abi.ColliderData:
        if self.shape != "mesh":
            return abi.ColliderData(self.shape, tuple(self.position), self.size, self.axes)
        from . import mesh
        pose = (tuple(self.position), tuple(self.rotation), self.spacing)
        if self._mask is None or self._mask[0] != pose:
            occupancy = mesh.voxelize(self.geometry, mesh.compose(self.position, self.rotation), self.spacing)
            self._mask = (pose, occupancy)
        return abi.ColliderData("mesh", tuple(self.position), (0.0, 0.0, 0.0), occupancy=self._mask[1])


class FireSimulation:
    """src/library/FireSimulation.ts: fixed-step driver over the HIP engine."""

    def __init__(self, options: Optional[dict] = None):
        self.options = resolve_simulation(options)
        self._emitters: List[Emitter] = []
        self._explosions: List[Explosion] = []
        self._forces: List[Force] = []
        self._colliders: List[Collider] = []
        self._pending = []
        self._next_handle = 1
        self._next_event = 1
        self.accumulator = 0.0
        self.dropped_time = 0.0
        self.fluid = None

    # initialize (FireSimulation.ts:177-229), without a renderer.
    def initialize(self):
        from .engine import FluidSimulation
        o = self.options
        self.fluid = FluidSimulation(o["voxelSize"], o["velocityDivisor"], o["brickSize"],
                                     o["smokeDivisor"], o["scalarMacCormack"])
        return self

    def configure(self, patch: dict):
        _keys(patch, GROUPS, "simulation.configure")
        self.options.update(resolve_settings(patch, {g: self.options[g] for g in GROUPS}))
        return self

    def _add(self, kind, options):
        e = Emitter(self, self._next_handle, kind, options or {})
        self._emitters.append(e)
        self._next_handle += 1
        return e

    def add_fire(self, options=None):
        return self._add("fire", options)

    def add_flame_jet(self, options=None):
        return self._add("jet", options)

    def add_smoke(self, options=None):
        return self._add("smoke", options)

    def add_emitter(self, options=None):
        return self._add("emitter", options)

    def add_explosion(self, options=None):
        x = Explosion(self, self._next_handle, options or {})
        self._explosions.append(x)
        self._next_handle += 1
        return x

    def add_force(self, options: dict, targets=None):
        if targets is not None and any(t not in self._emitters for t in targets):
            raise ValueError("Force target must be a live emitter owned by this simulation.")
        f = Force(resolve_force(options), targets)
        self._forces.append(f)
        return f

    def add_collider(self, shape: str, position, radius=None, size=None, axes=None, geometry=None,
                     rotation=(0.0, 0.0, 0.0)):
        c = Collider(shape, position, radius, size, axes, geometry, rotation, self.options["voxelSize"])
        self._colliders.append(c)
        return c

    def grid_settings(self) -> dict:
        g = self.options["grid"]
        return {"cutoff": g["cutoff"], "maxVoxels": g["maxVoxels"], "ground": g["ground"]}

    def build_frame(self) -> abi.FrameData:
        """The step's SimulationFrame (FireSimulation.ts:455-519); consumes pending bursts."""
        sources = [s for s in (e.source() for e in self._emitters) if s is not None]
        for owner, position, event in self._pending:
            r = owner.radius
            sources.append(abi.SourceData(
                id=(owner.id << 32) | event, shape="sphere", position=position, size=(r, r, r),
                flame=owner.charge["flame"], heat_rate=owner.charge["heat"] / FIXED_DT,
                smoke_rate=owner.charge["smoke"] / FIXED_DT, fuel_rate=owner.charge["fuel"] / FIXED_DT,
                velocity=(0, 0, 0), velocity_response=60 if owner.outward_speed > 0 else 0,
                outward_speed=owner.outward_speed,
                variation=(owner.variation["period"], owner.variation["strength"])))
        self._pending = []
        o = self.options
        fl, mo = o["flame"], o["motion"]
        fuel = o["fuel"]
        return abi.FrameData(
            buoyancy=mo["buoyancy"], smoke_weight=mo["smokeWeight"], velocity_damping=mo["damping"],
            cooling=fl["cooling"], dissipation=o["smoke"]["dissipation"], vorticity=mo["vorticity"],
            seed=o["seed"], lifespan=fl["lifespan"], flame_smoke_rate=fl["smokeRate"],
            flame_heat_rate=fl["heatRate"], expansion_rate=fl["expansionRate"],
            fuel=(fuel["enabled"], fuel["ignitionHeat"], fuel["burnRate"]),
            sources=sources, forces=[f for f in (x.frame() for x in self._forces) if f is not None],
            colliders=[c.frame() for c in self._colliders])

    def update(self, delta_seconds: float):
        """update (FireSimulation.ts:434-453): at most one fixed step a call."""
        if self.fluid is None:
            raise abi.FireProError("Call initialize() before updating or triggering effects.")
        _number(delta_seconds, "update deltaSeconds", 0, 1.7976931348623157e308)
        self.accumulator += delta_seconds
        if self.accumulator + 1e-12 >= FIXED_DT:
            self.accumulator = max(0.0, self.accumulator - FIXED_DT)
            behind = math.floor((self.accumulator + 1e-12) / FIXED_DT) * FIXED_DT
            self.accumulator = max(0.0, self.accumulator - behind)
            self.dropped_time += behind
            self.step()

    def step(self):
        g = self.grid_settings()
        self.fluid.cutoff = g["cutoff"]
        self.fluid.voxel_budget = g["maxVoxels"]
        self.fluid.ground = g["ground"]
        self.fluid.step(self.build_frame(), FIXED_DT)

    @property
    def stats(self) -> dict:
        """stats (FireSimulation.ts:295-334); the simulation's memory is memoryBytes +
        meshMemoryBytes + forceMemoryBytes (the render surface is not built here)."""
        s = self.fluid.stats()
        return {"simulationTime": s.time, "droppedTime": self.dropped_time,
                "gridLimited": bool(s.pool_limited),
                "activeVoxels": s.active_bricks * self.options["brickSize"] ** 3,
                "estimatedMemoryBytes": s.memory_bytes + s.mesh_memory_bytes + s.force_memory_bytes}

    def debug(self, options: dict):
        """debug (FireSimulation.ts:336-362) for the solver's part: a 'pressure' or
        'divergence' field makes the engine copy that field after every step
        (showSolverField); other fields are views of the render surface and stop the copies."""
        _keys(options, ["bricks", "field"], "debug")
        field = options.get("field", getattr(self, "debug_field", "beauty"))
        if field not in DEBUG_FIELDS:
            raise ValueError("Unknown debug field.")
        self.debug_field = field
        if self.fluid is not None:
            self.fluid.show_solver_field(field if field in ("pressure", "divergence") else None)
        return self

    def reset(self):
        self.fluid.reset()
        self._pending = []
        self.accumulator = 0.0
        self.dropped_time = 0.0
        self._next_event = 1

    def dispose(self):
        if self.fluid is not None:
            self.fluid.dispose()
            self.fluid = None
