"""The benchmark workloads as scenes built through the library API.

Randomness is PCG32 (O'Neill 2014, pcg32_srandom/pcg32_random) seeded by the scene seed.
Each scene is a FireSimulation with its handles, a step count and a per-step hook (bursts
trigger at their step). `frames()` yields each step's plain-data frame and grid settings
without touching a GPU, so any engine can replay the identical input.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Iterator, Optional, Tuple

from . import abi
from .library import FireSimulation


class PCG32:
    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed: int, stream: int = 54):
        self.state = 0
        self.inc = ((stream << 1) | 1) & self.MASK
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u32() / 4294967296.0)

    def integer(self, lo: int, hi: int) -> int:  # inclusive
        return lo + self.next_u32() % (hi - lo + 1)

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))


@dataclass
class Scene:
    name: str
    sim: FireSimulation
    steps: int
    hook: Optional[Callable[[int], None]] = None
    camera: Optional[dict] = None

    def frames(self, steps: Optional[int] = None) -> Iterator[Tuple[abi.FrameData, dict]]:
        for k in range(self.steps if steps is None else steps):
            if self.hook:
                self.hook(k)
            yield self.sim.build_frame(), self.sim.grid_settings()

    def grid(self) -> dict:
        o = self.sim.options
        return {"voxelSize": o["voxelSize"], "velocityDivisor": o["velocityDivisor"],
                "brickSize": o["brickSize"], "smokeDivisor": o["smokeDivisor"],
                "scalarMacCormack": o["scalarMacCormack"]}


def campfire(seed: int = 1, steps: int = 60, smoke: int = 1) -> Scene:
    """Plain method, small: one spherical fuel+heat emitter r 0.3 at (0, 0.2, 0)."""
    sim = FireSimulation({"voxelSize": 0.05, "brickSize": 16, "velocityDivisor": 2, "seed": seed,
                          "smokeDivisor": smoke,
                          "fuel": {"enabled": True, "ignitionHeat": 0.5, "burnRate": 4}})
    sim.add_emitter({"position": (0, 0.2, 0), "shape": {"radius": 0.3},
                     "emission": {"flame": 0, "heatRate": 10, "smokeRate": 0, "fuelRate": 2}})
    cam = {"position": (-4.12, 1.96, 5.30), "target": (0, 1.25, 0), "fov": 44,
           "width": 256, "height": 256}
    return Scene("campfire-plain", sim, steps, camera=cam)


def four_fires(seed: int = 1, steps: int = 300, brick: int = 8, max_voxels: int = 16_000_000) -> Scene:
    """Four fire spheres at uniform positions in [-3,3]x[0,4]x[-3,3], curl-noise turbulence
    (strength 1, scale 2) and one sphere collider of radius 0.5."""
    rng = PCG32(seed)
    sim = FireSimulation({"voxelSize": 0.025, "brickSize": brick, "velocityDivisor": 2,
                          "seed": seed, "grid": {"maxVoxels": max_voxels}})
    for _ in range(4):
        p = (rng.uniform(-3, 3), rng.uniform(0, 4), rng.uniform(-3, 3))
        sim.add_fire({"position": p})
    sim.add_force({"type": "turbulence", "strength": 1, "scale": 2})
    c = (rng.uniform(-3, 3), rng.uniform(0, 4), rng.uniform(-3, 3))
    sim.add_collider("sphere", c, radius=0.5)
    cam = {"position": (0, 3, 10), "target": (0, 2, 0), "fov": 50, "width": 1920, "height": 1080}
    return Scene(f"four-fires-b{brick}", sim, steps, camera=cam)


def scattered(count: int = 5, spacing: float = 1.6, steps: int = 4) -> Scene:
    """A count^3 lattice of small fires `spacing` apart (voxel 0.025 m, brick 8): each keeps
    its own tile and the margin around it, so the simulation holds over a thousand tile
    slots while its bricks stay few."""
    sim = FireSimulation({"voxelSize": 0.025, "brickSize": 8, "velocityDivisor": 2, "seed": 1,
                          "grid": {"maxVoxels": 16_000_000}})
    h = (count - 1) * spacing / 2
    for i in range(count):
        for j in range(count):
            for k in range(count):
                sim.add_fire({"position": (i * spacing - h, 0.5 + j * spacing, k * spacing - h),
                              "shape": {"radius": 0.08}})
    sim.add_collider("sphere", (0.2, 0.5 + h, 0.1), radius=0.3)
    return Scene(f"scattered-{count}", sim, steps)


def explosions(seed: int = 1, steps: int = 400) -> Scene:
    """Brick churn: 8 radial bursts in a 20 m cube at uniform steps 0-150."""
    rng = PCG32(seed)
    sim = FireSimulation({"voxelSize": 0.05, "brickSize": 8, "velocityDivisor": 2, "seed": seed,
                          "grid": {"maxVoxels": 16_000_000},
                          "fuel": {"enabled": True, "ignitionHeat": 0.5, "burnRate": 4}})
    bursts = []
    for _ in range(8):
        t = rng.integer(0, 150)
        pos = (rng.uniform(-10, 10), rng.uniform(0, 20), rng.uniform(-10, 10))
        fuel = rng.log_uniform(0.5, 5)
        x = sim.add_explosion({"radius": 0.5, "outwardSpeed": 8,
                               "charge": {"flame": 1, "heat": 2, "fuel": fuel}})
        bursts.append((t, pos, x))

    def hook(k: int):
        for t, pos, x in bursts:
            if t == k:
                x.trigger(pos)

    return Scene("explosions", sim, steps, hook=hook)


def small(seed: int = 1, steps: int = 12, brick: int = 8, divisor: int = 2,
          voxel: float = 0.05, collider: bool = True, turbulence: bool = True,
          burst: bool = True, smoke: int = 1, tilted: bool = False) -> Scene:
    """A seconds-sized scene touching every stage: an emitter with fuel and velocity, a
    smoke source, a burst with variation, a turbulence force and a box collider; `smoke` is
    the smoke grid's divisor. With `tilted` the box is turned about all three axes and a
    second, spherical collider sits in the plume."""
    sim = FireSimulation({"voxelSize": voxel, "brickSize": brick, "velocityDivisor": divisor,
                          "smokeDivisor": smoke,
                          "seed": seed, "fuel": {"enabled": True, "ignitionHeat": 0.5,
                                                 "burnRate": 4}})
    sim.add_fire({"position": (0.05, 0.3, 0.02), "shape": {"radius": 0.25},
                  "emission": {"fuelRate": 1.5}})
    sim.add_smoke({"position": (0.6, 0.4, -0.2), "shape": {"radius": 0.2}})
    if turbulence:
        sim.add_force({"type": "turbulence", "strength": 1.5, "scale": 2})
    sim.add_force({"type": "vortex", "center": (0.2, 0.5, 0), "strength": 2, "radius": 0.6,
                   "lift": 0.5, "inward": 0.3})
    if collider and tilted:
        import math
        a, b, g = 0.5, 0.3, 0.8
        ca, sa, cb, sb, cg, sg = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(g), math.sin(g)
        # The columns of Rz(g) Ry(b) Rx(a): the box's world axes.
        rot = [[cb * cg, sa * sb * cg - ca * sg, ca * sb * cg + sa * sg],
               [cb * sg, sa * sb * sg + ca * cg, ca * sb * sg - sa * cg],
               [-sb, sa * cb, ca * cb]]
        axes = tuple(tuple(rot[r][k] for r in range(3)) for k in range(3))
        sim.add_collider("box", (0.1, 0.95, 0.0), size=(0.5, 0.12, 0.4), axes=axes)
        sim.add_collider("sphere", (0.0, 0.75, 0.05), radius=0.11)
    elif collider:
        sim.add_collider("box", (0.1, 0.95, 0.0), size=(0.5, 0.12, 0.4))
    x = sim.add_explosion({"radius": 0.3, "outwardSpeed": 4,
                           "charge": {"flame": 1, "heat": 2, "fuel": 1},
                           "variation": {"period": 0.3, "strength": 0.6}})

    def hook(k: int):
        if burst and k == 3:
            x.trigger((-0.4, 0.5, 0.3))

    suffix = (f"-s{smoke}" if smoke > 1 else "") + ("-tilted" if tilted else "")
    return Scene(f"small-b{brick}-d{divisor}{suffix}", sim, steps, hook=hook)


def meshes(seed: int = 1, steps: int = 12, smoke: int = 1, colliders: bool = False,
           targeted: bool = False) -> Scene:
    """Mesh emitters of the three shapes presets use: a tilted box blowing flame and fuel
    sideways, a torus of heat and smoke, and a disk of flame on the ground, with turbulence;
    with `colliders`, a tilted mesh box over the disk and a mesh octahedron in the plume;
    with `targeted`, a vortex acting only on the box's and the torus's cells and a sphere
    smoke source's region."""
    from . import mesh
    sim = FireSimulation({"voxelSize": 0.05, "brickSize": 8, "velocityDivisor": 2, "smokeDivisor": smoke,
                          "seed": seed, "fuel": {"enabled": True, "ignitionHeat": 0.5, "burnRate": 4}})
    box = sim.add_emitter({"shape": {"type": "mesh", "geometry": mesh.box(0.4, 0.1, 0.3)},
                     "position": (0.3, 0.35, 0.0), "rotation": (0.2, 0.5, 0.0),
                     "emission": {"flame": 1, "heatRate": 2, "fuelRate": 1},
                     "velocity": {"direction": (1, 1, 0), "speed": 2}})
    ring = sim.add_emitter({"shape": {"type": "mesh", "geometry": mesh.document_geometry("torus", 0.25, (1, 1, 1))},
                     "position": (-0.4, 0.45, 0.1), "emission": {"heatRate": 3, "smokeRate": 2}})
    sim.add_emitter({"shape": {"type": "mesh", "geometry": mesh.document_geometry("disk", 0.3, (1, 1, 1))},
                     "position": (0.0, 0.15, -0.4), "emission": {"flame": 1, "heatRate": 1.5}})
    sim.add_force({"type": "turbulence", "strength": 1.5, "scale": 2})
    if colliders:
        sim.add_collider("mesh", (0.05, 0.55, -0.35), geometry=mesh.box(0.35, 0.08, 0.3), rotation=(0.3, 0.2, 0.1))
        sim.add_collider("mesh", (-0.35, 0.9, 0.1), geometry=mesh.octahedron(0.18), rotation=(0.0, 0.4, 0.0))
    if targeted:
        puff = sim.add_smoke({"position": (0.5, 0.6, 0.4), "shape": {"radius": 0.15}})
        sim.add_force({"type": "vortex", "center": (0.0, 0.5, 0.0), "strength": 6, "radius": 0.5,
                       "lift": 2, "inward": 1}, [box, ring, puff])
    suffix = ("-s%d" % smoke if smoke > 1 else "") + ("-colliders" if colliders else "") + ("-targeted" if targeted else "")
    return Scene(f"meshes{suffix}", sim, steps)


SCENES: Dict[str, Callable[..., Scene]] = {
    "campfire": campfire, "four_fires": four_fires, "explosions": explosions, "small": small,
    "meshes": meshes,
}
