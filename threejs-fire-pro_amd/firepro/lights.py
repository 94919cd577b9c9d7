"""Scene lights (src/library/SceneLights.ts): point lights for the scene's surfaces, one per
anchor (an emitter, or an explosion while it burns), placed at the centroid of the flames
nearest that anchor, coloured and as bright as those flames are.

The device part is fp_sim_scene_lights (FluidSimulation.scene_lights). Every active brick
samples its fields on a 4^3 lattice and gives its flames' luminance, luminance-weighted
offset and rgb to its nearest anchor. This module is the host part of SceneLights.ts. It
turns the totals into measured lights (SceneLights.apply, :220-255), ends explosions' lights
once they have faded (finished, :117-124), and hands the lights to point lights (place,
:258-286). The reference uses three.js PointLights. Here a light is a plain record, and the
application copies it into its own scene.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

SUM_FLOATS = 8          # floats summed per anchor (scene-lights.wgsl)
LIGHT_CUTOFF = 0.01     # a light ends where its illuminance falls to this
FIXED_LIGHTS = 4        # lights kept without clustered lighting (SceneLights.ts:21)
SETTLE_READINGS = 10    # an explosion's light ends after this many readings ...
FADED = 0.01            # ... once its brightness is below this share of its peak
MAX_INTENSITY = 400.0   # SceneLights.ts:245


@dataclass
class LightAnchor:
    """Where flames light the scene from (SceneLights.ts:28-34)."""
    id: str
    position: Sequence[float]
    transient: bool = False  # an explosion's light, which ends once its flames die down


@dataclass
class Light:
    """A point light as the application would place it (SceneLights.copy, :287-292)."""
    position: List[float]
    color: List[float]
    intensity: float
    distance: float


@dataclass
class Measured:
    """An anchor's light as last measured (SceneLights.ts:35-45)."""
    position: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    color: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])
    intensity: float = 0.0
    transient: bool = False
    peak: float = 0.0
    readings: int = 0
    light: Optional[Light] = None


def _copy(m: Measured) -> Light:
    return Light(list(m.position), list(m.color), m.intensity, math.sqrt(m.intensity / LIGHT_CUTOFF))


class SceneLights:
    """SceneLights (SceneLights.ts:63-328) without the three.js objects. With `clustered`
    every lit anchor keeps its own light, as with three.js ClusteredLighting. Otherwise a
    fixed few lights go to the anchors that light the viewer's surroundings most."""

    def __init__(self, clustered: bool = False):
        self.clustered = clustered
        self.measured: Dict[str, Measured] = {}
        self.fixed: List[Optional[Light]] = []
        self.viewer = [0.0, 0.0, 0.0]  # where the camera that last drew the simulation stood

    def finished(self, anchor_id: str) -> bool:
        """Whether an explosion's light has died down, so its anchor can go."""
        m = self.measured.get(anchor_id)
        return bool(m and m.transient and m.readings >= SETTLE_READINGS and m.intensity <= m.peak * FADED)

    def sample(self, simulation, anchors: Sequence[LightAnchor], intensity: float = 1.0, settings=None) -> None:
        """Collect the light of the simulation's latest step by anchor, then place the lights
        (SceneLights.sample, :127-218, read back at once)."""
        if anchors:
            sums = simulation.scene_lights([a.position for a in anchors], settings)
        else:
            sums = np.zeros((0, SUM_FLOATS), np.float32)
        self.apply(sums, anchors, intensity, simulation.brick_size, simulation.voxel_size)

    def apply(self, sums: np.ndarray, anchors: Sequence[LightAnchor], intensity: float, brick_cells: int,
              voxel_size: float) -> None:
        """Measure each anchor's light from its totals (SceneLights.apply, :220-255)."""
        # Each sample stands for a cube a quarter of a brick wide.
        sample = (brick_cells / 4) * voxel_size
        live = set()
        for i, anchor in enumerate(anchors):
            live.add(anchor.id)
            mass, x, y, z, r, g, b = (float(v) for v in sums[i][:7])
            m = self.measured.get(anchor.id)
            if m is None:
                m = Measured(transient=anchor.transient)
                self.measured[anchor.id] = m
            energy = max(0.0, mass) * sample ** 3
            m.position = [float(c) for c in anchor.position]
            if mass > 0:
                m.position = [m.position[0] + x / mass, m.position[1] + y / mass, m.position[2] + z / mass]
            peak = max(r, g, b, 1e-8)
            m.color = [r / peak, g / peak, b / peak]
            m.intensity = min(MAX_INTENSITY, energy * 2 * intensity)
            m.peak = max(m.peak, m.intensity)
            m.readings += 1
        for anchor_id in [k for k in self.measured if k not in live]:
            del self.measured[anchor_id]
        self.place()

    def place(self) -> None:
        """Give the measured lights to point lights (SceneLights.place, :258-286)."""
        if self.clustered:
            self.fixed = []
            for m in self.measured.values():
                m.light = _copy(m) if m.intensity > 0 else None
            return
        for m in self.measured.values():
            m.light = None
        lit = [m for m in self.measured.values() if m.intensity > 0]

        def at_viewer(m: Measured) -> float:
            d2 = sum((p - v) ** 2 for p, v in zip(m.position, self.viewer))
            return m.intensity / max(1e-6, d2)

        brightest = sorted(lit, key=at_viewer, reverse=True)
        self.fixed = [_copy(brightest[i]) if i < len(brightest) else None for i in range(FIXED_LIGHTS)]

    @property
    def lights(self) -> List[Light]:
        """The lights now lit."""
        if self.clustered:
            return [m.light for m in self.measured.values() if m.light is not None]
        return [light for light in self.fixed if light is not None and light.intensity > 0]

    def reset(self) -> None:
        """Turn every light off and drop the measurements (SceneLights.reset, :312-318)."""
        self.measured.clear()
        self.fixed = [None] * len(self.fixed)
