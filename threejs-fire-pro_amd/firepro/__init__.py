"""firepro: MI355X-native sparse-brick fire/smoke solver (HIP on gfx950).

The library layer (`library.FireSimulation`) mirrors Fire Pro's src/library API; the engine
layer (`engine.FluidSimulation`) drives the C-ABI HIP library, libfirepro.so.
"""
from . import abi  # noqa: F401
from .library import FireSimulation, resolve_simulation, resolve_settings  # noqa: F401

__all__ = ["FireSimulation", "resolve_simulation", "resolve_settings", "abi"]
