"""campfire_amd — MI355X (gfx950) implementation of campfire's per-timestep gas solver.

The product is libcampfire_hip.so (HIP kernels + the C ABI of
include/campfire_hip.h).  This package holds its sources, the build recipe,
the Python mirror of the reference `FluidSim` (campfire_amd.fluid) and the
seeded scene builder (campfire_amd.scenes).
"""
from ._lib import CampfireError  # noqa: F401

__all__ = ["CampfireError"]
