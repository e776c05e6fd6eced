"""sam_amd -- MI355X-native RK3 pseudo-spectral step of SAM (tslund/sam).

The compute path is libsamhip.so (HIP kernels for gfx950 behind the C ABI of
include/samhip.h).  ``sam_amd.samhip`` binds it; ``sam_amd.ic`` writes seeded
initial conditions; ``sam_amd.io`` reads/writes input.dat, vel.* and header.*.
"""
from .samhip import Sam, make_params, load, SamError  # noqa: F401

__all__ = ["Sam", "make_params", "load", "SamError"]
