"""MI355X-native MCML-style photon transport on triangle meshes.

The product is liblp.so: hand-written HIP kernels for gfx950 behind the C
ABI of include/lightprop.h, with a C++ host runtime (BVH, scene, tallies) and
the CalculatePropagation-compatible C++ facade (lightpropagation_amd/host).
This package exposes that ABI to Python through ctypes.
"""
from ._capi import Context, LPError, RNG_PHILOX, RNG_RAN3, TALLY, MISC, FIXED, FIXED_PHI, lib  # noqa: F401

__all__ = ["Context", "LPError", "RNG_PHILOX", "RNG_RAN3", "TALLY", "MISC", "FIXED", "FIXED_PHI", "lib"]
