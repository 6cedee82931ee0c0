"""sphexa_amd: MI355X-native (gfx950) single-rank SPH-EXA hydro hot path.

Hand-written HIP kernels in ``csrc/`` are built into ``libsphx.so`` (C ABI declared in
``include/sphx.h``); ``hydro`` mirrors the reference's C++ entry points on torch
device tensors.  Importing the package does not touch the GPU.
"""
from ._lib import (BC_FIXED, BC_OPEN, BC_PERIODIC, FP64, MIXED, Box, Groups, Octree, Particles, SphxError,
                   load)

__all__ = ["BC_FIXED", "BC_OPEN", "BC_PERIODIC", "FP64", "MIXED", "Box", "Groups", "Octree", "Particles",
           "SphxError", "load"]
