"""sprgpu: MI355X numeric factorization and solves for sam-lib's SPR solver.

* ``sprgpu.spr``    -- the reference entry points (SPRFC1, SPRFC3, SPRFS1,
                       SPRBS1, SPRSOL) on numpy arrays, running on the GPU.
* ``sprgpu.device`` -- the device-resident factor handle (FP64 or FP32
                       factor, FP64 iterative refinement, many RHS).
* ``sprgpu.fem``    -- model generators (hex8, quad4, 3D beam frames) that
                       produce SAM control arrays and element matrices.
"""
from ._lib import LIB_PATH, FLIB_PATH, SprGpuError, lib, last_error  # noqa: F401

__all__ = ["LIB_PATH", "FLIB_PATH", "SprGpuError", "lib", "last_error"]
