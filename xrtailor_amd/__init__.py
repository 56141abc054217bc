"""MI355X (gfx950) cloth-substep solver: HIP kernels behind a C ABI.

The product is ``libxrtailor_amd.so`` (C ABI in ``include/xrtailor_amd.h``).  This
package holds its in-tree build, a ctypes mirror of the ABI, and the procedural scene
generators the tests and benchmark share.
"""
from ._ffi import XrtError, XrtParams, load  # noqa: F401
from .solver import BendSolver, ClothSolver, StretchSolver, default_params  # noqa: F401

__all__ = ["ClothSolver", "StretchSolver", "BendSolver", "XrtParams", "XrtError",
           "default_params", "load"]
