"""lsmlib_amd -- LSMLIB's 3D level-set update path on MI355X (gfx950).

`toolbox(real)` returns the drop-in LSM3D_* library (double or float build),
callable with numpy arrays (host memory, staged) or torch tensors (device
memory, in place).  `pylsmlib` holds the torch-facing solvers (fused RK
stages, batched patches, the adjoint).
"""
from ._native import LsmError, lib_path, toolbox  # noqa: F401
from .signatures import CITATIONS, MACRO_OF, SIGNATURES  # noqa: F401

__all__ = ["toolbox", "lib_path", "LsmError", "SIGNATURES", "MACRO_OF", "CITATIONS"]
