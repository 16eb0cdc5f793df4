"""MI355X-native (gfx950) Moreau contact solve: projected Jacobi and graph-coloured
projected Gauss-Seidel in FP64 HIP kernels behind a C ABI (include/cdx/cdx.h).

Submodules: ``native`` (ctypes bindings, fails loudly without the built library), ``solver``
(SolverBase-shaped host interface, Moreau stepping), ``scenes`` (CPU contact generation and
scene builders), ``build`` (in-tree build recipe).
"""
from . import native  # noqa: F401

__all__ = ["native", "solver", "scenes", "build"]
