"""gpaw_amd: GPAW's real-space finite-difference / multigrid hot path on
MI355X (gfx950).

Drop-in modules for the reference API:
  gpaw_amd.grid_descriptor.GridDescriptor   (gpaw/grid_descriptor.py)
  gpaw_amd.fd_operators.Laplace / Gradient  (gpaw/fd_operators.py)
  gpaw_amd.transformers.Transformer         (gpaw/transformers.py)
  gpaw_amd.poisson.PoissonSolver            (gpaw/poisson.py)
All compute runs in libgpaw_amd.so (hand-written HIP, C ABI in
include/gpaw_amd.h); there is no CPU fallback.
"""
from ._lib import LIB_PATH, GpawAmdError, LibraryNotBuilt, lib  # noqa: F401

__all__ = ['LIB_PATH', 'GpawAmdError', 'LibraryNotBuilt', 'lib']
