"""gpaw_amd: GPAW's real-space finite-difference / multigrid hot path on
MI355X (gfx950).

Drop-in modules for the reference API:
  gpaw_amd.grid_descriptor.GridDescriptor   (gpaw/grid_descriptor.py)
  gpaw_amd.fd_operators.Laplace / Gradient  (gpaw/fd_operators.py)
  gpaw_amd.transformers.Transformer         (gpaw/transformers.py)
  gpaw_amd.poisson.PoissonSolver            (gpaw/poisson.py)
  gpaw_amd.preconditioner.Preconditioner    (gpaw/preconditioner.py)
Beyond the reference's API:
  gpaw_amd.poisson.MixedPoissonSolver       FP32 V-cycles inside an FP64 outer loop
  gpaw_amd.batch.BatchApply                 many grids / phases / band counts per launch
  gpaw_amd.serving.Server                   continuous batching of Poisson / Laplace requests
  gpaw_amd.callers                          the reference's callers of the path (Hamiltonian
                                            on bands, density interpolation, potential
                                            restriction, pair-density solves)
All compute runs in libgpaw_amd.so (hand-written HIP, C ABI in
include/gpaw_amd.h); there is no CPU fallback.
"""
from ._lib import LIB_PATH, GpawAmdError, LibraryNotBuilt, lib  # noqa: F401

__all__ = ['LIB_PATH', 'GpawAmdError', 'LibraryNotBuilt', 'lib']
