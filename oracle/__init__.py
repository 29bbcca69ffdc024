"""TEST INFRASTRUCTURE ONLY: the CPU parity oracle for the gpaw_amd hot path.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package.  The product never does.
"""
from .gpaw_oracle import (  # noqa: F401
    Kernels, Grid, laplace_stencil, gradient_stencil, guc_stencil, mehrstellen_a,
    mehrstellen_b, Operator, Transformer, PoissonSolver, flatten_offsets,
    transformer_plan, gaussian_charges, madelung, bohr_per_angstrom,
    Preconditioner, laplace, apply_pseudo_hamiltonian, interpolate_density,
    restrict_potential, pair_potentials, rpa_coupling,
)
