"""The reference's callers of the finite-difference hot path, on the device.

Each function restates one call site of the reference with the same
arguments and arithmetic, with the arrays on the GPU (torch tensors) or in
host memory (numpy arrays, copied in and out like every entry point of this
package):

* apply_pseudo_hamiltonian -- kinetic stencil over bands, then the local
  potential (gpaw/wavefunctions/fd.py:50-52, gpaw/hamiltonian.py:371-398);
* interpolate_density -- pseudo density of every spin to the fine grid, and
  the charge renormalisation of zero-boundary cells (gpaw/density.py:217-234);
* restrict_potential -- vt_g += vHt_g, then restriction to the coarse grid,
  for every spin (gpaw/hamiltonian.py:254-259);
* pair_potentials -- one Poisson solve per pair density, as the exact
  exchange does (gpaw/xc/hybrid.py:140-152);
* rpa_coupling -- one Poisson solve per pair density, then the coupling
  integrals of the RPA omega matrix (gpaw/lrtddft/omega_matrix.py:383-430).

The spins (or bands) of one call go to the device as the band axis of ONE
stencil / transfer launch.
"""
import numpy as np

from . import _arrays, ops
from ._arrays import Arg


def apply_local_potential(vt_G, psit_nG, Htpsit_nG):
    """Htpsit_nG += psit_nG * vt_G for every band (gpaw/hamiltonian.py:393-398)."""
    ops.add_potential(Htpsit_nG, psit_nG, vt_G)


def apply_pseudo_hamiltonian(kin, vt_G, psit_xG, Htpsit_xG, phase_cd=None):
    """Kinetic operator over the bands, then the local potential
    (gpaw/wavefunctions/fd.py:50-52 with
    Hamiltonian.apply_local_potential, gpaw/hamiltonian.py:371-398).

    kin: a Laplace(gd, -0.5, nn, dtype) of gpaw_amd.fd_operators."""
    out = Arg(Htpsit_xG, write=True, name='Htpsit_xG')
    psi = Arg(psit_xG, name='psit_xG')
    v = Arg(vt_G, name='vt_G')
    if not kin.is_allocated():
        kin.allocate()
    # one pass: the stencil sum, then + psi * vt, rounded as the two calls
    kin.operator.apply_potential(psi.tensor, out.tensor, v.tensor, phase_cd)
    out.finish()


def interpolate_density(interpolator, finegd, nt_sG, nt_sg, charge=0.0,
                        comp_charge=0.0):
    """Density.interpolate (gpaw/density.py:217-234): every spin of nt_sG to
    the fine grid; on cells that are not periodic in every direction the
    fine density is scaled so that it integrates to -(charge + comp_charge).

    interpolator: a Transformer(gd, finegd, nn) of gpaw_amd.transformers."""
    out = Arg(nt_sg, write=True, name='nt_sg')
    src = Arg(nt_sG, name='nt_sG')
    interpolator.apply(src.tensor, out.tensor)
    if not finegd.pbc_c.all():
        pseudo_charge = -(charge + comp_charge)
        if abs(pseudo_charge) > 1.0e-14:
            x = pseudo_charge / np.sum(finegd.integrate(out.tensor))
            ops.scale(out.tensor, x)
    out.finish()


def restrict_potential(restrictor, vt_sg, vHt_g, vt_sG):
    """vt_g += vHt_g and vt_G = R vt_g for every spin
    (gpaw/hamiltonian.py:254-259).  vt_sg is updated in place.

    restrictor: a Transformer(finegd, gd, nn) of gpaw_amd.transformers."""
    fine = Arg(vt_sg, write=True, name='vt_sg')
    vH = Arg(vHt_g, name='vHt_g')
    out = Arg(vt_sG, write=True, name='vt_sG')
    t = fine.tensor
    if t.dim() == 3:
        ops.axpby(t, vH.tensor, 1.0, 1.0)
    else:
        for s in range(t.shape[0]):
            ops.axpby(t[s], vH.tensor, 1.0, 1.0)
    restrictor.apply(t, out.tensor)
    fine.finish()
    out.finish()


def pair_potentials(poisson, rhot_pg, charges, eps=1e-12):
    """The exact-exchange loop's Poisson solves (gpaw/xc/hybrid.py:148-152):
    for every pair density p, vt_g = 0 and poisson.solve(vt_g, -rhot_g,
    charge=charges[p], eps=eps, zero_initial_phi=True).  Returns the
    potentials (same kind of array as rhot_pg) and the iteration counts.

    poisson: a PoissonSolver of gpaw_amd.poisson (its plan and graphs are
    reused for every pair)."""
    rho = Arg(rhot_pg, name='rhot_pg')
    t = _arrays.torch()
    vt = t.zeros_like(rho.tensor)
    neg = t.empty_like(rho.tensor[0])
    niter = []
    for p in range(rho.tensor.shape[0]):
        ops.cast(neg, rho.tensor[p])   # same dtype: an element copy
        ops.scale(neg, -1.0)           # -rhot_g: a sign flip, exact
        niter.append(poisson.solve(vt[p], neg, charge=charges[p], eps=eps,
                                   zero_initial_phi=True))
    if rho.host is not None:
        return vt.cpu().numpy(), niter
    return vt, niter


def rpa_coupling(poisson, rhot_pg, gd=None, restrictor=None, rhot_pG=None):
    """The Hartree part of the RPA omega matrix
    (gpaw/lrtddft/omega_matrix.py:383-430), without the pair prefactors and
    the atomic corrections: for every pair density ij, phit = 0 and
    poisson.solve(phit, rhot_ij, charge=None); with a restrictor
    (finegrid == 1) phit is restricted and integrated against the coarse
    densities rhot_pG; K[ij, kq] = gd.integrate(rhot_kq * phit) for kq >= ij
    (one launch of gpaw_dots per row) and K[kq, ij] = K[ij, kq].

    gd: the grid the integrals run over (default: the solver's).  Returns K
    as a numpy array and the iteration counts."""
    t = _arrays.torch()
    rho = Arg(rhot_pg, name='rhot_pg')
    if gd is None:
        gd = poisson.gd
    if restrictor is None:
        rint = rho.tensor
    else:
        if rhot_pG is None:
            raise ValueError('rpa_coupling: a restrictor needs the coarse densities rhot_pG')
        rint = Arg(rhot_pG, name='rhot_pG').tensor
    P = rho.tensor.shape[0]
    if rint.shape[0] != P:
        raise ValueError('rpa_coupling: %d coarse densities for %d pairs' % (rint.shape[0], P))
    K = t.zeros((P, P), dtype=t.float64, device='cuda')
    phit_p = t.empty_like(rho.tensor[0])
    phit = phit_p if restrictor is None else t.empty_like(rint[0])
    niter = []
    for ij in range(P):
        phit_p.zero_()
        niter.append(poisson.solve(phit_p, rho.tensor[ij], charge=None))
        if restrictor is not None:
            restrictor.apply(phit_p, phit)
        K[ij, ij:] = ops.dots(phit, rint[ij:], gd.dv)
    Kh = K.cpu().numpy()
    iu = np.triu_indices(P, 1)
    Kh[iu[1], iu[0]] = Kh[iu]
    return Kh, niter
