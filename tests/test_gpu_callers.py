"""Callers of the hot path on the device against the oracle's restatements
of the same call sites (gpaw/wavefunctions/fd.py:50-52,
gpaw/hamiltonian.py:254-259,371-398, gpaw/density.py:217-234,
gpaw/xc/hybrid.py:148-152).

Exact equality where the arithmetic is element-wise or a pinned kernel; the
density renormalisation divides by an integral summed in a different
(deterministic) order than numpy, so it is compared to 1e-14 relative; the
pair-density solves to 1e-13 with equal iteration counts (the Poisson
tolerance of test_gpu_poisson.py).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
BOHR = 0.52917720859


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a)).to('cuda')


def _gds(N, pbc, h=0.2 / BOHR):
    from gpaw_amd.grid_descriptor import GridDescriptor
    return GridDescriptor((N,) * 3, (N * h,) * 3, pbc), O.Grid((N,) * 3, (N * h,) * 3, pbc)


@pytest.mark.parametrize('dtype,pbc', [(complex, True), (float, True), (complex, False)])
def test_apply_pseudo_hamiltonian(dtype, pbc):
    print('case pseudo-Hamiltonian', dtype, pbc, flush=True)
    from gpaw_amd.callers import apply_pseudo_hamiltonian
    from gpaw_amd.fd_operators import Laplace
    gd, og = _gds(24, pbc)
    rng = np.random.default_rng(7)
    shape = (5,) + tuple(og.n_c)
    psi = rng.standard_normal(shape)
    if dtype == complex:
        psi = psi + 1j * rng.standard_normal(shape)
    v = rng.standard_normal(tuple(og.n_c))
    ph = og.phase_cd([0.15, -0.3, 0.45]) if dtype == complex else None
    want = np.zeros_like(psi)
    O.apply_pseudo_hamiltonian(O.laplace(og, -0.5, 3, dtype=dtype), v, psi, want, ph)
    kin = Laplace(gd, -0.5, 3, dtype)
    got = dev(np.zeros_like(psi))
    apply_pseudo_hamiltonian(kin, dev(v), dev(psi), got, ph)
    assert np.array_equal(got.cpu().numpy(), want)
    # numpy in / out, one band
    out1 = np.zeros_like(psi[0])
    apply_pseudo_hamiltonian(kin, v, psi[0], out1, ph)
    assert np.array_equal(out1, want[0])


def test_apply_pseudo_hamiltonian_generic_stencil():
    """A non-orthogonal cell (generic stencil): the stencil pass and the
    potential pass run separately, with the same result as the oracle."""
    print('case pseudo-Hamiltonian generic', flush=True)
    from gpaw_amd.callers import apply_pseudo_hamiltonian
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    cell = np.array([[6.0, 0.0, 0.0], [3.0, 5.196152422706632, 0.0], [0.0, 0.0, 7.0]])
    gd = GridDescriptor((16, 16, 20), cell, True)
    og = O.Grid((16, 16, 20), cell, True)
    rng = np.random.default_rng(11)
    psi = rng.standard_normal((3, 16, 16, 20)) + 1j * rng.standard_normal((3, 16, 16, 20))
    v = rng.standard_normal((16, 16, 20))
    ph = og.phase_cd([0.1, 0.25, -0.2])
    want = np.zeros_like(psi)
    O.apply_pseudo_hamiltonian(O.laplace(og, -0.5, 2, dtype=complex), v, psi, want, ph)
    got = dev(np.zeros_like(psi))
    apply_pseudo_hamiltonian(Laplace(gd, -0.5, 2, complex), dev(v), dev(psi), got, ph)
    assert np.array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize('pbc', [True, False])
def test_interpolate_density(pbc):
    print('case interpolate density', pbc, flush=True)
    from gpaw_amd.callers import interpolate_density
    from gpaw_amd.transformers import Transformer
    gd, og = _gds(20, pbc)
    fine, ofine = gd.refine(), og.refine()
    rng = np.random.default_rng(8)
    nt_sG = rng.uniform(0.0, 1.0, (2,) + tuple(og.n_c))
    want = np.zeros((2,) + tuple(ofine.n_c))
    O.interpolate_density(O.Transformer(og, ofine, 3), ofine, nt_sG, want,
                          charge=-4.0, comp_charge=1.25)
    got = dev(np.zeros_like(want))
    interpolate_density(Transformer(gd, fine, 3), fine, dev(nt_sG), got,
                        charge=-4.0, comp_charge=1.25)
    g = got.cpu().numpy()
    if pbc:
        assert np.array_equal(g, want)
    else:
        assert np.abs(g - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize('pbc', [True, False])
def test_restrict_potential(pbc):
    print('case restrict potential', pbc, flush=True)
    from gpaw_amd.callers import restrict_potential
    from gpaw_amd.transformers import Transformer
    gd, og = _gds(20, pbc)
    fine, ofine = gd.refine(), og.refine()
    rng = np.random.default_rng(9)
    vt_sg = rng.standard_normal((2,) + tuple(ofine.n_c))
    vH = rng.standard_normal(tuple(ofine.n_c))
    want_G = np.zeros((2,) + tuple(og.n_c))
    want_g = vt_sg.copy()
    O.restrict_potential(O.Transformer(ofine, og, 3), want_g, vH, want_G)
    got_g = dev(vt_sg)
    got_G = dev(np.zeros_like(want_G))
    restrict_potential(Transformer(fine, gd, 3), got_g, dev(vH), got_G)
    assert np.array_equal(got_g.cpu().numpy(), want_g)
    assert np.array_equal(got_G.cpu().numpy(), want_G)


def test_pair_potentials():
    print('case pair potentials', flush=True)
    _torch()
    from gpaw_amd.callers import pair_potentials
    from gpaw_amd.poisson import PoissonSolver
    gd, og = _gds(32, False)
    rng = np.random.default_rng(10)
    centers = [rng.uniform(2.5 / BOHR, (32 * 0.2 - 2.5) / BOHR, 3) for _ in range(3)]
    rho = np.stack([O.gaussian_charges(og, [c], [0.8 / BOHR], [q])
                    for c, q in zip(centers, (1.0, 0.0, 1.0))])
    rho[1] = O.gaussian_charges(og, centers[:2], [0.8 / BOHR] * 2, [1.0, -1.0])
    charges = [-1.0, 0.0, -1.0]
    ops_ = O.PoissonSolver(nn=3, relax='J')
    ops_.set_grid_descriptor(og)
    ops_.initialize()
    want, n_want = O.pair_potentials(ops_, rho, charges, eps=1e-12)
    ps = PoissonSolver(nn=3, relax='J')
    ps.set_grid_descriptor(gd)
    ps.initialize()
    got, n_got = pair_potentials(ps, dev(rho), charges, eps=1e-12)
    assert n_got == n_want
    g = got.cpu().numpy()
    for p in range(3):
        assert np.abs(g[p] - want[p]).max() <= 1e-13 * np.abs(want[p]).max()


def test_rpa_coupling_periodic_modes():
    """Cos-mode pair densities on a periodic grid (a diagonal coupling
    matrix in closed form): same matrix as the oracle to 1e-13 of its
    largest entry, same iteration counts."""
    print('case rpa coupling periodic', flush=True)
    _torch()
    from gpaw_amd.callers import rpa_coupling
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import PoissonSolver
    N, h = 16, 0.3
    og = O.Grid((N,) * 3, (N * h,) * 3, True)
    gd = GridDescriptor((N,) * 3, (N * h,) * 3, True)
    i0 = np.indices((N,) * 3)[0]
    rho = np.stack([np.cos(2 * np.pi * m * i0 / N) for m in (1, 2, 3)])
    ops_ = O.PoissonSolver(nn=3, relax='J', eps=1e-20)
    ops_.set_grid_descriptor(og)
    ops_.initialize()
    want, n_want = O.rpa_coupling(ops_, rho)
    ps = PoissonSolver(nn=3, relax='J', eps=1e-20)
    ps.set_grid_descriptor(gd)
    ps.initialize()
    got, n_got = rpa_coupling(ps, dev(rho))
    assert n_got == n_want
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    assert np.array_equal(got, got.T)


def test_rpa_coupling_restricted_zero_boundary():
    """finegrid == 1 on a zero-boundary cell: pair densities solved on the
    fine grid (one exactly neutral, which takes the neutral branch of
    charge=None, and two with a net charge, which take the Gaussian
    branch), potentials restricted with K=6, integrated against coarse
    densities."""
    print('case rpa coupling restricted', flush=True)
    _torch()
    from gpaw_amd.callers import rpa_coupling
    from gpaw_amd.poisson import PoissonSolver
    from gpaw_amd.transformers import Transformer
    gd, og = _gds(32, False)
    cg, oc = gd.coarsen(), og.coarsen()
    rng = np.random.default_rng(12)
    lo, hi = 2.0 / BOHR, (32 * 0.2 - 2.0) / BOHR
    cs = [rng.uniform(lo, hi, 3) for _ in range(4)]
    sig = [0.7 / BOHR] * 2
    rho = np.stack([O.gaussian_charges(og, cs[:2], sig, [1.0, -1.0]),
                    O.gaussian_charges(og, cs[2:], sig, [1.0, -1.0]),
                    O.gaussian_charges(og, cs[1:3], sig, [1.0, 0.5])])
    rho_c = np.stack([O.gaussian_charges(oc, cs[:2], sig, [1.0, -1.0]),
                      O.gaussian_charges(oc, cs[2:], sig, [1.0, -1.0]),
                      O.gaussian_charges(oc, cs[1:3], sig, [1.0, 0.5])])
    rho[1] -= rho[1].mean()
    rho_c[1] -= rho_c[1].mean()
    ops_ = O.PoissonSolver(nn=3, relax='J')
    ops_.set_grid_descriptor(og)
    ops_.initialize()
    want, n_want = O.rpa_coupling(ops_, rho, gd=oc, restrictor=O.Transformer(og, oc, 3),
                                  rhot_pG=rho_c)
    ps = PoissonSolver(nn=3, relax='J')
    ps.set_grid_descriptor(gd)
    ps.initialize()
    got, n_got = rpa_coupling(ps, dev(rho), gd=cg, restrictor=Transformer(gd, cg, 3),
                              rhot_pG=dev(rho_c))
    assert n_got == n_want
    err = np.abs(got - want).max() / np.abs(want).max()
    print('coupling matrix: relative max-norm difference %.2e' % err, flush=True)
    assert err <= 1e-13
