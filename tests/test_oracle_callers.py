"""The oracle's restatements of the hot path's callers (Hamiltonian on
bands, density interpolation, potential restriction, pair-density solves),
pinned on the CPU by what the mathematics fixes: closed-form eigenvalues of
the nn=1 Laplacian, charge conservation and renormalisation, constant
preservation and linearity of the transfers, and linearity of the solver."""
import numpy as np

import oracle as O


def _grid(N, pbc, h=0.3):
    return O.Grid((N,) * 3, (N * h,) * 3, pbc)


def test_pseudo_hamiltonian_closed_form():
    """psi = cos(2 pi m i0 / N) on a periodic grid is an eigenfunction of the
    nn=1 Laplacian: L psi = (2 cos t - 2) / h^2 psi.  With the local
    potential v(i1) the result is (-0.5 (2 cos t - 2) / h^2 + v) psi
    pointwise, for every band (band b uses m = b + 1)."""
    N, h = 12, 0.3
    g = _grid(N, True, h)
    kin = O.laplace(g, -0.5, 1)
    i0, i1, _ = np.indices((N, N, N))
    v = 0.3 + np.cos(2 * np.pi * i1 / N)
    psi = np.stack([np.cos(2 * np.pi * (b + 1) * i0 / N) for b in range(3)])
    out = np.zeros_like(psi)
    O.apply_pseudo_hamiltonian(kin, v, psi, out)
    for b in range(3):
        t = 2 * np.pi * (b + 1) / N
        want = (-0.5 * (2 * np.cos(t) - 2) / h ** 2 + v) * psi[b]
        assert np.abs(out[b] - want).max() < 1e-12


def test_pseudo_hamiltonian_complex_bloch():
    """Complex bands with Bloch phases: psi = exp(i 2 pi (m + k) i0 / N) has
    the eigenvalue (2 cos t - 2) / h^2 with t = 2 pi (m + k) / N when the
    phases of phase_cd(k) are applied at the cell boundary."""
    N, h = 10, 0.25
    g = _grid(N, True, h)
    kin = O.laplace(g, -0.5, 1, dtype=complex)
    k = np.array([0.2, 0.0, 0.0])
    ph = g.phase_cd(k)
    i0 = np.indices((N, N, N))[0]
    v = np.full((N, N, N), -0.7)
    psi = np.exp(2j * np.pi * (1 + k[0]) * i0 / N)[None]
    out = np.zeros_like(psi)
    O.apply_pseudo_hamiltonian(kin, v, psi, out, ph)
    t = 2 * np.pi * (1 + k[0]) / N
    want = (-0.5 * (2 * np.cos(t) - 2) / h ** 2 - 0.7) * psi[0]
    assert np.abs(out[0] - want).max() < 1e-12


def _density(g, seed, ns=2):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (ns,) + tuple(g.n_c))


def test_interpolate_density_conserves_on_periodic_cells():
    """'With periodic boundary conditions, the interpolation will conserve
    the number of electrons' (gpaw/density.py:226-227): per spin."""
    gd = _grid(12, True)
    fine = gd.refine()
    nt_sG = _density(gd, 1)
    nt_sg = np.zeros((2,) + tuple(fine.n_c))
    O.interpolate_density(O.Transformer(gd, fine, 3), fine, nt_sG, nt_sg,
                          charge=5.0)
    assert np.allclose(fine.integrate(nt_sg), gd.integrate(nt_sG), rtol=1e-12)


def test_interpolate_density_renormalises_zero_boundary_cells():
    gd = _grid(12, False)
    fine = gd.refine()
    nt_sG = _density(gd, 2)
    nt_sg = np.zeros((2,) + tuple(fine.n_c))
    O.interpolate_density(O.Transformer(gd, fine, 3), fine, nt_sG, nt_sg,
                          charge=-3.0, comp_charge=0.5)
    assert abs(fine.integrate(nt_sg).sum() - 2.5) < 1e-12


def test_restrict_potential_constant_and_linear():
    """K=6 restriction keeps constants (0.5 (1 + 2 sum w) = 1) and is
    linear; vt_g += vHt_g happens in place before it."""
    fine = _grid(16, True)
    gd = fine.coarsen()
    R = O.Transformer(fine, gd, 3)
    rng = np.random.default_rng(3)
    vH = np.full(tuple(fine.n_c), 0.25)
    vt_sg = np.stack([np.full(tuple(fine.n_c), 1.0), np.full(tuple(fine.n_c), -2.0)])
    vt_sG = np.zeros((2,) + tuple(gd.n_c))
    O.restrict_potential(R, vt_sg, vH, vt_sG)
    assert np.abs(vt_sG[0] - 1.25).max() < 1e-14
    assert np.abs(vt_sG[1] + 1.75).max() < 1e-14
    assert np.abs(vt_sg[1] + 1.75).max() == 0.0
    a = rng.standard_normal(tuple(fine.n_c))
    b = rng.standard_normal(tuple(fine.n_c))
    vt = np.stack([a.copy()])
    out = np.zeros((1,) + tuple(gd.n_c))
    O.restrict_potential(R, vt, b, out)
    ra = np.zeros(tuple(gd.n_c))
    rb = np.zeros(tuple(gd.n_c))
    R.apply(a, ra)
    R.apply(b, rb)
    assert np.abs(out[0] - (ra + rb)).max() < 1e-13 * np.abs(ra + rb).max()


def test_pair_potentials_are_minus_the_solutions():
    """solve(v, -rho) on neutral pair densities gives -solve(phi, rho)."""
    g = _grid(16, False, 0.4)
    ps = O.PoissonSolver(nn=3, relax='J')
    ps.set_grid_descriptor(g)
    ps.initialize()
    rng = np.random.default_rng(4)
    rho = rng.standard_normal((2,) + tuple(g.n_c))
    rho -= rho.mean(axis=(1, 2, 3), keepdims=True)
    vt, niter = O.pair_potentials(ps, rho, [0.0, 0.0], eps=1e-14)
    for p in range(2):
        phi = g.zeros()
        ps.solve(phi, rho[p].copy(), charge=0.0, eps=1e-14, zero_initial_phi=True)
        assert np.abs(vt[p] + phi).max() <= 1e-12 * np.abs(phi).max()


def _cos_modes(g, ms):
    i0 = np.indices(tuple(g.n_c))[0]
    return np.stack([np.cos(2 * np.pi * m * i0 / g.n_c[0]) for m in ms])


def _mu(t, h):
    """-(eigenvalue) of the nn=3 Laplacian for cos(t i0): the table
    [-49/18, 3/2, -3/20, 1/90] of gpaw/fd_operators.py:286-292."""
    c = [-49 / 18, 3 / 2, -3 / 20, 1 / 90]
    return -(c[0] + 2 * sum(c[k] * np.cos(k * t) for k in (1, 2, 3))) / h ** 2


def test_rpa_coupling_closed_form():
    """Pair densities cos(2 pi m i0 / N) on a periodic grid are eigenmodes of
    the solver's operator -1/(4 pi) L: phi = 4 pi / mu(t) rho, so the
    coupling matrix is diagonal with K[m, m] = 4 pi / mu(t) * V / 2.  With
    the K=6 restriction (finegrid == 1) of phi to the coarse grid and coarse
    cos modes, the diagonal carries the restriction's transfer factor
    0.5 (1 + 2 sum_j w_j cos(j t)), w = (0.5859375, -0.09765625, 0.01171875)."""
    N, h, ms = 16, 0.3, (1, 2, 3)
    g = _grid(N, True, h)
    V = (N * h) ** 3
    ps = O.PoissonSolver(nn=3, relax='J', eps=1e-24)
    ps.set_grid_descriptor(g)
    ps.initialize()
    rho = _cos_modes(g, ms)
    K, niter = O.rpa_coupling(ps, rho)
    want = np.diag([4 * np.pi / _mu(2 * np.pi * m / N, h) * V / 2 for m in ms])
    assert np.abs(K - want).max() <= 1e-12 * np.abs(want).max()
    assert all(n > 0 for n in niter)
    # finegrid == 1: solve on the fine grid, restrict, integrate on the coarse
    c = g.coarsen()
    R = O.Transformer(g, c, 3)
    rho_c = _cos_modes(c, ms)
    Kc, _ = O.rpa_coupling(ps, rho, gd=c, restrictor=R, rhot_pG=rho_c)
    w = (0.5859375, -0.09765625, 0.01171875)
    r = [0.5 * (1 + 2 * sum(wj * np.cos(j * 2 * np.pi * m / N) for wj, j in zip(w, (1, 3, 5))))
         for m in ms]
    want_c = np.diag([4 * np.pi / _mu(2 * np.pi * m / N, h) * r[i] * V / 2
                      for i, m in enumerate(ms)])
    assert np.abs(Kc - want_c).max() <= 1e-12 * np.abs(want_c).max()
