"""Pins the oracle's multigrid V-cycle schedule (gpaw/poisson.py:45-116,
193-218,245-276) with linear algebra on tiny hierarchies.

Every level's operator L, restriction R and interpolation P is turned into a
dense matrix by applying the golden-pinned kernels to unit vectors.  One
V-cycle is then an affine map phi -> M phi + N rho, built here from textbook
multigrid algebra and the reference's parameters:

* weighted Jacobi with weight w and diagonal d is phi -> S phi + (w/d) rho with
  S = I - (w/d) L (c/bmgs/relax.c:47-70 rearranged); k sweeps give S^k and
  T_k = sum_{j<k} S^j (w/d);
* a level below the coarsest: pre-smooth a times, residual L phi - rho,
  restrict, solve the coarser level from phi = 0 (its map N_{l+1}),
  subtract the interpolated correction, post-smooth b times:
      M_l = S^b (I - P N_{l+1} R L) S^a
      N_l = S^b ((I - P N_{l+1} R L) T_a + P N_{l+1} R) + T_b
* the coarsest level only post-smooths: M = S^b, N = T_b;
* sweeps (a, b) = (2, 1) and w = 2/3 on level 0, (4, 4) and w = 1 below,
  (., 8) on the coarsest level (gpaw/poisson.py:64-89,111-113).

The oracle's iterate2 must match this map cycle after cycle, its error must
be dv * |L phi - rho|^2, and solve_neutral must stop at the first cycle whose
error is <= eps.  Mistakes planted in a copy of the oracle's V-cycle (swapped
sweep counts, weight 1 on level 0, a coarse phi kept from the previous cycle,
the correction added instead of subtracted, 4 coarsest sweeps) each miss the
dense map by orders of magnitude more than the tolerance.
"""
import numpy as np
import pytest

import oracle as O

# (N_c, pbc, nn, levels): 8^3 -> 4^3 -> 2^3 -> 1 periodic; 16^3 zero
# boundary (15^3 -> 7^3 -> 3^3 -> 1^3 arrays); 12^3 zero boundary (11^3 ->
# 5^3 -> 2^3, a coarsest level that one sweep does not solve)
HIERARCHIES = [((8, 8, 8), True, 1, 3), ((8, 8, 8), True, 3, 3),
               ((16, 16, 16), False, 1, 3), ((16, 16, 16), False, 3, 3),
               ((12, 12, 12), False, 3, 2)]
PRE0, POST0, W0 = 2, 1, 2.0 / 3.0    # gpaw/poisson.py:78-84
PRE, POST, W = 4, 4, 1.0             # gpaw/poisson.py:92-94
COARSEST = 8                         # gpaw/poisson.py:112-113


def _solver(N, pbc, nn, cls=None):
    g = O.Grid(N, np.array(N, float) * 0.3, pbc)
    ps = (cls or O.PoissonSolver)(nn=nn)
    ps.set_grid_descriptor(g)
    ps.initialize()
    return g, ps


def _dense_operator(op):
    n = int(np.prod(op.shape))
    eye = np.ascontiguousarray(np.eye(n).reshape((n,) + op.shape))
    out = np.zeros_like(eye)
    op.apply(eye, out)
    return out.reshape(n, n).T.copy()


def _dense_transfer(tr):
    nin = int(np.prod(tr.ngpin))
    nout = int(np.prod(tr.ngpout))
    eye = np.ascontiguousarray(np.eye(nin).reshape((nin,) + tr.ngpin))
    out = np.zeros((nin,) + tr.ngpout)
    tr.apply(eye, out)
    return out.reshape(nin, nout).T.copy()


def _jacobi(Lm, d, w, k):
    """k sweeps; d is the stencil's centre coefficient (on grids narrower
    than the stencil the dense diagonal also holds wrapped terms)."""
    n = Lm.shape[0]
    S = np.eye(n) - (w / d) * Lm
    T = np.zeros((n, n))
    Sk = np.eye(n)
    for _ in range(k):
        T = T + Sk * (w / d)
        Sk = S @ Sk
    return Sk, T


def _dense_vcycle(ps):
    """(M, N) of one V-cycle on level 0, from the reference's parameters."""
    Ls = [_dense_operator(op) for op in ps.operators]
    Rs = [_dense_transfer(t) for t in ps.restrictors]
    Ps = [_dense_transfer(t) for t in ps.interpolators]
    ds = [op.get_diagonal_element() for op in ps.operators]
    top = ps.levels
    _, Nc = _jacobi(Ls[top], ds[top], W if top > 0 else W0, COARSEST)
    Mc = None
    for l in range(top - 1, -1, -1):
        a, b, w = (PRE0, POST0, W0) if l == 0 else (PRE, POST, W)
        Sa, Ta = _jacobi(Ls[l], ds[l], w, a)
        Sb, Tb = _jacobi(Ls[l], ds[l], w, b)
        C = Ps[l] @ Nc @ Rs[l]
        K = np.eye(Ls[l].shape[0]) - C @ Ls[l]
        Mc = Sb @ K @ Sa
        Nc = Sb @ (K @ Ta + C) + Tb
    return Mc, Nc, Ls[0]


def _inputs(g, seed):
    rng = np.random.default_rng(seed)
    rho = rng.standard_normal(tuple(g.n_c))
    if g.pbc_c.all():
        rho -= rho.mean()
    return rng.standard_normal(tuple(g.n_c)), rho


def _max_rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize('N,pbc,nn,levels', HIERARCHIES)
def test_oracle_vcycle_equals_dense_map(N, pbc, nn, levels):
    g, ps = _solver(N, pbc, nn)
    assert ps.levels == levels
    assert ps.presmooths == [PRE0] + [PRE] * (levels - 1) + [COARSEST]
    assert ps.postsmooths == [POST0] + [POST] * (levels - 1) + [COARSEST]
    M, Nm, L0 = _dense_vcycle(ps)
    phi0, rho = _inputs(g, 11)
    phi = phi0.copy()
    errs = ps.run_vcycles(phi, rho, 1)
    x, r = phi0.ravel(), rho.ravel()
    want = []
    for k in range(4):
        x = M @ x + Nm @ r
        res = L0 @ x - r
        want.append(g.dv * res @ res)
    assert _max_rel(phi.ravel(), M @ phi0.ravel() + Nm @ r) < 1e-12
    # three more cycles: the coarse phi must start from zero every cycle
    errs += [ps.iterate2(ps.step) for _ in range(3)]
    assert _max_rel(phi.ravel(), x) < 1e-11
    assert np.allclose(errs, want, rtol=1e-9, atol=0)
    # consistency: the discrete solution is a fixed point (M + N L = I)
    assert np.abs(M + Nm @ L0 - np.eye(M.shape[0])).max() < 1e-9


@pytest.mark.parametrize('N,pbc,nn,levels', HIERARCHIES[1:3])
def test_oracle_stopping_rule(N, pbc, nn, levels):
    """niter = the first cycle whose error is <= eps (gpaw/poisson.py:202-205)."""
    g, ps = _solver(N, pbc, nn)
    M, Nm, L0 = _dense_vcycle(ps)
    _, rho = _inputs(g, 5)
    x, r, errs = np.zeros(rho.size), rho.ravel(), []
    xs = []
    for k in range(8):
        x = M @ x + Nm @ r
        res = L0 @ x - r
        errs.append(g.dv * res @ res)
        xs.append(x)
    for stop in (1, 3, 6):
        eps = 0.5 * (errs[stop - 1] + errs[stop - 2]) if stop > 1 else 2 * errs[0]
        g2, ps2 = _solver(N, pbc, nn)
        phi = g2.zeros()
        niter = ps2.solve_neutral(phi, rho.copy(), eps=eps)
        assert niter == stop
        want = xs[stop - 1].reshape(phi.shape)
        if pbc:
            want = want - want.mean()
        assert _max_rel(phi, want) < 1e-10


# ---- planted mistakes: each must miss the dense map -----------------------
class _Mutant(O.PoissonSolver):
    """The oracle's V-cycle with one planted mistake (`mutation`)."""
    mutation = None

    def iterate2(self, step, level=0):
        m = self.mutation
        residual = self.residuals[level]
        w = self.weights[level]
        if m == 'weight1' and level == 0:
            w = 1.0
        if level < self.levels:
            pre = self.presmooths[level]
            if m == 'swap' and level == 0:
                pre = self.postsmooths[level]
            self.operators[level].relax(self.relax_method, self.phis[level],
                                        self.rhos[level], pre, w)
            self.operators[level].apply(self.phis[level], residual)
            residual -= self.rhos[level]
            self.restrictors[level].apply(residual, self.rhos[level + 1])
            if m != 'keep_coarse':
                self.phis[level + 1][:] = 0.0
            self.iterate2(4.0 * step, level + 1)
            self.interpolators[level].apply(self.phis[level + 1], residual)
            if m == 'add_correction':
                self.phis[level] += residual
            else:
                self.phis[level] -= residual
        post = self.postsmooths[level]
        if m == 'swap' and level == 0:
            post = self.presmooths[level]
        if m == 'coarsest4' and level == self.levels:
            post = 4
        self.operators[level].relax(self.relax_method, self.phis[level],
                                    self.rhos[level], post, w)
        if level == 0:
            self.operators[level].apply(self.phis[level], residual)
            residual -= self.rhos[level]
            return np.dot(residual.ravel(), residual.ravel()) * self.dv


MUTATIONS = [None, 'swap', 'weight1', 'keep_coarse', 'add_correction',
             'coarsest4']


# On the periodic 8^3 hierarchy the coarsest level is one point whose
# restricted density is zero for a neutral rho, so its sweep count cannot
# matter there; the 12^3 zero-boundary hierarchy has 2^3 coarsest points.
@pytest.mark.parametrize('mutation,N,pbc,nn,levels',
                         [(m,) + HIERARCHIES[4] for m in MUTATIONS] +
                         [(m,) + HIERARCHIES[1] for m in MUTATIONS[:-1]])
def test_planted_schedule_mistakes_are_caught(mutation, N, pbc, nn, levels):
    g, ps = _solver(N, pbc, nn)
    M, Nm, _ = _dense_vcycle(ps)
    cls = type('M', (_Mutant,), {'mutation': mutation})
    g, mut = _solver(N, pbc, nn, cls)
    phi0, rho = _inputs(g, 3)
    phi = phi0.copy()
    mut.run_vcycles(phi, rho, 2)
    x = phi0.ravel()
    for _ in range(2):
        x = M @ x + Nm @ rho.ravel()
    err = _max_rel(phi.ravel(), x)
    if mutation is None:
        assert err < 1e-11
    else:
        assert err > 1e-6, mutation
