"""Closed-form and invariant checks that pin the oracle independently of the
golden fixtures (the reference's own test scripts, restated with asserts):

* gpaw/test/laplace.py       polynomial exactness of Laplace / Gradient
* gpaw/test/gradient.py      closed-form gradient values
* gpaw/test/transformations.py, gpaw/test/non_periodic.py
                             integral conservation of restrict / interpolate
* gpaw/test/poisson.py       translation symmetry of the periodic solution
* gpaw/test/gauss_func.py    Gaussian potential vs the analytic erf form
* Lagrange exactness of the interpolation weights (gpaw/transformers.py:163)
* the simple-cubic Madelung (Wigner) constant for the Ewald shift
"""
from math import cos, pi, sin

import numpy as np
import pytest

import oracle as O
from oracle.gpaw_oracle import Operator, Transformer, laplace

CELLS = [
    ('distorted hexagonal', 4,
     [(1, 0, 0), (1.02 * cos(pi / 3 - 0.02), 1.02 * sin(pi / 3 - 0.02), 0),
      (0, 0, 1.0)]),
    ('hexagonal', 4, [(1, 0, 0), (0.5, 3 ** 0.5 / 2, 0), (0, 0, 1.1)]),
    ('fcc', 6, [(0, 1, 1), (1, 0, 1), (1, 1, 0)]),
    ('fcc-alternative', 6, [(1, 0, 0), (0.5, 3 ** 0.5 / 2, 0),
                            (0.5, 3 ** 0.5 / 6, (2.0 / 3) ** 0.5)]),
    ('bcc', 4, [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]),
    ('sc', 3, [1.1, 1.02, 1.03]),
    ('distorted sc', 6, [(1, 0, 0), (0.01, 1, 0), (0, 0.02, 1)]),
    ('rocksalt', 6,
     [(2 * np.sqrt(1.0 / 3), np.sqrt(1.0 / 8), -np.sqrt(1.0 / 24)),
      (2 * np.sqrt(1.0 / 3), -np.sqrt(1.0 / 8), -np.sqrt(1.0 / 24)),
      (2 * np.sqrt(1.0 / 3), 0, np.sqrt(1.0 / 6))]),
    ('nasty', 6, [(1, 0, 0), (0.0001, 1.03, 0), (0.0001, 0.0001, 1.0)]),
    ('Mike', 6, 5 * np.array([(5.565 / 28, 0, 0), (0.0001 / 28, 5.565 / 28, 0),
                              (0.0001 / 24, 0.0001 / 24, 4.684 / 24)])),
    ('MnO', 6, [(1, 0.5, 0.5), (0.5, 1, 0.5), (0.5, 0.5, 1)]),
]


@pytest.mark.parametrize('name,D,cell', CELLS, ids=[c[0] for c in CELLS])
def test_laplace_gradient_polynomial_exactness(name, D, cell):
    for n in range(1, 6):
        N = 2 * n + 2
        g = O.Grid((N, N, N), cell)
        b = g.zeros()
        r_gv = g.grid_point_coordinates().transpose((1, 2, 3, 0))
        r_gv = r_gv - g.cell_cv.sum(0) / 2
        c, o = O.laplace_stencil(g, 1.0, n)
        assert len(c) == D * 2 * n + 1
        lap = Operator(c, o, g)
        grads = [Operator(*O.gradient_stencil(g, v, 1.0, n), g)
                 for v in range(3)]
        for m in range(0, 2 * n + 1):
            for ix in range(m + 1):
                for iy in range(m - ix + 1):
                    iz = m - ix - iy
                    a = np.ascontiguousarray((r_gv ** (ix, iy, iz)).prod(3))
                    r = 2.0 if (m == 2 and max(ix, iy, iz) == 2) else 0.0
                    lap.apply(a, b)
                    assert abs(b[n + 1, n + 1, n + 1] - r) < 2e-12
                    for v in range(3):
                        grads[v].apply(a, b)
                        r = 1.0 if (m == 1 and (ix, iy, iz)[v] == 1) else 0.0
                        assert abs(b[n + 1, n + 1, n + 1] - r) < 2e-12


def test_gradient_closed_forms():
    g = O.Grid((8, 1, 1), (8.0, 1.0, 1.0))
    a = g.zeros()
    a[:, 0, 0] = np.arange(8)
    d = g.zeros()
    Operator(*O.gradient_stencil(g, 0), g).apply(a, d)
    assert d[3, 0, 0] == 1.0 and np.sum(d[:, 0, 0]) == 0.0
    g = O.Grid((1, 8, 1), (1.0, 8.0, 1.0), (1, 0, 1))
    a = g.zeros()
    a[0, :, 0] = np.arange(g.beg_c[1], g.end_c[1]) - 1
    d = g.zeros()
    Operator(*O.gradient_stencil(g, 1), g).apply(a, d)
    assert d[0, 0, 0] == 0.5 and np.sum(d[0, :, 0]) == 3.0
    g = O.Grid((1, 7, 1), ((1.0, 0.0, 0.0), (5.0, 5.0, 0.0), (0.0, 0.0, 0.7)))
    a = g.zeros()
    a[0, :, 0] = np.arange(g.beg_c[1], g.end_c[1]) - 1
    d = g.zeros()
    Operator(*O.gradient_stencil(g, 1), g).apply(a, d)
    assert d[0, 0, 0] == -3.5 and abs(np.sum(d[0, :, 0])) < 1e-12


@pytest.mark.parametrize('dtype', [float, complex])
def test_transformer_integral_conservation(dtype):
    n = 20
    g = O.Grid((n, n, n))
    rng = np.random.RandomState(8)
    a = g.zeros(dtype=dtype)
    a.real = rng.random_sample(a.shape)
    if dtype == complex:
        a.imag = rng.random_sample(a.shape)
    phase = np.ones((3, 2), complex)
    for g2 in (g.refine(), g.coarsen()):
        b = g2.zeros(dtype=dtype)
        for k in (2, 4, 6, 8):
            Transformer(g, g2, k // 2, dtype).apply(a, b, phase)
            assert abs(g.integrate(a) - g2.integrate(b)) < 1e-14


def test_transformer_non_periodic_conservation():
    n = 20
    g1 = O.Grid((n, n, n), (8.0, 8.0, 8.0), pbc_c=0)
    a1 = g1.zeros()
    a1[:] = np.random.RandomState(8).random_sample(a1.shape)
    g2 = g1.refine()
    a2 = g2.zeros()
    Transformer(g1, g2).apply(a1, a2)
    assert abs(g1.integrate(a1) - g2.integrate(a2)) < 1e-10
    for sl in (0, -1):
        a2[sl] = 0.0
        a2[:, sl] = 0.0
        a2[:, :, sl] = 0.0
    Transformer(g2, g1).apply(a2, a1)
    assert abs(g1.integrate(a1) - g2.integrate(a2)) < 1e-10


@pytest.mark.parametrize('nn', [1, 2, 3, 4])
def test_interpolation_is_lagrange_exact(nn):
    """Midpoints of a degree-(2nn-1) polynomial are reproduced exactly away
    from the boundary: this pins every interpolation weight."""
    K = 2 * nn
    gc = O.Grid((32, 4, 4), (32.0, 4.0, 4.0), (False, True, True))
    gf = gc.refine()
    x = gc.grid_point_coordinates()[0]
    a = np.ascontiguousarray(0.3 + 0.7 * x - 0.05 * x ** 2 + 0.002 * x ** 3
                             if K >= 4 else 0.3 + 0.7 * x)
    b = gf.zeros()
    Transformer(gc, gf, nn).apply(a, b)
    xf = gf.grid_point_coordinates()[0]
    ref = (0.3 + 0.7 * xf - 0.05 * xf ** 2 + 0.002 * xf ** 3
           if K >= 4 else 0.3 + 0.7 * xf)
    inner = slice(2 * K, -2 * K)
    assert np.abs(b[inner] - ref[inner]).max() < 1e-10


def test_madelung_simple_cubic():
    # Wigner-lattice constant of a simple cubic lattice in jellium
    assert abs(O.madelung(np.eye(3) * 3.0) * 3.0 - 2.837297479480) < 1e-9


def test_poisson_periodic_symmetry():
    """gpaw/test/poisson.py: two identical charges at (0,0,0) and the cell
    centre give equal potentials there (Jacobi)."""
    N = 16
    L = 2.87 / 0.529177
    g = O.Grid((N, N, N), (L, L, L))
    f = O.gaussian_charges(g, [(0.0, 0.0, 0.0)], [0.8], [1.0])
    rho = f + np.roll(f, (8, 8, 8), (0, 1, 2))
    rho -= g.integrate(rho) / L ** 3
    ps = O.PoissonSolver(nn=1, relax='J')
    ps.set_grid_descriptor(g)
    ps.initialize()
    phi = g.zeros()
    ps.solve(phi, rho, charge=0)
    assert abs(phi[0, 0, 0] - phi[8, 8, 8]) < 1e-12


def test_poisson_gaussian_potential():
    """gpaw/test/gauss_func.py (L=0): the charged zero-boundary branch."""
    g = O.Grid((48, 48, 48), (20.0, 20.0, 20.0), 0)
    ps = O.PoissonSolver(nn=3)
    ps.set_grid_descriptor(g)
    ps.initialize()
    ng, vg = O.gpaw_oracle.gaussian_monopole(g)
    pot = g.zeros()
    ps.solve(pot, ng, charge=None, zero_initial_phi=True)
    assert g.integrate((pot - vg) ** 2) ** 0.5 < 0.6


def test_poisson_residual_decreases():
    g = O.Grid((24, 24, 24), (8.0, 8.0, 8.0), 0)
    L = np.diag(g.cell_cv)
    rho = O.gaussian_charges(g, [L * 0.4, L * 0.6], [1.0, 1.2], [1.0, -1.0])
    ps = O.PoissonSolver(nn=3)
    ps.set_grid_descriptor(g)
    ps.initialize()
    errs = ps.run_vcycles(g.zeros(), rho, 6)
    assert all(b < 0.5 * a for a, b in zip(errs, errs[1:]))


def _guc_poly_error(c, o, g, n):
    """Largest deviation from the exact Laplacian at the central point over
    every monomial of degree <= 2n (gpaw/test/laplace.py's check)."""
    lap = Operator(c, o, g)
    r_gv = g.grid_point_coordinates().transpose((1, 2, 3, 0)) - g.cell_cv.sum(0) / 2
    b = g.zeros()
    worst = 0.0
    for m in range(0, 2 * n + 1):
        for ix in range(m + 1):
            for iy in range(m - ix + 1):
                iz = m - ix - iy
                a = np.ascontiguousarray((r_gv ** (ix, iy, iz)).prod(3))
                r = 2.0 if (m == 2 and max(ix, iy, iz) == 2) else 0.0
                lap.apply(a, b)
                worst = max(worst, abs(b[n + 1, n + 1, n + 1] - r))
    return worst


@pytest.mark.parametrize('name,D,cell', CELLS, ids=[c[0] for c in CELLS])
def test_guc_laplace_polynomial_exactness(name, D, cell):
    """GUCLaplace (the newgucstencil=False Laplacian, gpaw/fd_operators.py:
    317-367) is exact on every monomial of degree <= 2n, n = 1..5, on the
    eleven cells of gpaw/test/laplace.py; its mixed-derivative weights are
    derived in the oracle from first-difference weights, not tabulated."""
    for n in range(1, 6):
        g = O.Grid((2 * n + 2,) * 3, cell)
        c, o = O.guc_stencil(g, 1.0, n)
        assert _guc_poly_error(c, o, g, n) < 2e-12, (name, n)


def test_guc_laplace_planted_mistakes_are_caught():
    """A dropped mirror pair, a mixed term without its factor 2, and a
    flipped corner sign each break the exactness check."""
    g = O.Grid((8, 8, 8), CELLS[0][2])
    n = 3
    c, o = O.guc_stencil(g, 1.0, n)
    assert _guc_poly_error(c, o, g, n) < 2e-12
    mixed = np.flatnonzero((o != 0).sum(1) == 2)
    assert len(mixed) > 0
    # mixed term halved
    c2 = c.copy()
    c2[mixed] *= 0.5
    assert _guc_poly_error(c2, o, g, n) > 1e-3
    # one corner's sign flipped
    c3 = c.copy()
    c3[mixed[0]] *= -1
    assert _guc_poly_error(c3, o, g, n) > 1e-3
    # the (j, i) mirror of an (i, j) pair dropped: offsets whose two nonzero
    # components differ in magnitude, taken where the larger one is first
    drop = [p for p in mixed
            if abs(o[p][o[p] != 0][0]) > abs(o[p][o[p] != 0][1])]
    assert drop
    keep = np.setdiff1d(np.arange(len(c)), drop)
    assert _guc_poly_error(c[keep], o[keep], g, n) > 1e-3


def test_first_derivative_weights_closed_forms():
    from fractions import Fraction as F
    w = O.gpaw_oracle._first_derivative_weights
    assert w(1) == [F(1, 2)]
    assert w(2) == [F(2, 3), F(-1, 12)]
    assert w(3) == [F(3, 4), F(-3, 20), F(1, 60)]


@pytest.mark.parametrize('cell', [[3.0, 3.0, 3.0], [3.0, 3.3, 2.7]])
def test_mehrstellen_identity(cell):
    """LaplaceA / LaplaceB (gpaw/fd_operators.py:192-229) form the compact
    fourth-order (Collatz "Mehrstellen") pair: A p = -scale * B (nabla^2 p)
    holds exactly for every polynomial of degree <= 5 (the truncation errors
    of both sides agree through h^2 and differ first in sixth derivatives),
    and fails at degree 6, which shows the check has teeth."""
    g = O.Grid((10, 11, 9), cell, False)
    r = g.grid_point_coordinates() - np.array(cell)[:, None, None, None] / 2
    x, y, z = r
    inner = (slice(1, -1),) * 3
    for scale in (1.0, 0.25 / np.pi):
        A = Operator(*O.mehrstellen_a(g, scale), g)
        B = Operator(*O.mehrstellen_b(), g)
        for deg in range(0, 7):
            worst = 0.0
            for ix in range(deg + 1):
                for iy in range(deg - ix + 1):
                    iz = deg - ix - iy
                    p = np.ascontiguousarray(x ** ix * y ** iy * z ** iz)
                    lap = np.zeros_like(p)
                    if ix >= 2:
                        lap += ix * (ix - 1) * x ** (ix - 2) * y ** iy * z ** iz
                    if iy >= 2:
                        lap += iy * (iy - 1) * x ** ix * y ** (iy - 2) * z ** iz
                    if iz >= 2:
                        lap += iz * (iz - 1) * x ** ix * y ** iy * z ** (iz - 2)
                    Ap = g.zeros()
                    A.apply(p, Ap)
                    Bl = g.zeros()
                    B.apply(np.ascontiguousarray(lap), Bl)
                    ref = scale * Bl[inner]
                    worst = max(worst, np.abs(Ap[inner] + ref).max() /
                                max(1.0, np.abs(ref).max()))
            if deg <= 5:
                assert worst < 1e-13, (scale, deg, worst)
            else:
                assert worst > 1e-4, (scale, deg, worst)


def test_mehrstellen_planted_mistakes_are_caught():
    """A wrong face weight or edge weight in A, or a wrong centre weight in
    B, breaks the degree-4 identity."""
    cell = [3.0, 3.3, 2.7]
    g = O.Grid((10, 11, 9), cell, False)
    r = g.grid_point_coordinates() - np.array(cell)[:, None, None, None] / 2
    x, y, z = r
    inner = (slice(1, -1),) * 3
    p = np.ascontiguousarray(x ** 2 * y ** 2 + 0.3 * z ** 4 + x * y * z)
    lap = np.ascontiguousarray(2 * y ** 2 + 2 * x ** 2 + 3.6 * z ** 2)
    ca, oa = O.mehrstellen_a(g, 1.0)
    cb, ob = O.mehrstellen_b()

    def miss(ca, cb):
        Ap = g.zeros()
        Operator(ca, oa, g).apply(p, Ap)
        Bl = g.zeros()
        Operator(cb, ob, g).apply(lap, Bl)
        return np.abs(Ap[inner] + Bl[inner]).max()

    assert miss(ca, cb) < 1e-12
    for k in (1, 7):     # a face weight, an edge weight
        c2 = ca.copy()
        c2[k] *= 1.01
        assert miss(c2, cb) > 1e-4
    c3 = cb.copy()
    c3[0] = 0.6
    assert miss(ca, c3) > 1e-4
