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
    for n in range(1, 4):
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
