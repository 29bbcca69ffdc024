"""Randomised comparison of the oracle restatement with the reference's own C
(oracle/_ref, built in place from the reference tree).  Skipped when the
reference build is absent (e.g. on a machine without the reference)."""
import numpy as np
import pytest

import oracle as O
from oracle.gpaw_oracle import Operator, Transformer

pytestmark = pytest.mark.skipif(not O.Kernels.available('reference'),
                                reason='oracle/_ref not built')

PBCS = [(True, True, True), (False, False, False), (True, False, True),
        (False, True, False)]


@pytest.mark.parametrize('seed', range(6))
def test_operator_random(seed):
    rng = np.random.default_rng(seed)
    pbc = PBCS[seed % 4]
    N = tuple(int(x) for x in rng.integers(4, 9, 3) * 2)
    h = 0.3 + 0.1 * rng.random()
    g = O.Grid(N, np.array(N) * h, pbc)
    nn = int(rng.integers(1, 4))
    c, o = O.laplace_stencil(g, rng.normal(), nn)
    for dtype in (float, complex):
        a = rng.normal(size=(2,) + tuple(g.n_c))
        if dtype == complex:
            a = a + 1j * rng.normal(size=a.shape)
        ph = g.phase_cd(rng.uniform(-0.5, 0.5, 3))
        outs = []
        for kind in ('port', 'reference'):
            b = np.zeros_like(a)
            Operator(c, o, g, dtype, kernels=kind).apply(a, b, ph)
            outs.append(b)
        assert np.array_equal(outs[0], outs[1])
    s = rng.normal(size=tuple(g.n_c))
    for method in (1, 2):
        outs = []
        for kind in ('port', 'reference'):
            f = np.ascontiguousarray(rng.normal(size=tuple(g.n_c))) * 0 + s
            Operator(c, o, g, kernels=kind).relax(method, f, s * 0.5, 2, 0.7)
            outs.append(f)
        assert np.array_equal(outs[0], outs[1])


def _ref_interp_overflows(gin, gout, nn):
    """The reference's interpolation writes its first (z) pass into the
    output array (c/bmgs/interpolate.c:158-161); the size check of
    gpaw/transformers.py:48-49 ignores the extra padding of non-periodic
    axes, so small non-periodic grids with large nn overrun the output.
    Such cases corrupt the reference's heap and are not compared."""
    pad, npad, skip, interp = O.transformer_plan(gin, gout, nn)
    if not interp:
        return False
    m = np.array(gin.n_c) + pad.sum(1)
    q2 = (m[2] - (2 * nn - 1)) * 2 - skip[2, 0] + skip[2, 1]
    return m[0] * m[1] * q2 > np.prod(gout.n_c)


@pytest.mark.parametrize('seed', range(8))
def test_transformer_random(seed):
    rng = np.random.default_rng(100 + seed)
    pbc = PBCS[seed % 4]
    nn = 1 + seed % 4
    Nc = tuple(int(x) for x in rng.integers(3, 7, 3) * 2)
    gc = O.Grid(Nc, np.array(Nc) * 0.5, pbc)
    gf = gc.refine()
    for gin, gout in ((gf, gc), (gc, gf)):
        try:
            Transformer(gin, gout, nn)
        except AssertionError:
            continue      # grid too small for this nn
        if _ref_interp_overflows(gin, gout, nn):
            continue
        for dtype in (float, complex):
            a = rng.normal(size=(2,) + tuple(gin.n_c))
            if dtype == complex:
                a = a + 1j * rng.normal(size=a.shape)
            ph = gin.phase_cd(rng.uniform(-0.5, 0.5, 3))
            outs = []
            for kind in ('port', 'reference'):
                b = np.zeros((2,) + tuple(gout.n_c), dtype)
                Transformer(gin, gout, nn, dtype, kernels=kind).apply(a, b, ph)
                outs.append(b)
            assert np.array_equal(outs[0], outs[1])


@pytest.mark.skipif(not O.Kernels.available('reference_mt'),
                    reason='threaded reference build absent')
def test_threaded_reference_is_identical(monkeypatch):
    """The reference's GPAW_OMP_MONLY build (bmgs_fd / restrict /
    interpolate over OMP_NUM_THREADS pthreads, the all-cores CPU baseline)
    gives the serial build's bits: stencil applies, K = 2..8 transfers and a
    Poisson solve with its iteration count."""
    monkeypatch.setenv('OMP_NUM_THREADS', '4')
    rng = np.random.default_rng(31)
    g = O.Grid((20, 18, 22), (7.0, 6.5, 8.0), (True, False, True))
    a = rng.standard_normal((2,) + tuple(g.n_c))
    out = {}
    for kind in ('reference', 'reference_mt'):
        b = np.zeros_like(a)
        O.laplace(g, 1.0, 3, kernels=kind).apply(a, b)
        c = g.coarsen()
        r = np.zeros((2,) + tuple(c.n_c))
        O.Transformer(g, c, 3, kernels=kind).apply(a, r)
        f = g.refine()
        i = np.zeros((2,) + tuple(f.n_c))
        O.Transformer(g, f, 2, kernels=kind).apply(a, i)
        pg = O.Grid((32,) * 3, (12.0,) * 3, False)
        ps = O.PoissonSolver(nn=3, relax='J', kernels=kind)
        ps.set_grid_descriptor(pg)
        ps.initialize()
        rho = O.gaussian_charges(pg, [(5.0, 6.0, 6.5), (7.0, 6.0, 5.5)], [1.0, 1.1], [1.0, -1.0])
        phi = pg.zeros()
        n = ps.solve(phi, rho, eps=1e-14)
        out[kind] = (b, r, i, phi, n)
    for x, y in zip(out['reference'], out['reference_mt']):
        assert np.array_equal(x, y)
