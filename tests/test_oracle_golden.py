"""The oracle restatement reproduces the reference's own C bit for bit.

Golden fixtures (tests/golden/*.npz) were written by tests/golden/make_golden.py
from the reference's c/bmgs/bmgs.c + c/bc.c compiled in place.  Here our
restatement (oracle/liboracle.so) must reproduce every one exactly.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from oracle.gpaw_oracle import Operator, Transformer

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _files(prefix):
    return sorted(glob.glob(os.path.join(GOLD, prefix + '*.npz')))


def _grid(d, key='N_c'):
    return O.Grid(d[key], d['cell_cv'], d['pbc_c'])


@pytest.mark.parametrize('path', _files('op_'), ids=os.path.basename)
def test_operator_fixture(path):
    d = np.load(path)
    g = _grid(d)
    if str(d['kind']) == 'relax':
        op = Operator(d['coefs'], d['offsets_pc'], g, kernels='port')
        f = d['f0'].copy()
        op.relax(int(d['method']), f, d['s'], int(d['nrelax']), float(d['w']))
        assert np.array_equal(f, d['f'])
        return
    a = d['a']
    op = Operator(d['coefs'], d['offsets_pc'], g, a.dtype.type,
                  kernels='port')
    b = np.zeros_like(a)
    op.apply(a, b, d['phase_cd'] if 'phase_cd' in d else None)
    assert np.array_equal(b, d['b'])


@pytest.mark.parametrize('path', _files('tr_'), ids=os.path.basename)
def test_transformer_fixture(path):
    d = np.load(path)
    gi = O.Grid(d['N_in'], d['cell_cv'], d['pbc_c'])
    go = O.Grid(d['N_out'], d['cell_cv'], d['pbc_c'])
    a = d['a']
    t = Transformer(gi, go, int(d['nn']), a.dtype.type, kernels='port')
    b = np.zeros_like(d['b'])
    t.apply(a, b, d['phase_cd'] if 'phase_cd' in d else None)
    assert np.array_equal(b, d['b'])


@pytest.mark.parametrize('path', _files('poisson_'), ids=os.path.basename)
def test_poisson_vcycle_fixture(path):
    d = np.load(path)
    g = _grid(d)
    ps = O.PoissonSolver(nn=3, relax='J', kernels='port')
    ps.set_grid_descriptor(g)
    ps.initialize()
    phi = g.zeros()
    errs = ps.run_vcycles(phi, d['rho'], len(d['errors']))
    assert np.array_equal(phi, d['phi'])
    assert np.array_equal(np.array(errs), d['errors'])


def test_fixture_inventory():
    # every kernel family of the hot path is covered by at least one fixture
    names = ' '.join(os.path.basename(p) for p in _files(''))
    for key in ('lap_nn1', 'lap_nn2', 'lap_nn3', 'bloch', 'hex', 'mehrstellen',
                'grad', 'relax_jacobi', 'relax_gs', 'restrict_k2',
                'restrict_k4', 'restrict_k6', 'restrict_k8', 'interp_k2',
                'interp_k4', 'interp_k6', 'interp_k8', 'poisson'):
        assert key in names, key
