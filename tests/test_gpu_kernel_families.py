"""Every star-stencil kernel family against the oracle.

The library has four ways to evaluate a star stencil: the plane-sweeping
LDS tile kernel, the one-shot LDS box kernel, the register-queue column
kernel and the generic offset-list kernel.  The default choice (heuristic or
tuner) depends on the grid, so the ordinary parity tests reach only some of
them on each shape.  Here each family is forced in turn
(gpaw_stencil_kernel) on the same inputs and must reproduce the oracle's
bits exactly: applies of real and complex (Bloch-phase) data at ranges 1-4,
Jacobi sweeps, residuals, and the fused kinetic + potential apply.
"""
import numpy as np
import pytest

import oracle as O
from oracle.gpaw_oracle import Operator as OOp

pytestmark = pytest.mark.gpu

FAMILIES = ['default', 'tile', 'box', 'queue', 'generic']


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a)).to('cuda')


def host(t):
    return t.cpu().numpy()


class family:
    def __init__(self, kind):
        self.kind = kind

    def __enter__(self):
        from gpaw_amd import ops
        self.prev = ops.stencil_kernel(self.kind)

    def __exit__(self, *exc):
        from gpaw_amd import ops
        ops.stencil_kernel(self.prev)


def _fdop(coefs, offsets, gd, dtype):
    from gpaw_amd.fd_operators import FDOperator
    return FDOperator(coefs, offsets, gd, dtype)


def _gd(N, cell, pbc):
    from gpaw_amd.grid_descriptor import GridDescriptor
    return GridDescriptor(N, cell, pbc)


APPLY_CASES = [
    # (N_c, pbc, nn, dtype, nin)
    ((40, 36, 44), False, 3, float, 2),
    ((34, 33, 40), True, 2, complex, 2),
    ((24, 20, 70), True, 4, float, 1),
    ((48, 48, 48), True, 3, complex, 2),
    ((21, 17, 35), (True, False, True), 1, complex, 3),
    ((37, 41, 29), (False, True, False), 3, float, 3),
]


@pytest.mark.parametrize('kind', FAMILIES)
@pytest.mark.parametrize('N,pbc,nn,dtype,nin', APPLY_CASES)
def test_apply_every_family(kind, N, pbc, nn, dtype, nin):
    print('case apply', kind, N, pbc, nn, dtype, nin, flush=True)
    rng = np.random.default_rng(hash((N, nn, nin)) % 2 ** 32)
    cell = np.array(N) * 0.31
    og = O.Grid(N, cell, pbc)
    gd = _gd(N, cell, pbc)
    c, o = O.laplace_stencil(og, -0.5, nn)
    a = rng.uniform(-1, 1, (nin,) + tuple(og.n_c))
    ph = None
    if dtype == complex:
        a = a + 1j * rng.uniform(-1, 1, a.shape)
        ph = og.phase_cd(rng.uniform(-0.5, 0.5, 3))
    ref = np.zeros_like(a)
    OOp(c, o, og, dtype).apply(a, ref, ph)
    b = dev(np.zeros_like(a))
    with family(kind):
        _fdop(c, o, gd, dtype).apply(dev(a), b, ph)
    assert np.array_equal(host(b), ref)


@pytest.mark.parametrize('kind', FAMILIES)
@pytest.mark.parametrize('N,nn,pbc', [((95, 95, 95), 3, False), ((47, 47, 47), 1, False),
                                      ((40, 32, 36), 2, True)])
def test_jacobi_and_residual_every_family(kind, N, nn, pbc):
    print('case relax/residual', kind, N, nn, pbc, flush=True)
    rng = np.random.default_rng(sum(N) + nn)
    og = O.Grid(N, np.array(N) * 0.2, pbc)
    gd = _gd(N, np.array(N) * 0.2, pbc)
    c, o = O.laplace_stencil(og, -0.25 / np.pi, nn)
    f = rng.normal(size=tuple(og.n_c))
    s = rng.normal(size=tuple(og.n_c))
    ref = f.copy()
    OOp(c, o, og).relax(2, ref, s, 3, 2.0 / 3.0)
    rref = np.zeros_like(f)
    OOp(c, o, og).apply(f, rref)
    rref -= s
    op = _fdop(c, o, gd, float)
    fd, r = dev(f), dev(np.zeros_like(f))
    with family(kind):
        op.relax(2, fd, dev(s), 3, 2.0 / 3.0)
        op.operator.residual(dev(f), dev(s), r)
    assert np.array_equal(host(fd), ref)
    assert np.array_equal(host(r), rref)


@pytest.mark.parametrize('kind', FAMILIES)
def test_pseudo_hamiltonian_every_family(kind):
    print('case kinetic + potential', kind, flush=True)
    from gpaw_amd.callers import apply_pseudo_hamiltonian
    from gpaw_amd.fd_operators import Laplace
    N, cell = (36, 32, 40), (7.2, 6.4, 8.0)
    og, gd = O.Grid(N, cell, True), _gd(N, cell, True)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=(3,) + N) + 1j * rng.normal(size=(3,) + N)
    vt = rng.normal(size=N)
    ph = og.phase_cd([0.1, 0.27, -0.4])
    want = np.zeros_like(psi)
    O.apply_pseudo_hamiltonian(O.laplace(og, -0.5, 3, dtype=complex), vt, psi, want, ph)
    got = dev(np.zeros_like(psi))
    with family(kind):
        apply_pseudo_hamiltonian(Laplace(gd, -0.5, 3, complex), dev(vt), dev(psi), got, ph)
    assert np.array_equal(host(got), want)


TILE_SHAPES = [(256, 32, 1), (256, 32, 5), (256, 64, 10 ** 6), (512, 32, 10 ** 6),
               (512, 64, 10 ** 6), (512, 64, 8), (1024, 32, 10 ** 6), (1024, 64, 4)]


@pytest.mark.parametrize('threads,tz,planes', TILE_SHAPES)
@pytest.mark.parametrize('N,pbc,nn,dtype,nin', [
    ((64, 24, 64), True, 3, complex, 3),     # sweeps over whole periodic i0 axes
    ((7, 20, 36), True, 3, complex, 2),      # n0 = 2R + 1
    ((6, 18, 40), True, 3, complex, 2),      # n0 = 2R
    ((9, 33, 30), True, 2, complex, 2),
    ((11, 17, 66), (True, False, True), 1, complex, 2),
    ((20, 30, 40), False, 3, complex, 2),
    ((40, 36, 44), True, 3, float, 2)])
def test_tile_shapes_against_oracle(threads, tz, planes, N, pbc, nn, dtype, nin):
    """The tile kernel forced to each launch shape (block size, columns,
    planes per sweep, up to the whole i0 axis) against the oracle, bit for
    bit."""
    print('case tile shape', threads, tz, planes, N, pbc, nn, dtype.__name__, nin, flush=True)
    from gpaw_amd import ops
    rng = np.random.default_rng(sum(N) * 7 + nn)
    cell = np.array(N) * 0.3
    og = O.Grid(N, cell, pbc)
    gd = _gd(N, cell, pbc)
    c, o = O.laplace_stencil(og, -0.5, nn)
    a = rng.uniform(-1, 1, (nin,) + tuple(og.n_c))
    ph = None
    if dtype == complex:
        a = a + 1j * rng.uniform(-1, 1, a.shape)
        ph = og.phase_cd(rng.uniform(-0.5, 0.5, 3))
    ref = np.zeros_like(a)
    OOp(c, o, og, dtype).apply(a, ref, ph)
    b = dev(np.zeros_like(a))
    ops.tile_shape(threads, tz, planes)
    try:
        with family('tile'):
            _fdop(c, o, gd, dtype).apply(dev(a), b, ph)
    finally:
        ops.tile_shape(0)
    assert np.array_equal(host(b), ref)
