"""Edge cases of the device kernels against the oracle: the smallest grids a
stencil or transformer accepts (periodic extent equal to the stencil range,
one-point axes), the largest band count per call, and the errors beyond it.
Exact equality, as for every FP64 kernel."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def dev(a):
    return _torch().from_numpy(np.ascontiguousarray(a)).to('cuda')


@pytest.mark.parametrize('N,nn,pbc,dtype', [
    ((3, 4, 5), 3, True, float), ((3, 3, 3), 3, True, complex),
    ((1, 1, 1), 1, True, float), ((1, 2, 7), 1, True, complex),
    ((2, 3, 4), 1, False, float), ((4, 4, 4), 3, False, complex),
    ((1, 64, 1), 1, True, float), ((7, 1, 9), 1, (True, True, False), complex)])
def test_smallest_grids(N, nn, pbc, dtype):
    print('case smallest grid', N, nn, pbc, dtype, flush=True)
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    cell = tuple(0.4 * n for n in N)
    og = O.Grid(N, cell, pbc)
    gd = GridDescriptor(N, cell, pbc)
    rng = np.random.default_rng(sum(N) + nn)
    shape = (3,) + tuple(og.n_c)
    a = rng.standard_normal(shape)
    ph = None
    if dtype == complex:
        a = a + 1j * rng.standard_normal(shape)
        ph = og.phase_cd([0.2, -0.35, 0.1])
    ref = np.zeros_like(a)
    O.laplace(og, 1.0, nn, dtype=dtype).apply(a, ref, ph)
    got = dev(np.zeros_like(a))
    Laplace(gd, 1.0, nn, dtype).apply(dev(a), got, ph)
    assert np.array_equal(got.cpu().numpy(), ref)


# (the reference's transformers need (n_in + 2 nn - 1)^2 <= n_out^2 on the first two
# axes when interpolating, gpaw/transformers.py:48-49: nn=3 needs a coarse extent >= 5)
@pytest.mark.parametrize('N,nn,pbc', [((10, 10, 10), 3, True), ((6, 6, 6), 1, False),
                                      ((2, 2, 2), 1, True), ((10, 4, 6), 2, True)])
def test_smallest_transfers(N, nn, pbc):
    print('case smallest transfer', N, nn, pbc, flush=True)
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.transformers import Transformer
    cell = tuple(0.5 * n for n in N)
    og, gd = O.Grid(N, cell, pbc), GridDescriptor(N, cell, pbc)
    oc, gc = og.coarsen(), gd.coarsen()
    rng = np.random.default_rng(5)
    fine = rng.standard_normal((2,) + tuple(og.n_c))
    coarse = rng.standard_normal((2,) + tuple(oc.n_c))
    r_ref = np.zeros((2,) + tuple(oc.n_c))
    O.Transformer(og, oc, nn).apply(fine, r_ref)
    r = dev(np.zeros_like(r_ref))
    Transformer(gd, gc, nn).apply(dev(fine), r)
    assert np.array_equal(r.cpu().numpy(), r_ref)
    i_ref = np.zeros((2,) + tuple(og.n_c))
    O.Transformer(oc, og, nn).apply(coarse, i_ref)
    i = dev(np.zeros_like(i_ref))
    Transformer(gc, gd, nn).apply(dev(coarse), i)
    assert np.array_equal(i.cpu().numpy(), i_ref)


def test_largest_band_count_and_beyond():
    """65535 bands per call (the C ABI's limit) on a 4^3 grid, checked on a
    few bands against the oracle; 65536 bands are refused."""
    print('case 65535 bands', flush=True)
    t = _torch()
    from gpaw_amd._lib import GpawAmdError
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    N = (4, 4, 4)
    og, gd = O.Grid(N, (2.0,) * 3, True), GridDescriptor(N, (2.0,) * 3, True)
    lap = Laplace(gd, 1.0, 1)
    g = t.Generator(device='cuda').manual_seed(3)
    a = t.randn((65535,) + N, dtype=t.float64, device='cuda', generator=g)
    b = t.empty_like(a)
    lap.apply(a, b)
    idx = [0, 1, 32767, 65534]
    ref = np.zeros((len(idx),) + N)
    O.laplace(og, 1.0, 1).apply(a[idx].cpu().numpy(), ref)
    assert np.array_equal(b[idx].cpu().numpy(), ref)
    del a, b
    big = t.zeros((65536,) + N, dtype=t.float64, device='cuda')
    with pytest.raises(GpawAmdError):
        lap.apply(big, t.empty_like(big))


@pytest.mark.parametrize('N,nn,pbc,nb', [((45, 48, 40), 3, True, 64), ((33, 34, 35), 2, False, 40),
                                        ((64, 64, 64), 3, True, 32)])
def test_wide_complex_tiles(N, nn, pbc, nb):
    """Complex bands on grids large enough for the 1024-thread tiles (two
    planes per step; an odd plane count leaves a one-plane last step), plain
    and with the local potential, against the oracle."""
    print('case wide complex tiles', N, nn, pbc, nb, flush=True)
    t = _torch()
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    cell = tuple(0.35 * n for n in N)
    og, gd = O.Grid(N, cell, pbc), GridDescriptor(N, cell, pbc)
    rng = np.random.default_rng(nb)
    shape = (nb,) + tuple(og.n_c)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    v = rng.standard_normal(tuple(og.n_c))
    ph = og.phase_cd([0.3, -0.1, 0.2])
    ref = np.zeros_like(a)
    O.laplace(og, -0.5, nn, dtype=complex).apply(a, ref, ph)
    lap = Laplace(gd, -0.5, nn, complex)
    ad = dev(a)
    got = t.empty_like(ad)
    lap.apply(ad, got, ph)
    assert np.array_equal(got.cpu().numpy(), ref)
    ref += a * v
    lap.operator.apply_potential(ad, got, dev(v), ph)
    assert np.array_equal(got.cpu().numpy(), ref)


def test_c1_configuration():
    """C1: Laplace(nn=3), float64, 8 bands of 48^3, periodic, h = 0.2 A,
    inputs U[-1, 1]: bit-identical to the oracle."""
    print('case C1', flush=True)
    t = _torch()
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    h = 0.2 / 0.52917720859
    N = (48,) * 3
    og, gd = O.Grid(N, (48 * h,) * 3, True), GridDescriptor(N, (48 * h,) * 3, True)
    a = np.random.default_rng(1).uniform(-1, 1, (8,) + N)
    ref = np.zeros_like(a)
    O.laplace(og, 1.0, 3).apply(a, ref)
    got = t.empty((8,) + N, dtype=t.float64, device='cuda')
    Laplace(gd, 1.0, 3).apply(dev(a), got)
    assert np.array_equal(got.cpu().numpy(), ref)
