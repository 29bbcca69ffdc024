"""FP32 variants of the C3 operators (float32 / complex64 storage and
arithmetic) checked against the FP64 oracle on the same inputs.

The FP64 path is bit-exact; FP32 is not a parity path but a stated
approximation: each result's relative max-norm error against the oracle,
max|out32 - out64| / max|out64|, must stay below TOL = 2e-6 (float32
epsilon is 6e-8; a 19-term stencil or a 3-pass K=6 transfer on
unit-variance data accumulates a few epsilons).  The measured errors are
printed."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _relerr(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize('dtype,k', [(complex, (0.1, 0.25, -0.3)), (float, None)])
def test_laplace_fp32_against_oracle(dtype, k):
    torch = _torch()
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    N = (32, 32, 40)
    og = O.Grid(N, np.array(N) * 0.378)
    gd = GridDescriptor(N, np.array(N) * 0.378)
    rng = np.random.default_rng(5)
    shape = (8,) + N
    a = rng.standard_normal(shape)
    if dtype == complex:
        a = a + 1j * rng.standard_normal(shape)
    ph = og.phase_cd(k) if k is not None else None
    ref = np.empty_like(a)
    O.laplace(og, -0.5, 3, dtype).apply(a, ref, ph)
    t32 = torch.complex64 if dtype == complex else torch.float32
    a32 = torch.from_numpy(a).to(t32).cuda()
    b32 = torch.empty_like(a32)
    Laplace(gd, -0.5, 3, dtype).apply(a32, b32, ph)
    err = _relerr(b32.cpu().numpy().astype(a.dtype), ref)
    print('laplace fp32', dtype.__name__, 'relative max-norm error %.2e' % err, flush=True)
    assert err < TOL


@pytest.mark.parametrize('interp', [False, True])
def test_k6_transfers_fp32_against_oracle(interp):
    torch = _torch()
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.transformers import Transformer
    N = (32, 32, 32)
    og, gd = O.Grid(N, (10.0,) * 3), GridDescriptor(N, (10.0,) * 3)
    ogc, gdc = og.coarsen(), gd.coarsen()
    rng = np.random.default_rng(9)
    src_grid = ogc if interp else og
    shape = (4,) + tuple(src_grid.n_c)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dst_shape = (4,) + tuple((og if interp else ogc).n_c)
    ref = np.empty(dst_shape, complex)
    k = (0.2, -0.1, 0.35)
    ph = src_grid.phase_cd(k)
    if interp:
        O.Transformer(ogc, og, 3, complex).apply(a, ref, ph)
        tr = Transformer(gdc, gd, 3, complex)
    else:
        O.Transformer(og, ogc, 3, complex).apply(a, ref, ph)
        tr = Transformer(gd, gdc, 3, complex)
    a32 = torch.from_numpy(a).to(torch.complex64).cuda()
    b32 = torch.empty(dst_shape, dtype=torch.complex64, device='cuda')
    tr.apply(a32, b32, ph)
    err = _relerr(b32.cpu().numpy().astype(complex), ref)
    print('K=6', 'interpolate' if interp else 'restrict', 'fp32 error %.2e' % err, flush=True)
    assert err < TOL
