"""gpaw_dots_noalloc (partials in a static device array, no allocation)
against gpaw_dots (partials in a buffer it allocates): the same kernels, so
the same bits, on two calls in a row, the second with a smaller input, and
a clear error for a call beyond the static array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _call(name, a, b, nb, n, scale, out, stream):
    from gpaw_amd import _lib
    return getattr(_lib.lib(), name)(ctypes.c_void_p(a.data_ptr()),
                                     ctypes.c_void_p(b.data_ptr()), nb, n, scale,
                                     ctypes.c_void_p(out.data_ptr()), stream)


@pytest.mark.parametrize('order', ['noalloc_first', 'dots_first'])
def test_dots_noalloc_matches_dots(order):
    t = _torch()
    from gpaw_amd import _arrays, _lib
    rng = np.random.default_rng(11)
    stream = _arrays.stream()
    # two calls in a row, the second smaller: a stale partial of the first
    # call left in the shared array would change the second result
    for nb, shape in [(12, (63, 63, 63)), (5, (31, 29, 33))]:
        print('case dots_noalloc', order, nb, shape, flush=True)
        n = int(np.prod(shape))
        a = t.from_numpy(rng.normal(size=n)).to('cuda')
        b = t.from_numpy(rng.normal(size=(nb, n))).to('cuda')
        want = t.full((nb,), np.nan, dtype=t.float64, device='cuda')
        got = t.full((nb,), np.nan, dtype=t.float64, device='cuda')
        calls = [('gpaw_dots', want), ('gpaw_dots_noalloc', got)]
        if order == 'noalloc_first':
            calls.reverse()
        for name, out in calls:
            rc = _call(name, a, b, nb, n, 0.37, out, stream)
            assert rc == _lib.GPAW_OK, _lib.lib().gpaw_last_error()
        t.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
        ref = 0.37 * (b.cpu().numpy() @ a.cpu().numpy())
        assert np.allclose(got.cpu().numpy(), ref, rtol=1e-12, atol=0)


def test_dots_noalloc_rejects_beyond_its_array():
    t = _torch()
    from gpaw_amd import _arrays, _lib
    n = 256 * 1024                      # 256 chunks per row
    nb = (1 << 19) // 256 + 1           # one row more than the array holds
    a = t.zeros(n, dtype=t.float64, device='cuda')
    b = t.zeros(1, dtype=t.float64, device='cuda')   # never read: refused first
    out = t.zeros(nb, dtype=t.float64, device='cuda')
    rc = _call('gpaw_dots_noalloc', a, b, nb, n, 1.0, out, _arrays.stream())
    assert rc == _lib.GPAW_EINVAL
    assert b'static' in _lib.lib().gpaw_last_error()
