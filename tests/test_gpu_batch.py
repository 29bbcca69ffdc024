"""Batched apply (one persistent launch over many grids): every group's
result is bit-identical to its own FDOperator.apply, for bands of several
k-points, grids of uneven size and boundary, different grid spacings (so
different coefficients), stencil ranges 1..3 in one call, and groups that
cannot be tiled (applied by the single-operator kernels)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _groups(torch, dtype, rng):
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    specs = [  # N, pbc, h, nn, nbands, k
        ((48, 48, 48), True, 0.20, 3, 3, (0.1, 0.2, -0.3)),
        ((48, 48, 48), True, 0.20, 3, 2, (0.5, 0.0, 0.25)),
        ((48, 48, 48), False, 0.22, 3, 1, None),
        ((80, 64, 96), True, 0.18, 3, 2, (-0.2, 0.4, 0.1)),
        ((40, 24, 56), (True, False, True), 0.25, 2, 4, (0.3, 0.0, -0.1)),
        ((64, 64, 64), True, 0.20, 1, 5, (0.0, 0.25, 0.25)),
        ((6, 6, 6), True, 0.30, 3, 2, (0.2, 0.2, 0.2)),      # too small to tile
        ((32, 32, 32), False, 0.20, 3, 0, None),             # no bands
    ]
    out = []
    tdt = torch.complex128 if dtype == complex else torch.float64
    for N, pbc, h, nn, nb, k in specs:
        gd = GridDescriptor(N, np.array(N) * h, pbc)
        op = Laplace(gd, -0.5, nn, dtype)
        shape = (nb,) + tuple(gd.n_c)
        a = rng.standard_normal(shape)
        if dtype == complex:
            a = a + 1j * rng.standard_normal(shape)
        ph = gd.phase_cd(k) if (dtype == complex and k is not None) else None
        ta = torch.from_numpy(a).cuda()
        out.append((op, ta, torch.empty(shape, dtype=tdt, device='cuda'), ph))
    return out


@pytest.mark.parametrize('dtype', [float, complex])
def test_batch_equals_single_applies(dtype):
    print('case batch', dtype.__name__, flush=True)
    torch = _torch()
    from gpaw_amd.batch import BatchApply
    rng = np.random.default_rng(17)
    groups = _groups(torch, dtype, rng)
    batch = BatchApply(max_groups=16)
    for rep in range(3):      # the descriptor slots are reused
        for _, _, b, _ in groups:
            b.fill_(float('nan'))
        batch(groups)
        torch.cuda.synchronize()
        for i, (op, a, b, ph) in enumerate(groups):
            ref = torch.empty_like(a)
            op.apply(a, ref, ph)
            torch.cuda.synchronize()
            assert torch.equal(b, ref), (rep, i)


def test_batch_single_band_3d_and_errors():
    torch = _torch()
    from gpaw_amd.batch import BatchApply
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    gd = GridDescriptor((32, 48, 64), (6.0, 9.0, 12.0))
    op = Laplace(gd, 1.0, 3)
    a = torch.randn(tuple(gd.n_c), dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    BatchApply(4)([(op, a, b, None)])
    ref = torch.empty_like(a)
    op.apply(a, ref)
    torch.cuda.synchronize()
    assert torch.equal(b, ref)
    with pytest.raises(ValueError):
        BatchApply(1)([(op, a, b, None), (op, a, b, None)])
    with pytest.raises(TypeError):    # mixed dtypes
        BatchApply(4)([(op, a, b, None), (op, a.float(), b.float(), None)])


def test_autotuned_shapes_give_identical_results():
    """The launch-shape tuner times its candidates on the real launch; every
    shape computes the same bits, so tuned and heuristic results agree."""
    torch = _torch()
    from gpaw_amd import ops
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    gd = GridDescriptor((64, 48, 40), (12.0, 9.0, 7.5))
    ph = gd.phase_cd((0.1, -0.2, 0.3))
    lap = Laplace(gd, -0.5, 3, complex)
    a = torch.randn((16,) + tuple(gd.n_c), dtype=torch.complex128, device='cuda')
    ref = torch.empty_like(a)
    lap.apply(a, ref, ph)
    prev = ops.autotune(True)
    try:
        for _ in range(2):       # first call tunes, second reuses the entry
            b = torch.empty_like(a)
            lap.apply(a, b, ph)
            torch.cuda.synchronize()
            assert torch.equal(b, ref)
        rep = ops.autotune_report()
        assert 'n=64x48x40' in rep and 'nin=16' in rep, rep
        print(rep, flush=True)
    finally:
        ops.autotune(prev)


@pytest.mark.parametrize('dtype', [float, complex])
def test_batch_equals_oracle(dtype):
    """A batch of uneven grids, k-points, boundaries, spacings and stencil
    ranges against the oracle's restatement of the reference's apply (its
    own coefficients and phases per group): bit for bit."""
    print('case batch vs oracle', dtype.__name__, flush=True)
    torch = _torch()
    import oracle as O
    from gpaw_amd.batch import BatchApply
    rng = np.random.default_rng(29)
    groups = _groups(torch, dtype, rng)
    batch = BatchApply(max_groups=16)
    for _, _, b, _ in groups:
        b.fill_(float('nan'))
    batch(groups)
    torch.cuda.synchronize()
    for i, (op, a, b, ph) in enumerate(groups):
        gd = op.gd
        og = O.Grid(gd.N_c, gd.cell_cv, gd.pbc_c)
        assert np.array_equal(np.asarray(op.coef_p), O.laplace_stencil(og, -0.5, op.n)[0])
        ha = a.cpu().numpy()
        ref = np.zeros_like(ha)
        if ha.shape[0]:
            O.Operator(op.coef_p, op.offset_pc, og, dtype).apply(ha, ref, ph)
        assert np.array_equal(b.cpu().numpy(), ref), i


def test_batch_refuses_overlapping_outputs():
    """All groups run concurrently: an output overlapping another group's
    input or output is refused (inputs may be shared)."""
    torch = _torch()
    from gpaw_amd import GpawAmdError
    from gpaw_amd.batch import BatchApply
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    gd = GridDescriptor((32, 32, 32), (6.0, 6.0, 6.0))
    op = Laplace(gd, 1.0, 3)
    a = torch.randn((2,) + tuple(gd.n_c), dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    BatchApply(4)([(op, a, b, None), (op, a, c, None)])     # shared input: fine
    with pytest.raises(GpawAmdError):
        BatchApply(4)([(op, a, b, None), (op, b, c, None)])  # b written and read
    with pytest.raises(GpawAmdError):
        BatchApply(4)([(op, a[0], b[0], None), (op, a[1], b[0], None)])
