"""Mixed-precision multigrid Poisson (FP32 V-cycles inside an FP64 outer
loop): the stopping criterion is met when checked independently in FP64,
and the potential agrees with the plain FP64 solve.

Tolerances: the relative residual target is 1e-10; two solutions of the same
linear system that both meet it may differ by up to cond(L) * 1e-10 in
principle, in practice ~1e-10 relative for these smooth densities.  The test
asserts 1e-8 relative max-norm against the FP64 multigrid solution driven to
the same relative residual.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
BOHR = 0.52917720859


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _setup(N, h, pbc, count, seed):
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.utilities import gaussian_density, random_gaussian_charges
    gd = GridDescriptor((N,) * 3, (N * h,) * 3, pbc)
    rng = np.random.default_rng(seed)
    c, w, q = random_gaussian_charges(gd, rng, count, 0.4 / BOHR, 1.2 / BOHR,
                                      2.0 / BOHR)
    rho = gaussian_density(gd, c, w, q)
    return gd, rho


def _relres(gd, phi, rho):
    """||L phi - rho_n|| / ||rho_n|| with the oracle's FP64 operator."""
    og = O.Grid(gd.N_c, gd.cell_cv, gd.pbc_c)
    lap = O.gpaw_oracle.laplace(og, -0.25 / np.pi, 3)
    rn = rho - rho.sum() / np.prod(gd.get_size_of_global_array())
    out = np.zeros_like(phi)
    lap.apply(np.ascontiguousarray(phi), out)
    if gd.pbc_c.all():
        rn = rn - rn.mean()
    return np.linalg.norm(out - rn) / np.linalg.norm(rn)


@pytest.mark.parametrize('inner_precision', ['fp32', 'fp64'])
@pytest.mark.parametrize('outer', ['pcg', 'refine'])
@pytest.mark.parametrize('N,pbc', [(48, False), (64, False), (32, True)])
def test_mixed_meets_tolerance_and_matches_fp64(N, pbc, outer, inner_precision):
    print('case mixed', N, pbc, outer, inner_precision, flush=True)
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver, PoissonSolver
    gd, rho = _setup(N, 0.18 / BOHR, pbc, 6, N)
    ms = MixedPoissonSolver(nn=3, rel_tol=1e-10, outer=outer, inner=4,
                            inner_precision=inner_precision)
    ms.set_grid_descriptor(gd)
    ms.initialize()
    phi = t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda')
    it = ms.solve(phi, t.from_numpy(rho).to('cuda'))
    got = phi.cpu().numpy()
    rr = _relres(gd, got, rho)
    print('iterations', it, 'relres', rr, 'reported', ms.relres, flush=True)
    assert ms.relres <= 1e-10
    assert rr <= 1.5e-10
    # the reported ratio is the true residual's, not the CG recurrence's
    assert abs(rr - ms.relres) <= 1e-3 * rr
    # FP64 multigrid to the same relative residual
    ps = PoissonSolver(nn=3)
    ps.set_grid_descriptor(gd)
    ps.initialize()
    rn = rho - rho.sum() / np.prod(gd.get_size_of_global_array())
    eps = gd.dv * (1e-10 * np.linalg.norm(rn)) ** 2
    ref = t.zeros_like(phi)
    ps.solve(ref, t.from_numpy(rho).to('cuda'), eps=eps)
    ref = ref.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print('rel max-norm error vs FP64', err, flush=True)
    assert err < 1e-8


def test_mixed_fewer_cycles_than_fp64():
    """The PCG outer loop needs fewer V-cycles than plain FP64 V-cycles."""
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver, PoissonSolver
    gd, rho = _setup(64, 0.18 / BOHR, False, 8, 5)
    ms = MixedPoissonSolver(nn=3, rel_tol=1e-10)
    ms.set_grid_descriptor(gd)
    ms.initialize()
    it = ms.solve(t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda'),
                  t.from_numpy(rho).to('cuda'))
    ps = PoissonSolver(nn=3)
    ps.set_grid_descriptor(gd)
    ps.initialize()
    rn = rho - rho.sum() / np.prod(gd.get_size_of_global_array())
    eps = gd.dv * (1e-10 * np.linalg.norm(rn)) ** 2
    n64 = ps.solve(t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda'),
                   t.from_numpy(rho).to('cuda'), eps=eps)
    print('pcg iterations', it, 'fp64 V-cycles', n64, flush=True)
    assert it < n64


@pytest.mark.parametrize('N,pbc', [(64, False), (40, True)])
def test_mixed_fused_sweeps_identical(N, pbc, monkeypatch):
    """The FP32 V-cycle's levels below 0 run four Jacobi sweeps per fused
    launch, the first post-smoothing launch with the coarse-grid correction;
    the mixed solution must be bit-identical to single sweeps
    (GPAW_AMD_FUSED_JACOBI=0) and to a separate correction launch
    (GPAW_AMD_FOLD_CORRECTION=0), both read when the plan is built."""
    print('case mixed fused sweeps', N, pbc, flush=True)
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver
    out = []
    for fused, fold in (('1', '1'), ('1', '0'), ('0', '1')):
        monkeypatch.setenv('GPAW_AMD_FUSED_JACOBI', fused)
        monkeypatch.setenv('GPAW_AMD_FOLD_CORRECTION', fold)
        gd, rho = _setup(N, 0.18 / BOHR, pbc, 6, N + 3)
        ms = MixedPoissonSolver(nn=3, rel_tol=1e-10, outer='pcg', inner=4)
        ms.set_grid_descriptor(gd)
        ms.initialize()
        phi = t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda')
        n = ms.solve(phi, t.from_numpy(rho).to('cuda'))
        out.append((n, phi))
    for n, phi in out[1:]:
        assert n == out[0][0]
        assert t.equal(phi, out[0][1])


def test_inner_precision_iteration_counts():
    """The same outer loop with the inner V-cycle in FP32 and in FP64: the
    iteration counts differ by at most two (FP32 rounding in the
    preconditioner barely slows the Krylov iteration), so the speed of the
    FP32 variant over the FP64 one is the precision's own contribution."""
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver
    gd, rho = _setup(64, 0.18 / BOHR, False, 8, 9)
    its = {}
    for prec in ('fp32', 'fp64'):
        ms = MixedPoissonSolver(nn=3, rel_tol=1e-10, inner_precision=prec)
        ms.set_grid_descriptor(gd)
        ms.initialize()
        its[prec] = ms.solve(t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda'),
                             t.from_numpy(rho).to('cuda'))
    print('pcg iterations', its, flush=True)
    assert abs(its['fp32'] - its['fp64']) <= 2


def _weighted_level(ops_, level, wpre, wpost):
    """The oracle's V-cycle with the mixed solver's sweeps on one level
    (not the coarsest): each relax call of that level (pre, then post, per
    cycle) runs one Jacobi sweep per weight of its list."""
    ops_.presmooths[level] = len(wpre)
    ops_.postsmooths[level] = len(wpost)
    orig = ops_.operators[level].relax
    calls = [0]

    def relax(method, f, s, n, w=1.0):
        ws = wpre if calls[0] % 2 == 0 else wpost
        calls[0] += 1
        assert n == len(ws)
        for wi in ws:
            orig(method, f, s, 1, wi)
    ops_.operators[level].relax = relax


@pytest.mark.parametrize('N,pbc', [(64, False), (40, True), (160, False)])
def test_preconditioner_vcycle(N, pbc):
    """gpaw_poisson_precondition: the FP64 inner V-cycle from zero equals the
    oracle's V-cycle (gpaw/poisson.py:245-270) with the mixed solver's
    level-0 and level-1 sweeps and weights (gpaw_poisson_mixed_weights), bit
    for bit; the FP32 one agrees with it to the FP32 rounding level (1e-5
    relative max-norm).  On 159^3 level 1 (79^3) runs Chebyshev-weighted
    2 + 2 sweeps, on the smaller grids the reference's 4 + 4."""
    print('case precondition', N, pbc, flush=True)
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver
    gd, rho = _setup(N, 0.2 / BOHR, pbc, 6, N + 7)
    r = rho - rho.mean()
    ms = MixedPoissonSolver(nn=3)
    ms.set_grid_descriptor(gd)
    ms.initialize()
    wpre, wpost = ms.smoothing_weights()
    w1pre, w1post = ms.smoothing_weights(1)
    print('  level 1 weights', w1pre, w1post, flush=True)
    if N == 160:
        assert len(w1pre) == 2 and w1post == w1pre[::-1] and w1pre[0] != w1pre[1]
    else:
        assert len(w1pre) == len(w1post) == 4 and len(set(w1pre + w1post)) == 1
    z64 = t.full(tuple(gd.n_c), np.nan, dtype=t.float64, device='cuda')
    z32 = t.full(tuple(gd.n_c), np.nan, dtype=t.float64, device='cuda')
    rd = t.from_numpy(r).to('cuda')
    ms.precondition(rd, z64, 'fp64')
    ms.precondition(rd, z32, 'fp32')
    og = O.Grid(gd.N_c, gd.cell_cv, gd.pbc_c)
    ops_ = O.PoissonSolver(nn=3)
    ops_.set_grid_descriptor(og)
    ops_.initialize()
    _weighted_level(ops_, 0, wpre, wpost)
    if N == 160:
        _weighted_level(ops_, 1, w1pre, w1post)
    else:
        assert w1pre[0] == ops_.weights[1]
    want = og.zeros()
    ops_.run_vcycles(want, r, 1)
    z64 = z64.cpu().numpy()
    assert np.array_equal(z64, want)
    err = np.abs(z32.cpu().numpy() - want).max() / np.abs(want).max()
    print('  FP32 V-cycle vs FP64: %.2e' % err, flush=True)
    assert err < 1e-5


def _lmax_dense(h_c, nn):
    """Largest eigenvalue of D^-1 A for the nn Laplacian of spacings h_c on
    an 8^3 periodic grid (the symbol's maximum, reached at theta = pi on
    every axis of an even grid), from the dense matrix the oracle's apply
    builds on unit vectors."""
    n = 8
    og = O.Grid((n,) * 3, np.array(h_c) * n, True)
    lap = O.gpaw_oracle.laplace(og, 1.0, nn)
    A = np.empty((n ** 3, n ** 3))
    e = np.zeros((n,) * 3)
    out = np.zeros((n,) * 3)
    for i in range(n ** 3):
        e.flat[i] = 1.0
        lap.apply(e, out)
        A[:, i] = out.ravel()
        e.flat[i] = 0.0
    d = A[0, 0]
    return np.linalg.eigvalsh(A / d).max()


@pytest.mark.parametrize('h_c', [(0.2, 0.2, 0.2), (0.2, 0.24, 0.22)])
def test_chebyshev_smoothing_weights(h_c):
    """The mixed V-cycle's level-0 weights: 2 + 2 sweeps, post in reverse
    order, and prod_i (1 - w_i x) equals the degree-2 Chebyshev polynomial
    T_2((b + a - 2x) / (b - a)) / T_2((b + a) / (b - a)) on [0, b] with
    b = the largest eigenvalue of D^-1 A (dense eigensolve) and a = b / 7.
    Roots placed at cos(i pi / k) instead, or b taken as the 7-point
    stencil's 2, each fail the comparison."""
    _torch()
    from numpy.polynomial import chebyshev as C
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import MixedPoissonSolver
    N = 48
    gd = GridDescriptor((N,) * 3, np.array(h_c) * N / BOHR, False)
    ms = MixedPoissonSolver(nn=3)
    ms.set_grid_descriptor(gd)
    ms.initialize()
    wpre, wpost = ms.smoothing_weights()
    print('weights', wpre, wpost, flush=True)
    assert len(wpre) == 2 and wpost == wpre[::-1]
    b = _lmax_dense(np.array(h_c) / BOHR, 3)
    a = b / 7.0
    x = np.linspace(0.0, b, 101)

    def cheb(x, a, b):
        return C.chebval((b + a - 2 * x) / (b - a), [0, 0, 1]) / \
            C.chebval((b + a) / (b - a), [0, 0, 1])

    def prod(ws):
        return np.prod([1 - w * x for w in ws], axis=0)
    assert np.abs(prod(wpre) - cheb(x, a, b)).max() < 1e-12
    # planted mistakes
    mid, half = 0.5 * (b + a), 0.5 * (b - a)
    wrong_roots = [1 / (mid + half * np.cos(i * np.pi / 2)) for i in (1, 2)]
    assert np.abs(prod(wrong_roots) - cheb(x, a, b)).max() > 1e-3
    b7 = 2.0
    wrong_lmax = [1 / (0.5 * (b7 + b7 / 7) + 0.5 * (b7 - b7 / 7) * np.cos((2 * i + 1) * np.pi / 4))
                  for i in range(2)]
    assert np.abs(prod(wrong_lmax) - cheb(x, a, b)).max() > 1e-3


@pytest.mark.parametrize('N,pbc', [(64, False), (48, True), (96, False)])
def test_vcycle_fusions_identical(N, pbc, monkeypatch):
    """The residual and its restriction in one launch or in two
    (GPAW_AMD_FUSED_RESTRICT=0), and the coarse-grid corrections applied by
    the first post-smoothing sweeps or by launches of their own
    (GPAW_AMD_FOLD_CORRECTION=0), both read when the plan is built: the FP32
    and FP64 inner V-cycles give the same bits every way."""
    print('case V-cycle fusions', N, pbc, flush=True)
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver
    gd, rho = _setup(N, 0.2 / BOHR, pbc, 5, N + 3)
    rd = t.from_numpy(rho - rho.mean()).to('cuda')
    out = []
    for frest, fold in (('1', '1'), ('0', '1'), ('1', '0')):
        monkeypatch.setenv('GPAW_AMD_FUSED_RESTRICT', frest)
        monkeypatch.setenv('GPAW_AMD_FOLD_CORRECTION', fold)
        ms = MixedPoissonSolver(nn=3)
        ms.set_grid_descriptor(gd)
        ms.initialize()
        z = [t.full(tuple(gd.n_c), np.nan, dtype=t.float64, device='cuda') for _ in range(2)]
        ms.precondition(rd, z[0], 'fp32')
        ms.precondition(rd, z[1], 'fp64')
        out.append(z)
    for z in out[1:]:
        assert t.equal(out[0][0], z[0])
        assert t.equal(out[0][1], z[1])


@pytest.mark.parametrize('N,pbc', [(64, False), (40, True)])
def test_solve_history(N, pbc):
    """reuse_history=True: the first chunk of a solve is as long as the last
    converged solve's CG steps up to the first step whose |d| met the
    tolerance, so a repeated solve takes no more steps than the first and
    still meets 1e-10 (checked with the oracle's FP64 operator).
    reuse_history=False: repeated solves take the same steps and give the
    same bits (results depend on the inputs only)."""
    print('case solve history', N, pbc, flush=True)
    t = _torch()
    from gpaw_amd.poisson import MixedPoissonSolver
    gd, rho = _setup(N, 0.18 / BOHR, pbc, 6, N + 11)
    rd = t.from_numpy(rho).to('cuda')
    for reuse in (True, False):
        ms = MixedPoissonSolver(nn=3, rel_tol=1e-10, reuse_history=reuse)
        ms.set_grid_descriptor(gd)
        ms.initialize()
        runs = []
        for _ in range(3):
            phi = t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda')
            n = ms.solve(phi, rd)
            rr = _relres(gd, phi.cpu().numpy(), rho)
            print('  reuse=%s steps %d relres %.2e' % (reuse, n, rr), flush=True)
            assert rr <= 1.5e-10 and ms.relres <= 1e-10
            runs.append((n, phi))
        if reuse:
            assert runs[1][0] <= runs[0][0] and runs[2][0] == runs[1][0]
            assert t.equal(runs[1][1], runs[2][1])
        else:
            for n, phi in runs[1:]:
                assert n == runs[0][0] and t.equal(phi, runs[0][1])


def test_mixed_degenerate_inputs():
    """Error and edge paths of the CG loop: a zero density converges at
    once with phi left at zero (the guarded step length is 0, not 0/0); a
    non-finite density reports a breakdown (GPAW_EDEVICE) and leaves phi
    untouched; too few iterations raise PoissonConvergenceError."""
    print('case mixed degenerate inputs', flush=True)
    t = _torch()
    from gpaw_amd._lib import GpawAmdError
    from gpaw_amd.poisson import MixedPoissonSolver, PoissonConvergenceError
    gd, rho = _setup(32, 0.2 / BOHR, False, 4, 3)
    n = tuple(gd.n_c)
    ms = MixedPoissonSolver(nn=3, rel_tol=1e-10)
    ms.set_grid_descriptor(gd)
    ms.initialize()
    phi = t.zeros(n, dtype=t.float64, device='cuda')
    ms.solve_neutral(phi, t.zeros(n, dtype=t.float64, device='cuda'))
    assert t.count_nonzero(phi).item() == 0 and ms.relres == 0.0
    bad = t.from_numpy(rho - rho.mean()).to('cuda')
    bad[3, 4, 5] = float('nan')
    phi = t.full(n, 7.0, dtype=t.float64, device='cuda')
    with pytest.raises(GpawAmdError):
        ms.solve_neutral(phi, bad)
    assert bool((phi == 7.0).all())
    few = MixedPoissonSolver(nn=3, rel_tol=1e-10, maxiter=2)
    few.set_grid_descriptor(gd)
    few.initialize()
    with pytest.raises(PoissonConvergenceError):
        few.solve_neutral(t.zeros(n, dtype=t.float64, device='cuda'),
                          t.from_numpy(rho - rho.mean()).to('cuda'))


@pytest.mark.parametrize('N,pbc', [(48, False), (32, True)])
def test_mixed_device_solve_equals_python_branches(N, pbc):
    """MixedPoissonSolver.solve: the charge integral and the mean
    subtraction on the device (gpaw_poisson_neutralize), then the mixed
    solve of the density left in the plan, equal the reference's branch
    steps run from Python (gd.integrate, the background subtraction,
    solve_neutral) bit for bit, with the same step count."""
    print('case mixed device solve', N, pbc, flush=True)
    t = _torch()
    from gpaw_amd import ops
    from gpaw_amd.poisson import MixedPoissonSolver
    gd, rho = _setup(N, 0.2 / BOHR, pbc, 6, N + 5)
    npts = np.prod(gd.get_size_of_global_array())
    rho = rho - rho.sum() / npts + 4e-7 / gd.dv / npts
    out = []
    for device in (True, False):
        ms = MixedPoissonSolver(nn=3, rel_tol=1e-10, reuse_history=False)
        ms.set_grid_descriptor(gd)
        ms.initialize()
        phi = t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda')
        rd = t.from_numpy(rho).to('cuda')
        if device:
            n = ms.solve(phi, rd)
        else:
            q = gd.integrate(rd)
            assert abs(q) <= 1e-6
            rn = t.empty_like(rd)
            ops.sub_scalar(rn, rd, q / gd.dv / npts)
            n = ms.solve_neutral(phi, rn)
        out.append((n, phi))
    assert out[0][0] == out[1][0]
    assert t.equal(out[0][1], out[1][1])
