"""Benchmark: multigrid Poisson solves/s on one MI355X, with the FD roofline.

Workload (a "step" = one full Poisson solve from phi = 0):
  96^3 non-periodic grid (95^3 arrays), h = 0.2 A, PoissonSolver(nn=3,
  relax='J'), 4 coarsenings, FP64; density = 8 seeded normalised Gaussians
  (centres >= 2 A from the walls, widths U[0.5, 1.5] A, charges +-1, sum 0).
  Solved until dv*sum(r^2) <= 2e-10 AND ||L phi - rho||_2 / ||rho||_2 <= 1e-10
  (the reference criterion and the relative-residual target; the second is
  the stricter one here).

Also measured (same process, not the headline value):
  * roofline of the dominant kernel of the solve (level-0 Jacobi sweep),
    timed with HIP events on the launch stream;
  * FD Laplace (nn=3, complex128, 256 bands x 64^3, Bloch phases) HBM GB/s,
    and the K=6 restriction / interpolation of the same bands;
  * the eigensolver preconditioner on the same bands;
  * mixed-precision vs FP64 multigrid on 160^3 (159^3 arrays), h = 0.18 A;
  * a serving run: Poisson-process stream of mixed Poisson / Laplace
    requests, p50 / p99 latency;
  * the reference's own C kernels (oracle/_ref, built from the reference
    sources) driven by the restated V-cycle, on one host core.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
                [--only fd_laplace|preconditioner|mixed_precision|serving|copy]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BOHR = 0.52917720859          # Angstrom per bohr (ASE of the reference era)
PEAK_HBM_GBS = 8000.0         # MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--no-extra', action='store_true')
    p.add_argument('--cpu-seconds', type=float, default=12.0)
    p.add_argument('--serve-n', type=int, default=2000)
    p.add_argument('--serve-rate', type=float, default=None,
                   help='arrival rate of the latency run (default: 80%% of '
                        'the measured capacity)')
    p.add_argument('--only', choices=sorted(SECTIONS),
                   help='run one measurement section alone and print its '
                        'JSON (for profiling a single workload)')
    return p.parse_args()


def c2_inputs(np, seed=42):
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.utilities import gaussian_density
    N = 96
    h = 0.2 / BOHR
    gd = GridDescriptor((N,) * 3, (N * h,) * 3, False)
    rng = np.random.default_rng(seed)
    L = np.diag(gd.cell_cv)
    margin = 2.0 / BOHR
    centers = [rng.uniform(margin, L - margin) for _ in range(8)]
    widths = rng.uniform(0.5, 1.5, 8) / BOHR
    charges = np.array([1.0, -1.0] * 4)
    rho = gaussian_density(gd, centers, widths, charges)
    return gd, rho


def eps_for(np, gd, rho, rel=1e-10, eps_ref=2e-10):
    """Reference-criterion eps that also guarantees ||r||/||rho|| <= rel
    for the neutralised density the solver works on."""
    bg = rho.sum() * gd.dv / gd.dv / np.prod(gd.get_size_of_global_array())
    rn = np.linalg.norm((rho - bg).ravel())
    return min(eps_ref, gd.dv * (rel * rn) ** 2), rn


def event_time(torch, fn, reps, warm_s=0.03):
    """Average seconds per call of fn over `reps` calls, HIP events on the
    current stream, after at least `warm_s` seconds of untimed calls (the
    GPU clocks ramp up under load; short warm-ups read 10% slow)."""
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    for _ in range(min(200, max(2, int(warm_s / max(first, 1e-6))))):
        fn()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def jacobi_kernel_name():
    """The level-0 Jacobi kernel of the C2 solve (mode 2 = Jacobi, 95^3): the
    tuner's choice when it tuned that shape, else the library's default for
    real grids of up to 2^23 points (the box kernel, 4 planes per block)."""
    from gpaw_amd import ops
    for line in ops.autotune_report().splitlines():
        if 'mode=2' in line and 'n=95x95x95' in line and 'dtype=0' in line:
            if '-> tile' in line:
                tz = line.split('TZ=')[1].split()[0]
                q = line.split(' q=')[1].split()[0]
                return ('k_star_tile<double,3,%s,1,JACOBI> level-0 sweep '
                        '(tiles of %s planes)' % (tz, q))
            if '-> box' in line:
                q = line.split(' q=')[1].split()[0]
                return ('k_star_box<double,3,%s,JACOBI> level-0 sweep '
                        '(32 x 8 columns x %s planes per block)' % (q, q))
            return 'k_star<double,3,JACOBI> level-0 sweep'
    return ('k_star_box<double,3,4,JACOBI> level-0 sweep '
            '(32 x 8 columns x 4 planes per block, the default shape)')


def measured_traffic(label):
    """HBM bytes per launch of a kernel as measured by rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs (scaled by a copy kernel of known size) and
    recorded in profiles/hbm_traffic_counters.md; None when absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        'hbm_traffic_counters.md')
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split('|')]
            if len(cells) > 6 and cells[1].endswith('(%s)' % label):
                return round((float(cells[4]) + float(cells[5])) * 1e9)
    except (OSError, ValueError):
        pass
    return None


def jacobi_roofline(torch, np, ps, rho_d):
    """Average duration of one level-0 Jacobi sweep (the dominant kernel of
    the V-cycle): 100 relax calls of 2 sweeps captured once in a graph (as
    the V-cycles run them, back to back without host launch gaps) and the
    graph replayed 4 times, HIP events on the capture's stream."""
    op = ps.operators[0].operator
    n = int(np.prod(ps.gd.n_c))
    f = torch.zeros(tuple(ps.gd.n_c), dtype=torch.float64, device='cuda')
    op.relax(2, f, rho_d, 2, 2.0 / 3.0)        # launch shape and scratch, before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(100):
            op.relax(2, f, rho_d, 2, 2.0 / 3.0)
    t = event_time(torch, g.replay, 4) / 200
    algo = 24.0 * n          # read phi, read rho, write phi (8 B each)
    return {'bound': 'hbm', 'achieved': round(algo / t / 1e9, 1),
            'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
            'frac': round(algo / t / 1e9 / PEAK_HBM_GBS, 4),
            'traffic': measured_traffic('C2 level-0 Jacobi'), 'kernel': jacobi_kernel_name(),
            'bytes_per_launch': algo, 'avg_launch_us': round(t * 1e6, 2),
            'note': '95^3 FP64 working set (20.6 MB) is cache resident; traffic = '
                    'FETCH_SIZE + WRITE_SIZE per launch from profiles/hbm_traffic_counters.md '
                    '(L2 fills from the fabric, MALL hits included)'}


def copy_bandwidth(torch):
    """Device-to-device copy of 2 GiB (read + write 4 GiB) timed with events
    with the library's copy (k_copy16_flat: one 16-byte element per thread
    over the whole range, the fastest copy measured here): the attainable
    HBM bandwidth on this box, reported beside the peak."""
    from gpaw_amd.ops import copy
    n = 1 << 28
    a = torch.empty(n, dtype=torch.float64, device='cuda').fill_(1.0)
    b = torch.empty_like(a)
    t = event_time(torch, lambda: copy(b, a), 10)
    del a, b
    return round(2 * 8 * n / t / 1e9, 1)


def fd_laplace_c3(torch, np):
    """C3 Laplace: complex128 nn=3 on 256 bands of 64^3, periodic, phases;
    its K=6 transfers and FP32 variant; and the C1 real apply (8 x 48^3)."""
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    N, nb = 64, 256
    gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, True)
    rng = np.random.default_rng(3)
    k_c = rng.uniform(-0.5, 0.5, 3)
    ph = gd.phase_cd(k_c)
    lap = Laplace(gd, -0.5, 3, complex)
    g = torch.Generator(device='cuda').manual_seed(11)
    a = torch.randn((nb, N, N, N), dtype=torch.complex128, device='cuda',
                    generator=g)
    b = torch.empty_like(a)
    t = event_time(torch, lambda: lap.apply(a, b, ph), 10)
    algo = 2.0 * 16 * nb * N ** 3
    gbs = algo / t / 1e9
    res = {'config': 'complex128 nn=3 Laplace, 256 bands x 64^3, Bloch phases',
           'avg_apply_ms': round(t * 1e3, 3), 'GBps': round(gbs, 1),
           'pct_peak': round(100 * gbs / PEAK_HBM_GBS, 1),
           'bytes_per_apply': algo}
    # nn=3 (K=6) restriction 64^3 -> 32^3 and interpolation back, same bands
    from gpaw_amd.transformers import Transformer
    gc = gd.coarsen()
    c = torch.empty((nb, N // 2, N // 2, N // 2), dtype=torch.complex128,
                    device='cuda')
    rst = Transformer(gd, gc, 3, complex)
    itp = Transformer(gc, gd, 3, complex)
    tr = event_time(torch, lambda: rst.apply(a, c, ph), 5)
    ti = event_time(torch, lambda: itp.apply(c, b, gc.phase_cd(k_c)), 5)
    tb = 16.0 * nb * (N ** 3 + (N // 2) ** 3)
    res['restrict_K6_ms'] = round(tr * 1e3, 3)
    res['restrict_K6_GBps'] = round(tb / tr / 1e9, 1)
    res['interpolate_K6_ms'] = round(ti * 1e3, 3)
    res['interpolate_K6_GBps'] = round(tb / ti / 1e9, 1)
    res['transfer_bytes'] = tb
    # FP32 variant: complex64 storage and arithmetic on the same bands,
    # error against the FP64 result (which is bit-identical to the reference)
    a32 = a.to(torch.complex64)
    b32 = torch.empty_like(a32)
    t32 = event_time(torch, lambda: lap.apply(a32, b32, ph), 10)
    lap.apply(a, b, ph)
    err = float((b32.to(torch.complex128) - b).abs().max() / b.abs().max())
    res['fp32'] = {'avg_apply_ms': round(t32 * 1e3, 3),
                   'GBps': round(algo / 2 / t32 / 1e9, 1),
                   'pct_peak': round(100 * algo / 2 / t32 / 1e9 / PEAK_HBM_GBS, 1),
                   'rel_maxnorm_err_vs_fp64': err}
    del a, b, c, a32, b32
    # C1: real float64 nn=3 Laplace on 8 bands of 48^3, periodic (14.2 MB
    # per apply: cache resident, a latency / cache number)
    N1, nb1 = 48, 8
    gd1 = GridDescriptor((N1,) * 3, (N1 * 0.2 / BOHR,) * 3, True)
    lap1 = Laplace(gd1, 1.0, 3)
    a1 = torch.rand((nb1, N1, N1, N1), dtype=torch.float64, device='cuda', generator=g)
    b1 = torch.empty_like(a1)
    t1 = event_time(torch, lambda: lap1.apply(a1, b1), 50)
    algo1 = 16.0 * nb1 * N1 ** 3
    res['c1_real_8x48'] = {'avg_apply_us': round(t1 * 1e6, 2),
                           'GBps': round(algo1 / t1 / 1e9, 1),
                           'bytes_per_apply': algo1}
    del a1, b1
    res['batch'] = batch_c3(torch, np)
    return res


def batch_c3(torch, np):
    """Several k-points and uneven grids in one batched apply (one
    persistent launch per stencil range) against the same groups applied
    one after another: complex128 nn=3 Laplace, 4 seeded k-points x 48
    bands of 64^3 periodic, plus 32 bands of 48^3 and 24 bands of
    80 x 64 x 72 (periodic, own spacing and k-point), 16 bands of 56^3 zero
    boundary (Gamma).  Both bit-identical; algorithmic bytes 32 B per
    point and band."""
    from gpaw_amd.batch import BatchApply
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    rng = np.random.default_rng(41)
    g = torch.Generator(device='cuda').manual_seed(43)
    specs = [((64, 64, 64), True, 0.2, 48)] * 4 + [
        ((48, 48, 48), True, 0.22, 32), ((80, 64, 72), True, 0.18, 24),
        ((56, 56, 56), False, 0.2, 16)]
    groups, algo = [], 0.0
    for N, pbc, h, nb in specs:
        gd = GridDescriptor(N, np.array(N) * h / BOHR, pbc)
        op = Laplace(gd, -0.5, 3, complex)
        a = torch.randn((nb,) + tuple(gd.n_c), dtype=torch.complex128,
                        device='cuda', generator=g)
        ph = gd.phase_cd(rng.uniform(-0.5, 0.5, 3)) if pbc else None
        groups.append((op, a, torch.empty_like(a), ph))
        algo += 32.0 * nb * np.prod(gd.n_c)
    batch = BatchApply(max_groups=len(groups))
    tb = event_time(torch, lambda: batch(groups), 10)
    outs = [b.clone() for _, _, b, _ in groups]

    def sequential():
        for op, a, b, ph in groups:
            op.apply(a, b, ph)
    ts = event_time(torch, sequential, 10)
    same = all(torch.equal(o, b) for o, (_, _, b, _) in zip(outs, groups))
    del groups, outs
    out = {'config': 'complex128 nn=3 Laplace: 4 k-points x 48 bands of 64^3, '
                     '32 x 48^3, 24 x 80x64x72 (periodic, Bloch phases), '
                     '16 x 56^3 zero boundary; 7 groups',
           'batched_ms': round(tb * 1e3, 3), 'batched_GBps': round(algo / tb / 1e9, 1),
           'batched_pct_peak': round(100 * algo / tb / 1e9 / PEAK_HBM_GBS, 1),
           'sequential_ms': round(ts * 1e3, 3),
           'sequential_GBps': round(algo / ts / 1e9, 1),
           'speedup_batched_over_sequential': round(ts / tb, 3),
           'bit_identical': bool(same), 'bytes': algo}
    # many small uneven groups (k-points of small cells, serving-sized
    # requests): there a launch per group costs more than its data
    groups, algo = [], 0.0
    for i in range(96):
        e = int(rng.choice([24, 32, 40, 48]))
        N = (e, e + 8 * int(rng.integers(0, 2)), e)
        pbc = bool(i % 4)
        gd = GridDescriptor(N, np.array(N) * 0.2 / BOHR, pbc)
        nb = int(rng.integers(1, 5))
        a = torch.randn((nb,) + tuple(gd.n_c), dtype=torch.complex128,
                        device='cuda', generator=g)
        ph = gd.phase_cd(rng.uniform(-0.5, 0.5, 3)) if pbc else None
        groups.append((Laplace(gd, -0.5, 3, complex), a, torch.empty_like(a), ph))
        algo += 32.0 * nb * np.prod(gd.n_c)
    batch = BatchApply(max_groups=len(groups))
    tb = event_time(torch, lambda: batch(groups), 10)
    outs = [b.clone() for _, _, b, _ in groups]
    ts = event_time(torch, sequential, 10)
    same = all(torch.equal(o, b) for o, (_, _, b, _) in zip(outs, groups))
    del groups, outs
    out['many_small'] = {
        'config': 'complex128 nn=3 Laplace: 96 groups of 1-4 bands on grids of edges '
                  '24-56, 3 of 4 periodic with their own k-point',
        'batched_ms': round(tb * 1e3, 3), 'batched_GBps': round(algo / tb / 1e9, 1),
        'sequential_ms': round(ts * 1e3, 3),
        'speedup_batched_over_sequential': round(ts / tb, 3),
        'bit_identical': bool(same), 'bytes': algo}
    return out


def preconditioner_c3(torch, np):
    """Eigensolver preconditioner (gpaw/preconditioner.py) on the C3 bands:
    256 complex128 bands of 64^3, periodic, Bloch phases, nn=3 kinetic
    operator on the fine grid.  Algorithmic bytes per call: every array a
    launch reads or writes counted once per launch, 16 B per point:
    5 fine + 14 coarse-1 + 7 coarse-2 grid functions per band."""
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.preconditioner import Preconditioner
    N, nb = 64, 256
    gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, True)
    k_c = np.random.default_rng(5).uniform(-0.5, 0.5, 3)

    class Kpt:
        phase_cd = gd.phase_cd(k_c)
    pc = Preconditioner(gd, Laplace(gd, -0.5, 3, complex), complex, nb)
    pc.allocate()
    g = torch.Generator(device='cuda').manual_seed(13)
    r = torch.randn((nb, N, N, N), dtype=torch.complex128, device='cuda',
                    generator=g)
    t = event_time(torch, lambda: pc(r, Kpt), 5)
    n0 = N ** 3
    algo = 16.0 * nb * (5 * n0 + 14 * n0 / 8 + 7 * n0 / 64)
    del r
    return {'config': 'Preconditioner(nn=3 kin), complex128, 256 bands x '
                      '64^3, Bloch phases',
            'avg_call_ms': round(t * 1e3, 3),
            'GBps': round(algo / t / 1e9, 1),
            'bytes_per_call': algo}


def callers_c3(torch, np):
    """The hot path's callers (gpaw_amd.callers) at production shapes:
    * pseudo-Hamiltonian: nn=3 kinetic stencil over 256 complex128 bands of
      64^3 (Bloch phases) plus the local potential, one fused pass;
      algorithmic bytes: 32 per band point (read psi, write Htpsi) and the
      potential read once per call;
    * density interpolation: 2 spins, 64^3 -> 128^3 zero boundary, nn=3,
      with the charge renormalisation (interpolation reads 8 B per coarse
      and writes 8 B per fine point; the scaling reads and writes the fine
      grid once more and the integral reads it once);
    * potential restriction: 2 spins, vt_g += vHt_g (24 B per fine point)
      and nn=3 restriction to 64^3 (8 B per fine and per coarse point);
    * pair potentials: 8 pair-density solves on a 64^3 zero-boundary grid
      (63^3 arrays), eps = 1e-12, half of them charged;
    * RPA coupling (finegrid == 1): 12 pair densities solved on the 128^3
      zero-boundary fine grid (127^3, eps = 2e-10, charge=None), potentials
      restricted (nn=3) to 64^3, upper triangle of the coupling integrals."""
    from gpaw_amd import callers
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import PoissonSolver
    from gpaw_amd.transformers import Transformer
    from gpaw_amd.utilities import gaussian_density, random_gaussian_charges
    out = {}
    N, nb = 64, 256
    gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, True)
    g = torch.Generator(device='cuda').manual_seed(21)
    psi = torch.randn((nb, N, N, N), dtype=torch.complex128, device='cuda', generator=g)
    hpsi = torch.empty_like(psi)
    v = torch.randn((N, N, N), dtype=torch.float64, device='cuda', generator=g)
    kin = Laplace(gd, -0.5, 3, complex)
    ph = gd.phase_cd(np.random.default_rng(5).uniform(-0.5, 0.5, 3))
    t = event_time(torch, lambda: callers.apply_pseudo_hamiltonian(kin, v, psi, hpsi, ph), 10)
    algo = 32.0 * nb * N ** 3 + 8.0 * N ** 3
    out['pseudo_hamiltonian'] = {'config': 'nn=3 kinetic + local potential in one pass, '
                                           'complex128, 256 bands x 64^3, Bloch phases',
                                 'avg_call_ms': round(t * 1e3, 3),
                                 'GBps': round(algo / t / 1e9, 1)}
    del psi, hpsi
    # density interpolation and potential restriction, 2 spins
    cg = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, False)
    fg = cg.refine()
    ns = 2
    nt_sG = torch.rand((ns,) + tuple(cg.n_c), dtype=torch.float64, device='cuda', generator=g)
    nt_sg = torch.empty((ns,) + tuple(fg.n_c), dtype=torch.float64, device='cuda')
    itp = Transformer(cg, fg, 3)
    t = event_time(torch, lambda: callers.interpolate_density(itp, fg, nt_sG, nt_sg,
                                                              charge=-8.0), 10)
    nc, nf = float(np.prod(cg.n_c)), float(np.prod(fg.n_c))
    algo = ns * (8 * nc + 8 * nf) + ns * nf * 24
    out['interpolate_density'] = {'config': '2 spins, 64^3 -> 128^3 zero boundary, nn=3, '
                                            'charge renormalisation',
                                  'avg_call_ms': round(t * 1e3, 3),
                                  'GBps': round(algo / t / 1e9, 1)}
    vt_sg = torch.randn((ns,) + tuple(fg.n_c), dtype=torch.float64, device='cuda', generator=g)
    vH = torch.randn(tuple(fg.n_c), dtype=torch.float64, device='cuda', generator=g)
    vt_sG = torch.empty((ns,) + tuple(cg.n_c), dtype=torch.float64, device='cuda')
    rst = Transformer(fg, cg, 3)
    t = event_time(torch, lambda: callers.restrict_potential(rst, vt_sg, vH, vt_sG), 10)
    algo = ns * (24 * nf + 8 * nf + 8 * nc)
    out['restrict_potential'] = {'config': '2 spins, 128^3 -> 64^3 zero boundary, nn=3',
                                 'avg_call_ms': round(t * 1e3, 3),
                                 'GBps': round(algo / t / 1e9, 1)}
    del nt_sG, nt_sg, vt_sg, vH, vt_sG
    # pair potentials
    pg = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, False)
    rng = np.random.default_rng(64)
    rho = []
    for p in range(8):
        c, w, q = random_gaussian_charges(pg, rng, 2, 0.5 / BOHR, 1.0 / BOHR, 2.0 / BOHR)
        if p % 2:
            q = np.abs(q)   # a charged pair density (n1 == n2 in the exchange loop)
        rho.append(gaussian_density(pg, c, w, q))
    rho_d = torch.from_numpy(np.stack(rho)).to('cuda')
    charges = [-float(pg.integrate(r)) if p % 2 else 0.0 for p, r in enumerate(rho)]
    ps = PoissonSolver(nn=3, relax='J')
    ps.set_grid_descriptor(pg)
    ps.initialize()
    callers.pair_potentials(ps, rho_d, charges, eps=1e-12)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, niter = callers.pair_potentials(ps, rho_d, charges, eps=1e-12)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['pair_potentials'] = {'config': '8 pair densities, 64^3 zero boundary (63^3), '
                                        'eps=1e-12, 4 charged',
                              'solves_per_s': round(8 / dt, 2),
                              'iterations': niter}
    del rho_d
    # RPA coupling on the fine grid, integrals on the coarse one
    fz = cg.refine()
    rng = np.random.default_rng(128)
    rho_f, rho_c = [], []
    for p in range(12):
        c, w, q = random_gaussian_charges(fz, rng, 2, 0.5 / BOHR, 1.0 / BOHR, 2.0 / BOHR)
        rho_f.append(gaussian_density(fz, c, w, q))
        rho_c.append(gaussian_density(cg, c, w, q))
    rf = torch.from_numpy(np.stack(rho_f)).to('cuda')
    rc = torch.from_numpy(np.stack(rho_c)).to('cuda')
    ps = PoissonSolver(nn=3, relax='J')
    ps.set_grid_descriptor(fz)
    ps.initialize()
    rst = Transformer(fz, cg, 3)
    callers.rpa_coupling(ps, rf[:2], gd=cg, restrictor=rst, rhot_pG=rc[:2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, niter = callers.rpa_coupling(ps, rf, gd=cg, restrictor=rst, rhot_pG=rc)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['rpa_coupling'] = {'config': '12 pair densities, solved on 128^3 zero boundary (127^3), '
                                     'restricted to 64^3, coupling integrals',
                           'pairs_per_s': round(12 / dt, 2),
                           'iterations': niter}
    return out


METRIC = ('mixed vs FP64 Poisson solves/s @1e-10 resid; FD GB/s % of peak; '
          'p99 latency')


def mixed_vs_fp64(torch, np, N=160, h=0.18, count=32, reps=10):
    """C4: mixed-precision vs plain FP64 multigrid on the 160^3 zero-boundary
    configuration (159^3 arrays, h = 0.18 A, 32 seeded neutral Gaussians of
    widths 0.4-1.2 A), all to ||L phi - rho|| / ||rho|| <= 1e-10.

    Lines: 'fp64' = the reference's V-cycles (FP64 PoissonSolver);
    'mixed' = FP32 V-cycle preconditioning FP64 flexible CG (the product's
    mixed solver); 'pcg_fp64_vcycle' = the same CG with the same V-cycle in
    FP64 (what the outer iteration alone gains); 'refine_fp32' = FP32
    V-cycles inside FP64 iterative refinement (two V-cycles per
    correction).  mixed / pcg_fp64_vcycle is what FP32 itself contributes;
    the per-iteration time ratio compares one FP32 step with one FP64 step
    of the same loop."""
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import MixedPoissonSolver, PoissonSolver
    from gpaw_amd.utilities import gaussian_density, random_gaussian_charges
    gd = GridDescriptor((N,) * 3, (N * h / BOHR,) * 3, False)
    rng = np.random.default_rng(160)
    c, w, q = random_gaussian_charges(gd, rng, count, 0.4 / BOHR, 1.2 / BOHR,
                                      2.0 / BOHR)
    rho = gaussian_density(gd, c, w, q)
    rho_d = torch.from_numpy(rho).to('cuda')
    eps, _ = eps_for(np, gd, rho)
    res = {}
    phis = {}
    solvers = (('fp64', lambda: PoissonSolver(nn=3, eps=eps)),
               ('mixed', lambda: MixedPoissonSolver(nn=3, rel_tol=1e-10)),
               ('pcg_fp64_vcycle', lambda: MixedPoissonSolver(
                   nn=3, rel_tol=1e-10, inner_precision='fp64')),
               ('refine_fp32', lambda: MixedPoissonSolver(
                   nn=3, rel_tol=1e-10, outer='refine', inner=2)))
    for name, make in solvers:
        solver = make()
        solver.set_grid_descriptor(gd)
        solver.initialize()
        phi = torch.zeros(tuple(gd.n_c), dtype=torch.float64, device='cuda')
        for _ in range(2):                     # warm-up (tuning, graphs)
            phi.zero_()
            it = solver.solve(phi, rho_d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            phi.zero_()
            it = solver.solve(phi, rho_d)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = {'solves_per_s': round(1.0 / dt, 3),
                     'ms_per_solve': round(dt * 1e3, 3),
                     'iterations': it,
                     'ms_per_iteration': round(dt * 1e3 / max(it, 1), 4)}
        phis[name] = phi
        del solver
    bg = float(rho.sum()) / np.prod(gd.get_size_of_global_array())
    rho_n = torch.from_numpy(rho - bg).to('cuda')
    from gpaw_amd.fd_operators import Laplace
    lap = Laplace(gd, -0.25 / np.pi, 3)
    r = torch.empty_like(rho_n)
    for name in phis:
        lap.operator.residual(phis[name], rho_n, r)
        res[name]['rel_residual'] = float(torch.linalg.vector_norm(r).item() /
                                          torch.linalg.vector_norm(rho_n).item())
        if name != 'fp64':
            d = (phis[name] - phis['fp64']).abs().max().item()
            res[name]['rel_maxnorm_err_vs_fp64'] = d / phis['fp64'].abs().max().item()
    ms = {k: res[k]['ms_per_solve'] for k in res}
    res['speedup_mixed_over_fp64'] = round(ms['fp64'] / ms['mixed'], 3)
    res['speedup_pcg_fp64_vcycle_over_fp64'] = round(ms['fp64'] / ms['pcg_fp64_vcycle'], 3)
    res['speedup_refine_fp32_over_fp64'] = round(ms['fp64'] / ms['refine_fp32'], 3)
    res['fp32_contribution'] = round(ms['pcg_fp64_vcycle'] / ms['mixed'], 3)
    res['fp32_over_fp64_iteration_time'] = round(
        res['mixed']['ms_per_iteration'] / res['pcg_fp64_vcycle']['ms_per_iteration'], 3)
    # one inner V-cycle alone (gpaw_poisson_precondition: the same schedule
    # in FP32 and in FP64; the FP32 call also converts r in and z out)
    pc = MixedPoissonSolver(nn=3)
    pc.set_grid_descriptor(gd)
    pc.initialize()
    z = torch.empty_like(rho_n)
    tv = {p: event_time(torch, lambda p=p: pc.precondition(rho_n, z, p), 20)
          for p in ('fp32', 'fp64')}
    res['inner_vcycle'] = {
        'fp32_ms': round(tv['fp32'] * 1e3, 4), 'fp64_ms': round(tv['fp64'] * 1e3, 4),
        'fp32_over_fp64': round(tv['fp32'] / tv['fp64'], 3),
        'note': 'one preconditioning V-cycle from zero (2 + 2 Chebyshev-weighted sweeps on '
                'levels 0 and 1, reference hierarchy below), FP32 including the FP64 <-> '
                'FP32 conversion of r and z'}
    del pc
    res['config'] = ('C4: 160^3 zero boundary (159^3 arrays), h=0.18 A, 32 neutral '
                     'Gaussians, 1e-10 relative residual')
    return res


def serving(torch, np, n=2000, rate=None, slots=4, budget_gb=64):
    """C5: a seeded Poisson-arrival stream of n requests (80 % mixed-
    precision Poisson solves to 1e-10, 20 % Laplace applies on 1-64 bands;
    edges 48..160, half periodic), data resident in HBM, served by the
    continuous-batching server; results are taken by the caller as they
    complete.  First the capacity: every request released at t = 0 (the
    slots never idle), requests/s over the span.  Then latency: the same
    requests arriving as a Poisson process at 80 % of that capacity (or at
    `rate` when given), p50 / p99 from arrival to result.  One warm-up
    request per key first, so that hierarchies are resident."""
    from gpaw_amd import ops
    from gpaw_amd.serving import Request, Server, synthetic_stream
    # requests come with arbitrary band counts: no shape tuning on the
    # serving path (the heuristic shapes are used)
    ops.autotune(False)
    t0 = time.perf_counter()
    base_rate = 300.0
    reqs = synthetic_stream(n=n, rate=base_rate, seed=5)
    setup = time.perf_counter() - t0

    def drop(r):
        r.result = None
    srv = Server(slots=slots, hbm_budget=budget_gb << 30, on_result=drop)
    seen = {}
    for r in reqs:
        seen.setdefault(r.key, r)
    warm = [Request(-1 - i, r.kind, r.N, r.pbc, r.payload, 0.0)
            for i, r in enumerate(seen.values())]
    srv.run(warm, realtime=False)

    def fresh(scale):
        return [Request(r.rid, r.kind, r.N, r.pbc, r.payload, r.arrival * scale)
                for r in reqs]
    cap = srv.run(fresh(1.0), realtime=False)
    capacity = cap['requests_per_s']
    load = rate if rate else 0.8 * capacity
    rep = srv.run(fresh(base_rate / load), realtime=True)
    out = {'capacity': {'requests_per_s': round(capacity, 1),
                        'poisson_solves_per_s': round(cap['poisson_solves_per_s'], 1),
                        'span_s': round(cap['span_s'], 3),
                        'completed': cap['completed'],
                        'note': 'all requests released at t=0; latency there is '
                                'queueing time and not reported'},
           'load_requests_per_s': round(load, 1),
           'load_fraction_of_capacity': round(load / capacity, 3) if capacity else None}
    out.update(rep)
    out.update({'slots': slots, 'hbm_budget_GB': budget_gb,
                'stream_setup_s': round(setup, 2), 'plans': srv.stats(),
                'config': 'C5: %d requests, 80%% mixed-precision Poisson (1e-10) / '
                          '20%% Laplace nn=3 on 1-64 bands, edges 48-160, half '
                          'periodic; Laplace requests waiting together go out as '
                          'one batched apply' % n})
    for k in ('p50_ms', 'p99_ms', 'span_s', 'requests_per_s',
              'poisson_solves_per_s'):
        out[k] = round(out[k], 3)
    return out


def _cpu_solves(O, kind, gd_shape_args, rho, eps, budget):
    """Full solves of the workload with the oracle V-cycle over `kind`'s
    kernels for about `budget` seconds (at most 10): (solves, seconds,
    V-cycles per solve)."""
    og = O.Grid(*gd_shape_args)
    ps = O.PoissonSolver(nn=3, relax='J', kernels=kind)
    ps.set_grid_descriptor(og)
    ps.initialize()
    n = 0
    t0 = time.perf_counter()
    niter = None
    while True:
        phi = og.zeros()
        niter = ps.solve(phi, rho.copy(), eps=eps)
        n += 1
        if time.perf_counter() - t0 > budget or n >= 10:
            break
    return n, time.perf_counter() - t0, niter


def cpu_baseline(np, gd_shape_args, rho, eps, budget):
    """The reference's own C kernels (oracle/_ref, compiled from the
    reference sources) under the restated V-cycle driver: on one pinned host
    core, and beside it on every core of the process with the reference's
    GPAW_OMP_MONLY threads (its fd / restrict / interpolate split over
    OMP_NUM_THREADS pthreads per call; its relax, most of the V-cycle, is
    serial in the reference)."""
    import oracle as O
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    kind = 'reference' if O.Kernels.available('reference') else 'port'
    ctx = threadpool_limits(1) if threadpool_limits else None
    try:
        cores = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cores = [0]
    omp = os.environ.get('OMP_NUM_THREADS')
    try:
        try:
            os.sched_setaffinity(0, {cores[0]})
        except (AttributeError, OSError):
            pass
        if ctx:
            ctx.__enter__()
        n, dt, niter = _cpu_solves(O, kind, gd_shape_args, rho, eps, budget)
        out = {'value': round(n / dt, 4), 'unit': 'solves/s', 'cores': 1, 'kind': kind,
               'sample': '%d full solve(s) of the same 95^3 workload (%d V-cycles each), '
                         '%.1f s, one pinned core, BLAS threads = 1' % (n, niter, dt)}
        try:
            os.sched_setaffinity(0, set(cores))
        except (AttributeError, OSError):
            pass
        if O.Kernels.available('reference_mt') and len(cores) > 1:
            nth = len(cores)
            if omp and omp.isdigit() and 0 < int(omp) < nth:
                nth = int(omp)
            os.environ['OMP_NUM_THREADS'] = str(nth)
            n, dt, niter = _cpu_solves(O, 'reference_mt', gd_shape_args, rho, eps, budget)
            out['all_cores'] = {
                'value': round(n / dt, 4), 'unit': 'solves/s', 'cores': nth,
                'kind': 'reference',
                'sample': '%d full solve(s), %.1f s, the reference built with GPAW_OMP_MONLY '
                          '(%d threads in fd / restrict / interpolate; relax serial as in the '
                          'reference)' % (n, dt, nth)}
    finally:
        if omp is None:
            os.environ.pop('OMP_NUM_THREADS', None)
        else:
            os.environ['OMP_NUM_THREADS'] = omp
        try:
            os.sched_setaffinity(0, set(cores))
        except (AttributeError, OSError):
            pass
        if ctx:
            ctx.__exit__(None, None, None)
    return out


SECTIONS = {
    'fd_laplace': lambda torch, np, args: fd_laplace_c3(torch, np),
    'preconditioner': lambda torch, np, args: preconditioner_c3(torch, np),
    'callers': lambda torch, np, args: callers_c3(torch, np),
    'mixed_precision': lambda torch, np, args: mixed_vs_fp64(torch, np),
    'serving': lambda torch, np, args: serving(torch, np, n=args.serve_n,
                                               rate=args.serve_rate),
    'copy': lambda torch, np, args: {'copy_GBps': copy_bandwidth(torch)},
}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only (--gpus %d requested): this '
              'build measures the single-GPU share of the path; multi-GPU '
              'domain decomposition is not built' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from gpaw_amd import ops
    from gpaw_amd.poisson import PoissonSolver

    # launch shapes of the stencil kernels are picked by the library's
    # autotuner on first use (untimed warm-up below)
    ops.autotune(True)

    if args.only:
        print(json.dumps({args.only: SECTIONS[args.only](torch, np, args)}),
              flush=True)
        return
    # the headline solve runs the library's default launch shapes (for its
    # level-0 passes the box kernel with 4 planes per block, which the tuner
    # also picks there, within the spread of its timings): fixed shapes keep
    # every run of it, profiled runs included, on the same kernels
    ops.autotune(False)
    gd, rho = c2_inputs(np)
    eps, rho_norm = eps_for(np, gd, rho)
    ps = PoissonSolver(nn=3, relax='J', eps=eps)
    ps.set_grid_descriptor(gd)
    ps.initialize()
    rho_d = torch.from_numpy(rho).to('cuda')
    phi = torch.zeros(tuple(gd.n_c), dtype=torch.float64, device='cuda')

    def step():
        phi.zero_()
        return ps.solve(phi, rho_d)

    for _ in range(args.warmup):
        niter = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        niter = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    # verify the stopping point independently: r = L phi - (rho - bg)
    bg = float(rho.sum()) / np.prod(gd.get_size_of_global_array())
    rho_n = torch.from_numpy(rho - bg).to('cuda')
    r = torch.empty_like(phi)
    ps.operators[0].operator.residual(phi, rho_n, r)
    rel = float(torch.linalg.vector_norm(r).item()) / float(
        torch.linalg.vector_norm(rho_n).item())

    out = {
        # the benchmark's compound metric; `value` is its solves/s on the
        # plain FP64 configuration (96^3), the other parts are named in
        # `metric_parts` with the configuration each comes from
        'metric': METRIC,
        'value': round(args.steps / dt, 3),
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(dt / args.steps * 1e3, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': 'PoissonSolver(nn=3, J), 96^3 zero-boundary '
                               '(95^3 arrays), h=0.2 A, 8 neutral Gaussians',
                   'grid': [95, 95, 95], 'levels': ps.levels,
                   'vcycles_per_solve': niter,
                   'eps_reference_criterion': eps,
                   'final_rel_residual': rel,
                   'parallelism': 'single GPU'},
    }
    # the same solve with host (numpy) arrays: copies in and out over PCIe
    # included; reported beside `value`, which keeps inputs resident in HBM
    phi_h = np.zeros(tuple(gd.n_c))
    ps.solve(phi_h, rho.copy())
    t0 = time.perf_counter()
    for _ in range(3):
        phi_h[:] = 0.0
        ps.solve(phi_h, rho.copy())
    out['config']['host_arrays_ms_per_solve'] = round((time.perf_counter() - t0) / 3 * 1e3, 3)
    out['roofline'] = jacobi_roofline(torch, np, ps, rho_n)
    out['config']['launch_shapes'] = (
        'level-0 Jacobi and residual passes: the default shapes (box kernel, 32 x 8 columns '
        'x 4 planes per block); launch-shape tuning is off for this solve and on for the '
        'other sections')
    out['roofline']['copy_GBps_measured'] = copy_bandwidth(torch)
    ops.autotune(True)
    if not args.no_extra:
        out['fd_laplace'] = fd_laplace_c3(torch, np)
        out['preconditioner'] = preconditioner_c3(torch, np)
        out['callers'] = callers_c3(torch, np)
        out['mixed_precision'] = mixed_vs_fp64(torch, np)
        out['serving'] = serving(torch, np, n=args.serve_n, rate=args.serve_rate)
    if not args.no_cpu:
        cb = cpu_baseline(np, ((96,) * 3, gd.cell_cv, False), rho, eps,
                          args.cpu_seconds)
        out['cpu_baseline'] = cb
    out['metric_parts'] = metric_parts(out)
    print(json.dumps(out), flush=True)


def metric_parts(out):
    """Each part of the compound metric, its value and where it comes from
    in this line."""
    parts = {'FP64 Poisson solves/s at 1e-10 (value)': {
        'value': out['value'], 'config': 'C2: 96^3 zero boundary, nn=3 Jacobi, '
                                          'reference V-cycles'}}
    mp = out.get('mixed_precision')
    if mp:
        parts['mixed vs FP64 Poisson solves/s at 1e-10 (mixed_precision)'] = {
            'mixed_solves_per_s': mp['mixed']['solves_per_s'],
            'fp64_solves_per_s': mp['fp64']['solves_per_s'],
            'ratio': mp['speedup_mixed_over_fp64'],
            'config': 'C4: 160^3 zero boundary, 32 Gaussians'}
    fd = out.get('fd_laplace')
    if fd:
        parts['FD GB/s, % of peak (fd_laplace)'] = {
            'GBps': fd['GBps'], 'pct_peak': fd['pct_peak'],
            'config': 'C3: complex128 nn=3 Laplace, 256 bands x 64^3, Bloch phases'}
    sv = out.get('serving')
    if sv:
        parts['p99 latency (serving)'] = {
            'p99_ms': sv['p99_ms'], 'p50_ms': sv['p50_ms'],
            'load_requests_per_s': sv.get('load_requests_per_s'),
            'capacity_requests_per_s': sv.get('capacity', {}).get('requests_per_s'),
            'config': 'C5: 2000 mixed requests at 80 % of the measured capacity'}
    return parts


if __name__ == '__main__':
    main()
