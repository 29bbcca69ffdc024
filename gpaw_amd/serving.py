"""Request server for Poisson solves and batched Laplace applies.

A stream of requests (mixed grid sizes and boundary conditions) is served
with continuous batching on one GPU:

* slots      S worker threads, each with its own HIP stream.  A request runs
             on a free slot; requests on different slots overlap on the
             device.  Library calls release the GIL (ctypes), so the host
             side of one solve (one 8-byte readback per V-cycle) does not
             stall the others.
* residency  multigrid hierarchies (PoissonSolver plans) and Laplace
             operators stay on the device between requests, keyed by
             (kind, grid, boundary); a plan serves one request at a time and
             more than one plan per key is built when requests of the same
             key overlap.
* admission  a request is admitted when a slot is free AND its device bytes
             (resident plan if not yet built + its arrays) fit the HBM budget
             together with everything already resident or in flight; idle
             plans are evicted least-recently-used to make room.  Requests
             are admitted in arrival order (no starvation), so latency
             includes queueing.
* batching   Laplace requests of the same key that are waiting together are
             fused into one launch over all their bands.

Latency is measured from a request's arrival time to the moment its result
is ready on the host; the server reports p50/p99 and throughput.
"""
import collections
import heapq
import threading
import time

import numpy as np

BOHR = 0.52917720859


class Request:
    """kind: 'poisson' or 'laplace'; N: cubic grid edge; pbc: bool;
    payload: density (poisson, host array of the grid's shape) or bands
    (laplace, host array [nbands, *grid]); arrival: seconds on the
    server's clock."""
    __slots__ = ('rid', 'kind', 'N', 'pbc', 'payload', 'arrival', 'h',
                 'result', 'done', 'niter', 'error', 'start', 'finish')

    def __init__(self, rid, kind, N, pbc, payload, arrival=0.0, h=0.2 / BOHR):
        self.rid = rid
        self.kind = kind
        self.N = N
        self.pbc = pbc
        self.payload = payload
        self.arrival = arrival
        self.h = h
        self.result = None
        self.niter = None
        self.error = None
        self.start = None
        self.finish = None

    @property
    def key(self):
        return (self.kind, self.N, bool(self.pbc))

    def n_points(self):
        n = self.N if self.pbc else self.N - 1
        return n ** 3


def _neutral_norm(rho):
    """2-norm of rho minus its mean (the neutral density the solver sees)."""
    import torch
    return float(torch.linalg.vector_norm(rho - rho.mean()).item())


def plan_bytes(kind, N, pbc):
    """Device bytes a resident plan of this key holds: the Poisson plan
    keeps 4 FP64 arrays per level (phi pair, rho, residual) plus 8-byte
    FP32 mirrors for mixed precision; a Laplace operator holds none."""
    if kind != 'poisson':
        return 0
    n = N if pbc else N - 1
    total = 0
    for _ in range(5):
        total += 4 * 8 * n ** 3 + 4 * 4 * n ** 3
        n = max(1, (n + (0 if pbc else 1)) // 2 - (0 if pbc else 1))
    return total


def _is_complex(a):
    try:
        return bool(a.is_complex())
    except AttributeError:
        return bool(np.iscomplexobj(a))


def request_bytes(r):
    """Device bytes of a request's own arrays while it runs."""
    n = r.n_points()
    if r.kind == 'poisson':
        return 3 * 8 * n                      # rho, rho - background, phi
    nb = r.payload.shape[0]
    return 2 * (16 if _is_complex(r.payload) else 8) * n * nb


class Admission:
    """HBM-budget bookkeeping (pure host logic)."""

    def __init__(self, budget_bytes):
        self.budget = budget_bytes
        self.resident = collections.OrderedDict()   # plan id -> (key, bytes, busy)
        self.inflight = 0
        self.evicted = []                           # ids evicted, not yet dropped

    def used(self):
        return self.inflight + sum(b for _, b, _ in self.resident.values())

    def try_admit(self, need_plan, plan_key, plan_nbytes, req_nbytes):
        """Return True (and reserve) if the request fits, evicting idle
        plans least-recently-used if that makes room."""
        extra = req_nbytes + (plan_nbytes if need_plan else 0)
        while self.used() + extra > self.budget:
            victim = next((pid for pid, (_, _, busy) in self.resident.items()
                           if not busy), None)
            if victim is None:
                return False
            self.evict(victim)
        self.inflight += req_nbytes
        return True

    def add_plan(self, pid, key, nbytes):
        self.resident[pid] = (key, nbytes, True)

    def use_plan(self, pid):
        key, b, _ = self.resident.pop(pid)
        self.resident[pid] = (key, b, True)     # most recently used

    def release(self, pid, req_nbytes):
        key, b, _ = self.resident[pid]
        self.resident[pid] = (key, b, False)
        self.inflight -= req_nbytes

    def evict(self, pid):
        self.resident.pop(pid)
        self.evicted.append(pid)


class Server:
    """Continuous-batching server.  `executor(plan, requests, stream)` runs
    one batch on a slot and fills request.result; `make_plan(key)` builds a
    resident plan.  Both default to the GPU implementation."""

    def __init__(self, slots=4, hbm_budget=64 << 30, executor=None,
                 make_plan=None, rel_tol=1e-10, mixed=True, clock=None):
        self.slots = slots
        self.adm = Admission(hbm_budget)
        self.rel_tol = rel_tol
        self.mixed = mixed
        self.executor = executor or self._gpu_execute
        self.make_plan = make_plan or self._gpu_plan
        self.clock = clock or time.perf_counter
        self.plans = {}              # pid -> plan object
        self.plan_key = {}           # pid -> key
        self.free_plans = collections.defaultdict(list)
        self.next_pid = 0
        self.lock = threading.Condition()
        self.waiting = collections.deque()
        self.done = []
        self.busy_slots = 0
        self.t0 = None
        self._streams = None

    # ------------------------------------------------------------ GPU parts
    def _gpu_plan(self, key):
        kind, N, pbc = key
        from .grid_descriptor import GridDescriptor
        gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, pbc)
        if kind == 'poisson':
            from .poisson import MixedPoissonSolver, PoissonSolver
            if self.mixed:
                ps = MixedPoissonSolver(nn=3, rel_tol=self.rel_tol)
            else:
                ps = PoissonSolver(nn=3)
            ps.set_grid_descriptor(gd)
            ps.initialize()
            return ps
        from .fd_operators import Laplace
        return {'gd': gd, 'real': Laplace(gd, -0.5, 3, float),
                'complex': Laplace(gd, -0.5, 3, complex)}

    def _gpu_execute(self, plan, reqs, slot):
        """Payloads may be GPU tensors (served data already in HBM: no
        copies, the result stays on the device) or numpy arrays (copied in,
        result copied back)."""
        import torch
        stream = self._streams[slot]
        with torch.cuda.stream(stream):
            on_dev = isinstance(reqs[0].payload, torch.Tensor)
            if reqs[0].kind == 'poisson':
                r = reqs[0]
                rho = r.payload if on_dev else torch.from_numpy(r.payload).to('cuda')
                phi = torch.zeros_like(rho)
                if self.mixed:
                    r.niter = plan.solve(phi, rho)
                else:
                    eps = plan.gd.dv * (self.rel_tol * _neutral_norm(rho)) ** 2
                    r.niter = plan.solve(phi, rho, eps=eps)
                r.result = phi if on_dev else phi.cpu().numpy()
            else:
                if len(reqs) == 1:
                    a = reqs[0].payload if on_dev else torch.from_numpy(reqs[0].payload).to('cuda')
                else:
                    a = torch.cat([q.payload if on_dev else torch.from_numpy(q.payload)
                                   for q in reqs]).to('cuda')
                b = torch.empty_like(a)
                cplx = a.is_complex()
                op = plan['complex' if cplx else 'real']
                op.apply(a, b, np.ones((3, 2), complex) if cplx else None)
                k = 0
                for q in reqs:
                    nb = q.payload.shape[0]
                    q.result = b[k:k + nb] if on_dev else b[k:k + nb].cpu().numpy()
                    k += nb
            stream.synchronize()

    # ------------------------------------------------------------- serving
    def _get_plan(self, key):
        if self.free_plans[key]:
            pid = self.free_plans[key].pop()
            self.adm.use_plan(pid)
            return pid
        pid = self.next_pid
        self.next_pid += 1
        self.plans[pid] = None        # built by the worker, outside the lock
        self.plan_key[pid] = key
        self.adm.add_plan(pid, key, plan_bytes(*key))
        return pid

    def _admit_locked(self):
        """Admit requests in arrival order while slots and memory allow."""
        batches = []
        while self.waiting and self.busy_slots + len(batches) < self.slots:
            r = self.waiting[0]
            key = r.key
            need_plan = not self.free_plans[key]
            group = [r]
            if r.kind == 'laplace':     # fuse waiting applies of the same key
                for q in list(self.waiting)[1:]:
                    if q.key == key and _is_complex(q.payload) == \
                            _is_complex(r.payload) and \
                            sum(g.payload.shape[0] for g in group) + \
                            q.payload.shape[0] <= 256:
                        group.append(q)
            nbytes = sum(request_bytes(q) for q in group)
            if not self.adm.try_admit(need_plan, key, plan_bytes(*key), nbytes):
                if self.busy_slots == 0 and not batches:
                    # nothing runs and nothing can be evicted: it never fits
                    r.error = ('needs %d device bytes, HBM budget is %d' %
                               (nbytes + plan_bytes(*key), self.adm.budget))
                    r.finish = self.clock() - self.t0
                    self.waiting.popleft()
                    self.done.append(r)
                    continue
                break
            for pid in self.adm.evicted:
                self._drop_plan(pid)
            self.adm.evicted.clear()
            for q in group:
                self.waiting.remove(q)
            pid = self._get_plan(key)
            batches.append((pid, group, nbytes))
        return batches

    def _drop_plan(self, pid):
        key = self.plan_key.pop(pid)
        if pid in self.free_plans[key]:
            self.free_plans[key].remove(pid)
        self.plans.pop(pid, None)

    def _worker(self, slot, pid, group, nbytes):
        try:
            if self.plans.get(pid) is None:
                self.plans[pid] = self.make_plan(self.plan_key[pid])
            now = self.clock() - self.t0
            for q in group:
                q.start = now
            self.executor(self.plans[pid], group, slot)
        except Exception as e:          # the request fails, the server runs on
            for q in group:
                q.error = repr(e)
        fin = self.clock() - self.t0
        with self.lock:
            for q in group:
                q.finish = fin
                self.done.append(q)
            self.adm.release(pid, nbytes)
            self.free_plans[self.plan_key[pid]].append(pid)
            self.busy_slots -= 1
            self.free_slot_ids.append(slot)
            self.lock.notify_all()

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from start).  With
        realtime=False every request is available at t=0 (throughput mode,
        latencies then include queueing from t=0)."""
        import torch
        if self._streams is None and self.executor == self._gpu_execute:
            self._streams = [torch.cuda.Stream() for _ in range(self.slots)]
        pending = sorted(requests, key=lambda r: r.arrival)
        self.free_slot_ids = list(range(self.slots))
        self.t0 = self.clock()
        threads = []
        i = 0
        with self.lock:
            while i < len(pending) or self.waiting or self.busy_slots:
                now = self.clock() - self.t0
                while i < len(pending) and (not realtime or
                                            pending[i].arrival <= now):
                    self.waiting.append(pending[i])
                    i += 1
                for pid, group, nb in self._admit_locked():
                    slot = self.free_slot_ids.pop()
                    self.busy_slots += 1
                    t = threading.Thread(target=self._worker,
                                         args=(slot, pid, group, nb))
                    t.start()
                    threads.append(t)
                if i < len(pending) and realtime:
                    wait = pending[i].arrival - (self.clock() - self.t0)
                    self.lock.wait(timeout=max(0.0, min(wait, 0.05)))
                elif self.waiting or self.busy_slots:
                    self.lock.wait(timeout=0.05)
        for t in threads:
            t.join()
        return self.report(requests)

    @staticmethod
    def report(requests):
        ok = [r for r in requests if r.error is None and r.finish is not None]
        lat = np.array([r.finish - r.arrival for r in ok]) if ok else np.zeros(1)
        span = max(r.finish for r in ok) - min(r.arrival for r in ok) if ok else 0
        npois = sum(1 for r in ok if r.kind == 'poisson')
        return {'requests': len(requests), 'completed': len(ok),
                'failed': len(requests) - len(ok),
                'p50_ms': float(np.percentile(lat, 50) * 1e3),
                'p99_ms': float(np.percentile(lat, 99) * 1e3),
                'span_s': float(span),
                'requests_per_s': len(ok) / span if span > 0 else 0.0,
                'poisson_solves_per_s': npois / span if span > 0 else 0.0}


def synthetic_stream(n=2000, rate=200.0, seed=0, edges=(48, 64, 80, 96, 128, 160),
                     frac_poisson=0.8, max_bands=64, variants=3, device=True):
    """Seeded Poisson-arrival stream: 80 % Poisson solves, 20 % Laplace
    applies on 1-64 bands; cubic edges from `edges`, half periodic.
    Densities are sums of 4 neutral grid-normalised Gaussians (widths
    0.5-1.5 A, >= 2 A from the walls, or a quarter of the cell on grids
    smaller than 16 A); bands are normal(0, 1).  Each
    (edge, boundary) key draws from `variants` seeded densities and one
    64-band block (requests use its first nb bands), generated once and
    kept in HBM when device=True -- the served data are resident, the
    stream's requests reference them."""
    from .grid_descriptor import GridDescriptor
    from .utilities import gaussian_density, random_gaussian_charges
    rng = np.random.default_rng(seed)
    dens, bands = {}, {}
    if device:
        import torch
        gen = torch.Generator(device='cuda').manual_seed(seed)

    def density(N, pbc, v):
        k = (N, pbc, v)
        if k not in dens:
            gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, pbc)
            r2 = np.random.default_rng(hash(k) % 2 ** 32)
            # 2 A from the walls, or a quarter of the cell on small grids
            margin = min(2.0, 0.25 * N * 0.2) / BOHR
            c, w, q = random_gaussian_charges(gd, r2, 4, 0.5 / BOHR, 1.5 / BOHR,
                                              margin)
            d = gaussian_density(gd, c, w, q)
            dens[k] = torch.from_numpy(d).to('cuda') if device else d
        return dens[k]

    def band_block(N, pbc):
        k = (N, pbc)
        if k not in bands:
            m = N if pbc else N - 1
            if device:
                bands[k] = torch.randn((max_bands, m, m, m), dtype=torch.float64,
                                       device='cuda', generator=gen)
            else:
                bands[k] = np.random.default_rng(hash(k) % 2 ** 32).normal(
                    size=(max_bands, m, m, m))
        return bands[k]

    t = 0.0
    reqs = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        N = int(rng.choice(edges))
        pbc = bool(rng.integers(0, 2))
        if rng.random() < frac_poisson:
            v = int(rng.integers(0, variants))
            reqs.append(Request(k, 'poisson', N, pbc, density(N, pbc, v), t))
        else:
            nb = int(rng.integers(1, max_bands + 1))
            reqs.append(Request(k, 'laplace', N, pbc, band_block(N, pbc)[:nb], t))
    return reqs
