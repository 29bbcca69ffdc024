"""Request server for Poisson solves and batched Laplace applies.

A stream of requests (mixed grid sizes and boundary conditions) is served
with continuous batching on one GPU:

* slots      S worker threads, each with its own HIP stream.  A request runs
             on a free slot; requests on different slots overlap on the
             device.  Library calls release the GIL (ctypes), so the host
             side of one solve (one 8-byte readback per chunk of V-cycles)
             does not stall the others.
* residency  multigrid hierarchies (PoissonSolver plans) stay on the device
             between requests, keyed by (grid edge, boundary); a plan serves
             one request at a time and more than one plan per key is built
             when requests of the same key overlap.  Laplace operators hold
             no device memory and are kept per key.
* admission  a request is admitted when a slot is free AND its device bytes
             (resident plan if not yet built + its arrays) fit the HBM budget
             together with everything resident, in flight or held as
             results; idle plans are evicted least-recently-used to make
             room (the plan a request is about to reuse is never the one
             evicted for it).  Requests are admitted in arrival order (no
             starvation), so latency includes queueing.
* results    a finished request's result stays charged to the budget until
             the caller takes it: with `on_result` the server hands each
             result to that callback as it completes and then releases it;
             without, results stay charged until `take(request)`.
* batching   Laplace requests waiting together -- of any grid, boundary and
             band count, one dtype -- go to the GPU as one batched apply
             (one persistent launch per stencil range, each request with its
             own operator and output; results bit-identical to single
             applies).

Latency is measured from a request's arrival time to the moment its result
is ready; the server reports p50/p99 and throughput.
"""
import collections
import threading
import time

import numpy as np

BOHR = 0.52917720859


class Request:
    """kind: 'poisson' or 'laplace'; N: cubic grid edge; pbc: bool;
    payload: density (poisson, array of the grid's shape) or bands
    (laplace, array [nbands, *grid]), host numpy or GPU tensor; arrival:
    seconds on the server's clock."""
    __slots__ = ('rid', 'kind', 'N', 'pbc', 'payload', 'arrival', 'h',
                 'result', 'done', 'niter', 'error', 'start', 'finish',
                 'held')

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
        self.held = 0

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


def _bands(a):
    return int(a.shape[0]) if len(a.shape) == 4 else 1


def request_bytes(r):
    """Device bytes of a request's own arrays while it runs."""
    n = r.n_points()
    if r.kind == 'poisson':
        return 3 * 8 * n                      # rho, rho - background, phi
    return 2 * (16 if _is_complex(r.payload) else 8) * n * _bands(r.payload)


def result_bytes(r):
    """Device bytes of a finished request's result (phi, or the output
    bands)."""
    n = r.n_points()
    if r.kind == 'poisson':
        return 8 * n
    return (16 if _is_complex(r.payload) else 8) * n * _bands(r.payload)


class Admission:
    """HBM-budget bookkeeping (pure host logic)."""

    def __init__(self, budget_bytes):
        self.budget = budget_bytes
        self.resident = collections.OrderedDict()   # plan id -> (key, bytes, busy)
        self.inflight = 0
        self.results = 0                            # finished, not yet taken
        self.evicted = []                           # ids evicted, not yet dropped
        self.n_evicted = 0

    def used(self):
        return (self.inflight + self.results +
                sum(b for _, b, _ in self.resident.values()))

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
        """Mark an idle plan busy (and most recently used)."""
        key, b, _ = self.resident.pop(pid)
        self.resident[pid] = (key, b, True)

    def idle_plan(self, pid):
        key, b, _ = self.resident[pid]
        self.resident[pid] = (key, b, False)

    def release(self, pid, req_nbytes):
        if pid is not None:
            self.idle_plan(pid)
        self.inflight -= req_nbytes

    def evict(self, pid):
        self.resident.pop(pid)
        self.evicted.append(pid)
        self.n_evicted += 1


class Server:
    """Continuous-batching server.

    executor(plans, requests, slot) runs one batch on a slot and fills
    request.result (plans[i] serves requests[i]); make_plan(key) builds a
    resident plan.  Both default to the GPU implementation.  on_result(r),
    if given, receives every finished request (result set) and the server
    releases the result's budget charge once it returns."""

    def __init__(self, slots=4, hbm_budget=64 << 30, executor=None,
                 make_plan=None, rel_tol=1e-10, mixed=True, clock=None,
                 max_fuse_bands=256, on_result=None):
        self.slots = slots
        self.adm = Admission(hbm_budget)
        self.rel_tol = rel_tol
        self.mixed = mixed
        self.executor = executor or self._gpu_execute
        self.make_plan = make_plan or self._gpu_plan
        self.clock = clock or time.perf_counter
        self.max_fuse_bands = max_fuse_bands
        self.on_result = on_result
        self.plans = {}              # pid -> plan object
        self.plan_key = {}           # pid -> key
        self.free_plans = collections.defaultdict(list)
        self.laplace_plans = {}      # key -> plan (no device memory)
        self.next_pid = 0
        self.lock = threading.Condition()
        self.waiting = collections.deque()
        self.done = []
        self.busy_slots = 0
        self.t0 = None
        self._streams = None
        self._batches = None

    # ------------------------------------------------------------ GPU parts
    def _gpu_plan(self, key):
        kind, N, pbc = key
        from .grid_descriptor import GridDescriptor
        gd = GridDescriptor((N,) * 3, (N * 0.2 / BOHR,) * 3, pbc)
        if kind == 'poisson':
            from .poisson import MixedPoissonSolver, PoissonSolver
            if self.mixed:
                # results must not depend on which requests a plan served
                # before (each equals the same request served alone)
                ps = MixedPoissonSolver(nn=3, rel_tol=self.rel_tol, reuse_history=False)
            else:
                ps = PoissonSolver(nn=3)
            ps.set_grid_descriptor(gd)
            ps.initialize()
            return ps
        from .fd_operators import Laplace
        return {'gd': gd, 'real': Laplace(gd, -0.5, 3, float),
                'complex': Laplace(gd, -0.5, 3, complex)}

    def _gpu_execute(self, plans, reqs, slot):
        """Payloads may be GPU tensors (served data already in HBM: no
        copies, the result stays on the device) or numpy arrays (copied in,
        result copied back)."""
        import torch
        stream = self._streams[slot]
        with torch.cuda.stream(stream):
            on_dev = isinstance(reqs[0].payload, torch.Tensor)
            if reqs[0].kind == 'poisson':
                r, plan = reqs[0], plans[0]
                rho = r.payload if on_dev else torch.from_numpy(r.payload).to('cuda')
                phi = torch.zeros_like(rho)
                if self.mixed:
                    r.niter = plan.solve(phi, rho)
                else:
                    eps = plan.gd.dv * (self.rel_tol * _neutral_norm(rho)) ** 2
                    r.niter = plan.solve(phi, rho, eps=eps)
                r.result = phi if on_dev else phi.cpu().numpy()
            else:
                groups = []
                for q, plan in zip(reqs, plans):
                    a = q.payload if on_dev else torch.from_numpy(q.payload).to('cuda')
                    cplx = a.is_complex()
                    groups.append((plan['complex' if cplx else 'real'], a,
                                   torch.empty_like(a),
                                   np.ones((3, 2), complex) if cplx else None))
                self._batches[slot](groups)
                for q, (_, _, b, _) in zip(reqs, groups):
                    q.result = b if on_dev else b.cpu().numpy()
            stream.synchronize()

    # ------------------------------------------------------------- serving
    def _new_plan(self, key):
        pid = self.next_pid
        self.next_pid += 1
        self.plans[pid] = None        # built by the worker, outside the lock
        self.plan_key[pid] = key
        self.adm.add_plan(pid, key, plan_bytes(*key))
        return pid

    def _fail_never_fits(self, r, nbytes):
        r.error = ('needs %d device bytes, HBM budget is %d' %
                   (nbytes, self.adm.budget))
        r.finish = self.clock() - self.t0
        self.waiting.popleft()
        self.done.append(r)

    def _admit_locked(self):
        """Admit requests in arrival order while slots and memory allow."""
        batches = []
        while self.waiting and self.busy_slots + len(batches) < self.slots:
            r = self.waiting[0]
            key = r.key
            idle_now = self.busy_slots == 0 and not batches
            if r.kind == 'poisson':
                # reserve an idle plan of this key first, so that eviction
                # for this very request cannot drop it
                pid = self.free_plans[key].pop() if self.free_plans[key] else None
                if pid is not None:
                    self.adm.use_plan(pid)
                nbytes = request_bytes(r)
                if not self.adm.try_admit(pid is None, key, plan_bytes(*key), nbytes):
                    if pid is not None:
                        self.adm.idle_plan(pid)
                        self.free_plans[key].append(pid)
                    if idle_now and not self.adm.results:
                        # nothing runs, nothing can be evicted: it never fits
                        self._fail_never_fits(r, nbytes + plan_bytes(*key))
                        continue
                    break
                self._drop_evicted()
                self.waiting.popleft()
                if pid is None:
                    pid = self._new_plan(key)
                batches.append(([pid], [r], nbytes))
                continue
            # Laplace: this request and the Laplace requests waiting behind
            # it (any grid, one dtype) in one batched apply
            cplx = _is_complex(r.payload)
            group, nb = [r], _bands(r.payload)
            for q in list(self.waiting)[1:]:
                if q.kind == 'laplace' and _is_complex(q.payload) == cplx and \
                        nb + _bands(q.payload) <= self.max_fuse_bands:
                    group.append(q)
                    nb += _bands(q.payload)
            while group:
                nbytes = sum(request_bytes(q) for q in group)
                if self.adm.try_admit(False, key, 0, nbytes):
                    break
                group.pop()           # fewer fused requests may still fit
            if not group:
                if idle_now and not self.adm.results:
                    self._fail_never_fits(r, request_bytes(r))
                    continue
                break
            self._drop_evicted()
            for q in group:
                self.waiting.remove(q)
            batches.append(([None] * len(group), group, nbytes))
        return batches

    def _drop_evicted(self):
        for pid in self.adm.evicted:
            key = self.plan_key.pop(pid)
            if pid in self.free_plans[key]:
                self.free_plans[key].remove(pid)
            self.plans.pop(pid, None)
        self.adm.evicted.clear()

    def _laplace_plan(self, key):
        with self.lock:
            plan = self.laplace_plans.get(key)
        if plan is None:
            plan = self.make_plan(key)
            with self.lock:
                plan = self.laplace_plans.setdefault(key, plan)
        return plan

    def _worker(self, slot, pids, group, nbytes):
        try:
            plans = []
            for pid, q in zip(pids, group):
                if pid is None:
                    plans.append(self._laplace_plan(q.key))
                else:
                    if self.plans.get(pid) is None:
                        self.plans[pid] = self.make_plan(self.plan_key[pid])
                    plans.append(self.plans[pid])
            now = self.clock() - self.t0
            for q in group:
                q.start = now
            self.executor(plans, group, slot)
        except Exception as e:          # the request fails, the server runs on
            for q in group:
                q.error = repr(e)
                q.result = None
        fin = self.clock() - self.t0
        with self.lock:
            for q in group:
                q.finish = fin
                q.held = result_bytes(q) if q.error is None else 0
                self.adm.results += q.held
                self.done.append(q)
            for pid in pids:
                if pid is not None:
                    self.adm.idle_plan(pid)
                    self.free_plans[self.plan_key[pid]].append(pid)
            self.adm.inflight -= nbytes
            self.busy_slots -= 1
            self.free_slot_ids.append(slot)
            self.lock.notify_all()
        if self.on_result is not None:
            for q in group:
                try:
                    self.on_result(q)
                finally:
                    self.take(q)

    def take(self, r):
        """Hand a finished request's result to the caller and release its
        budget charge; returns the result."""
        with self.lock:
            self.adm.results -= r.held
            r.held = 0
            self.lock.notify_all()
        return r.result

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from start).  With
        realtime=False every request is available at t=0 (throughput mode,
        latencies then include queueing from t=0)."""
        import torch
        if self._streams is None and self.executor == self._gpu_execute:
            from .batch import BatchApply
            self._streams = [torch.cuda.Stream() for _ in range(self.slots)]
            self._batches = [BatchApply(max_groups=max(64, self.max_fuse_bands))
                             for _ in range(self.slots)]
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
                for pids, group, nb in self._admit_locked():
                    slot = self.free_slot_ids.pop()
                    self.busy_slots += 1
                    t = threading.Thread(target=self._worker,
                                         args=(slot, pids, group, nb))
                    t.start()
                    threads.append(t)
                if self.waiting and not self.busy_slots and self.adm.results \
                        and self.on_result is None:
                    # everything left is blocked by results nobody took
                    for r in list(self.waiting):
                        r.error = 'blocked: results held by the caller fill the HBM budget'
                        r.finish = self.clock() - self.t0
                        self.done.append(r)
                    self.waiting.clear()
                    continue
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

    def stats(self):
        """Plans built and evicted, largest number resident so far."""
        return {'plans_built': self.next_pid, 'plans_evicted': self.adm.n_evicted}


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
