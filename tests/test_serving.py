"""Request server: admission by HBM budget, plan residency, continuous
batching (host logic on the CPU with a stand-in executor), and on the GPU
that every result is identical alone and in the mix."""
import threading
import time

import numpy as np
import pytest

from gpaw_amd.serving import Admission, Request, Server, plan_bytes


def test_admission_budget_and_lru_eviction():
    a = Admission(100)
    assert a.try_admit(True, 'k1', 40, 10)
    a.add_plan(1, 'k1', 40)
    assert a.used() == 50
    # a second plan does not fit while the first is busy
    assert not a.try_admit(True, 'k2', 60, 10)
    a.release(1, 10)
    # now plan 1 is idle and gets evicted to make room
    assert a.try_admit(True, 'k2', 60, 10)
    assert a.evicted == [1]
    a.add_plan(2, 'k2', 60)
    assert a.used() == 70


def test_plan_bytes_scale():
    assert plan_bytes('laplace', 64, True) == 0
    b = plan_bytes('poisson', 96, False)
    assert 48 * 95 ** 3 < b < 60 * 95 ** 3


class _FakeGPU:
    """Stand-in executor: sleeps proportionally to the work, records the
    maximum number of concurrently running batches."""

    def __init__(self):
        self.running = 0
        self.max_running = 0
        self.lock = threading.Lock()
        self.batches = []

    def plan(self, key):
        return {'key': key}

    def execute(self, plans, reqs, slot):
        assert len(plans) == len(reqs)
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.batches.append([r.rid for r in reqs])
        time.sleep(0.002 * len(reqs))
        for r in reqs:
            r.result = r.rid
        with self.lock:
            self.running -= 1


def _stream(n, rng):
    reqs = []
    t = 0.0
    for k in range(n):
        t += rng.exponential(0.001)
        kind = 'poisson' if rng.random() < 0.8 else 'laplace'
        N = int(rng.choice([16, 24]))
        payload = np.zeros((int(rng.integers(1, 4)),) + (N - 1,) * 3) \
            if kind == 'laplace' else np.zeros((N - 1,) * 3)
        reqs.append(Request(k, kind, N, False, payload, t))
    return reqs


def test_server_completes_everything_with_bounded_concurrency():
    fake = _FakeGPU()
    srv = Server(slots=3, hbm_budget=1 << 40, executor=fake.execute,
                 make_plan=fake.plan)
    reqs = _stream(60, np.random.default_rng(1))
    rep = srv.run(reqs, realtime=True)
    assert rep['completed'] == 60 and rep['failed'] == 0
    assert all(r.result == r.rid for r in reqs)
    assert 1 <= fake.max_running <= 3
    assert rep['p99_ms'] >= rep['p50_ms'] > 0


def test_server_fuses_waiting_laplace_requests():
    """Waiting Laplace requests of different grids and boundaries go out as
    one batch (one batched apply), each with its own key's plan."""
    fake = _FakeGPU()
    srv = Server(slots=1, hbm_budget=1 << 40, executor=fake.execute,
                 make_plan=fake.plan)
    reqs = [Request(k, 'laplace', (16, 24)[k % 2], k % 3 == 0,
                    np.zeros((2,) + (((16, 24)[k % 2] - (k % 3 != 0)),) * 3), 0.0)
            for k in range(6)]
    srv.run(reqs, realtime=False)
    assert max(len(b) for b in fake.batches) > 1
    keys = {r.key for r in reqs}
    assert len(keys) > 2 and set(srv.laplace_plans) == keys


def test_server_respects_budget():
    fake = _FakeGPU()
    # room for one 24^3 Poisson plan + request at a time
    budget = plan_bytes('poisson', 24, False) + 3 * 8 * 23 ** 3
    srv = Server(slots=4, hbm_budget=budget, executor=fake.execute,
                 make_plan=fake.plan, on_result=lambda r: None)
    reqs = [Request(k, 'poisson', 24, False, np.zeros((23,) * 3), 0.0)
            for k in range(6)]
    rep = srv.run(reqs, realtime=False)
    assert rep['completed'] == 6
    assert fake.max_running == 1


def test_server_reports_failures_and_continues():
    def bad(plan, reqs, slot):
        if reqs[0].rid == 2:
            raise RuntimeError('boom')
        for r in reqs:
            r.result = 1
    srv = Server(slots=2, hbm_budget=1 << 40, executor=bad,
                 make_plan=lambda k: {})
    reqs = [Request(k, 'poisson', 16, True, np.zeros((16,) * 3), 0.0)
            for k in range(4)]
    rep = srv.run(reqs, realtime=False)
    assert rep['failed'] == 1 and rep['completed'] == 3
    assert 'boom' in reqs[2].error


def _alone(reqs, keep):
    """Serve each kept request alone on a fresh one-slot server; returns
    rid -> result."""
    alone = Server(slots=1, hbm_budget=64 << 30)
    out = {}
    for r in reqs:
        if r.rid not in keep:
            continue
        q = Request(r.rid, r.kind, r.N, r.pbc, r.payload, 0.0)
        alone.run([q], realtime=False)
        assert q.error is None, q.error
        out[r.rid] = alone.take(q)
    return out


@pytest.mark.gpu
def test_budget_below_working_set_on_device():
    """An HBM budget that holds about one hierarchy at a time: plans are
    evicted and rebuilt and requests queue for memory on the device, and
    every result still equals the same request served alone."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from gpaw_amd.serving import synthetic_stream
    print('case serving tight budget', flush=True)
    reqs = synthetic_stream(n=48, rate=2000.0, seed=11, edges=(32, 48, 64),
                            max_bands=8)
    budget = plan_bytes('poisson', 64, True) + 4 * 3 * 8 * 64 ** 3
    kept = {}
    srv = Server(slots=4, hbm_budget=budget,
                 on_result=lambda r: kept.__setitem__(r.rid, r.result))
    rep = srv.run(reqs, realtime=False)
    assert rep['completed'] == len(reqs), [r.error for r in reqs if r.error]
    st = srv.stats()
    print('plans built %d evicted %d' % (st['plans_built'], st['plans_evicted']),
          flush=True)
    assert st['plans_evicted'] >= 2
    assert srv.adm.used() <= budget and srv.adm.results == 0
    ref = _alone(reqs, set(kept))
    for rid, res in kept.items():
        assert torch.equal(res, ref[rid]), rid


@pytest.mark.gpu
def test_c5_stream_alone_and_in_mix():
    """The C5 configuration at its stated size: 2000 seeded Poisson-arrival
    requests (80 % mixed-precision Poisson solves to 1e-10, 20 % Laplace
    applies on 1-64 bands; edges 48..160, half periodic), all released at
    once so that the four slots stay full; a seeded sample of 250 results
    is compared bit for bit with the same requests served alone."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from gpaw_amd.serving import synthetic_stream
    print('case serving C5 2000 requests', flush=True)
    reqs = synthetic_stream(n=2000, rate=300.0, seed=5)
    rng = np.random.default_rng(2000)
    sample = set(int(i) for i in rng.choice(len(reqs), 250, replace=False))
    kept = {}

    def keep(r):
        if r.rid in sample:
            kept[r.rid] = r.result
        r.result = None
    srv = Server(slots=4, hbm_budget=64 << 30, on_result=keep)
    rep = srv.run(reqs, realtime=False)
    print('C5 mix: %d completed, %.1f req/s, p99 %.1f ms' % (
        rep['completed'], rep['requests_per_s'], rep['p99_ms']), flush=True)
    assert rep['completed'] == len(reqs), [r.error for r in reqs if r.error][:3]
    assert len(kept) == len(sample)
    assert sum(1 for r in reqs if r.rid in sample and r.kind == 'laplace') > 20
    ref = _alone(reqs, sample)
    for rid in sorted(sample):
        assert torch.equal(kept[rid], ref[rid]), rid


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from gpaw_amd.serving import synthetic_stream
    print('case serving mix vs alone', flush=True)
    reqs = synthetic_stream(n=40, rate=2000.0, seed=3, edges=(24, 32, 48),
                            max_bands=8)
    mix = Server(slots=4, hbm_budget=8 << 30)
    rep = mix.run(reqs, realtime=True)
    assert rep['completed'] == len(reqs), [r.error for r in reqs if r.error]
    mixed_results = [r.result.clone() for r in reqs]
    alone = Server(slots=1, hbm_budget=8 << 30)
    for r, ref in zip(reqs, mixed_results):
        q = Request(r.rid, r.kind, r.N, r.pbc, r.payload, 0.0)
        alone.run([q], realtime=False)
        assert q.error is None
        assert torch.equal(q.result, ref), (r.rid, r.kind, r.N, r.pbc)


def test_request_larger_than_budget_fails_cleanly():
    fake = _FakeGPU()
    srv = Server(slots=2, hbm_budget=4096, executor=fake.execute,
                 make_plan=fake.plan)
    reqs = [Request(0, 'poisson', 64, False, np.zeros((63,) * 3), 0.0),
            Request(1, 'laplace', 4, True, np.zeros((1, 4, 4, 4)), 0.0)]
    rep = srv.run(reqs, realtime=False)
    assert reqs[0].error and 'budget' in reqs[0].error
    assert reqs[1].result == 1 and rep['completed'] == 1


def test_reused_plan_is_not_evicted_for_its_own_request():
    """With room for one plan, a request whose key has an idle plan reuses
    it instead of evicting it and rebuilding."""
    fake = _FakeGPU()
    built = []

    def plan(key):
        built.append(key)
        return {'key': key}
    budget = plan_bytes('poisson', 24, False) + 3 * 8 * 23 ** 3
    srv = Server(slots=1, hbm_budget=budget + 1000, executor=fake.execute,
                 make_plan=plan, on_result=lambda r: None)
    reqs = [Request(k, 'poisson', 24, False, np.zeros((23,) * 3), 0.0)
            for k in range(4)]
    rep = srv.run(reqs, realtime=False)
    assert rep['completed'] == 4
    assert built == [('poisson', 24, False)]


def test_results_stay_charged_until_taken():
    fake = _FakeGPU()
    each = 3 * 8 * 15 ** 3
    budget = plan_bytes('poisson', 16, False) + each + 2 * 8 * 15 ** 3
    srv = Server(slots=1, hbm_budget=budget, executor=fake.execute,
                 make_plan=fake.plan)
    reqs = [Request(k, 'poisson', 16, False, np.zeros((15,) * 3), 0.0)
            for k in range(4)]
    rep = srv.run(reqs, realtime=False)
    # three requests run one after another while their results pile up; the
    # fourth no longer fits beside the plan and three held results, nothing
    # can be evicted (its plan is reserved for it) and nobody takes results
    assert rep['completed'] == 3 and 'results held' in reqs[3].error
    assert srv.adm.results == sum(8 * 15 ** 3 for r in reqs if r.error is None)
    for r in reqs:
        if r.error is None:
            assert srv.take(r) == r.rid
    assert srv.adm.results == 0
    # with a consumer the same stream completes
    taken = []
    srv2 = Server(slots=1, hbm_budget=budget, executor=fake.execute,
                  make_plan=fake.plan, on_result=lambda r: taken.append(r.rid))
    reqs = [Request(k, 'poisson', 16, False, np.zeros((15,) * 3), 0.0)
            for k in range(4)]
    assert srv2.run(reqs, realtime=False)['completed'] == 4
    assert sorted(taken) == [0, 1, 2, 3] and srv2.adm.results == 0
