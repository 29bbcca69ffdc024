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

    def execute(self, plan, reqs, slot):
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
    fake = _FakeGPU()
    srv = Server(slots=1, hbm_budget=1 << 40, executor=fake.execute,
                 make_plan=fake.plan)
    reqs = [Request(k, 'laplace', 16, True, np.zeros((2, 16, 16, 16)), 0.0)
            for k in range(5)]
    srv.run(reqs, realtime=False)
    assert max(len(b) for b in fake.batches) > 1


def test_server_respects_budget():
    fake = _FakeGPU()
    # room for one 24^3 Poisson plan + request at a time
    budget = plan_bytes('poisson', 24, False) + 3 * 8 * 23 ** 3
    srv = Server(slots=4, hbm_budget=budget, executor=fake.execute,
                 make_plan=fake.plan)
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
