"""Host logic of the serving loop (no GPU): the instance list, the arrival
process and the admission by HBM footprint."""
import numpy as np
import pytest

from tbfsolver_amd import serving as S
from tbfsolver_amd.batch import NXZ, NY


def test_instance_list_sizes_and_mix():
    specs = S.serving_instances(256, seed=1004)
    assert [s[0] for s in specs] == list(range(256))
    assert all(s[1][0] in S.SERVE_NXZ and s[1][2] in S.SERVE_NXZ and s[1][1] in S.SERVE_NY for s in specs)
    assert specs[0][1] == (256, 192, 256)   # the workload spans up to the largest channel
    assert set(NXZ) < set(S.SERVE_NXZ) and set(NY) < set(S.SERVE_NY)
    mg = [s for s in specs if s[2] == "MG"]
    assert 0.6 < len(mg) / 256 < 0.8  # 70 % PCG + MG
    assert all(4 <= s[3] <= 24 for s in mg)
    assert all(s[3] == 0 for s in specs if s[2] == "FAST")
    assert specs == S.serving_instances(256, seed=1004)  # seeded
    assert specs != S.serving_instances(256, seed=1005)


def test_poisson_arrivals_rate_and_order():
    reqs = S.poisson_arrivals(1000.0, 4.0, 256, seed=3)
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) > 0) and t[0] > 0 and t[-1] < 4.0
    assert abs(len(reqs) - 4000) < 4 * np.sqrt(4000)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1e-3) < 0.1e-3
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1  # exponential: coefficient of variation 1
    idx = np.array([r[1] for r in reqs])
    assert idx.min() >= 0 and idx.max() < 256


def test_admission_lru_skips_busy():
    a = S.HbmAdmission(100)
    for i, b in ((0, 30), (1, 30), (2, 30)):
        a.add(i, b)
    assert a.used == 90
    assert a.victims(10, busy=set()) == []
    assert a.victims(20, busy=set()) == [0]          # least recently used first
    a.touch(0)
    assert a.victims(20, busy=set()) == [1]
    assert a.victims(20, busy={1}) == [2]             # a busy instance is never evicted
    assert a.victims(70, busy={1}) == [2, 0]
    assert a.victims(70, busy={1, 2}) is None         # the budget is held by busy instances
    with pytest.raises(ValueError):
        a.victims(101, busy=set())
    a.remove(1)
    assert a.used == 60 and list(a.held) == [2, 0]


def test_footprint_estimate_covers_fields():
    n = (64, 32, 64)
    cells = 64 * 32 * 64
    for kind in ("MG", "FAST"):
        for prec in ("mixed", "fp64"):
            est = S.estimate_footprint(n, kind, prec)
            assert est > 8 * 17 * cells
    assert S.estimate_footprint(n, "MG", "fp64") > S.estimate_footprint(n, "MG", "mixed")


class _Stand:
    """A stand-in instance: a step takes a few milliseconds of host time and
    its result is a function of the instance and the step number only."""

    log = []

    def __init__(self, spec, precision, seed, state=None):
        self.index, self.n, self.kind = spec[0], spec[1], spec[2]
        self.steps = 0 if state is None else state["steps"]
        self.acc = 0 if state is None else state["acc"]
        self.footprint = S.estimate_footprint(self.n, self.kind, precision)
        self.busy = False

    def step(self, stream=None):
        import time

        assert not self.busy  # one request of an instance at a time
        self.busy = True
        time.sleep(0.002)
        self.acc = self.acc * 31 + self.index * 7 + self.steps
        _Stand.log.append((self.index, self.steps))
        self.steps += 1
        self.busy = False
        return _Val(self.acc)

    def state(self):
        return {"steps": self.steps, "acc": self.acc}

    def close(self):
        pass


class _Val(int):
    def clone(self):
        return _Val(self)


class _HostLoop(S.ServingLoop):
    resident = _Stand

    def _streams(self):
        return [None] * self.workers

    def _on(self, s):
        import contextlib

        return contextlib.nullcontext()

    def _wait(self, s):
        pass

    def _sync(self):
        pass


SPECS = [(i, (32 + 32 * (i % 3), 32, 32 + 32 * (i % 2)), "MG" if i % 3 else "FAST", 4) for i in range(10)]


def _serve(workers, budget=1e12, rate=3000.0, n=60):
    reqs = S.poisson_arrivals(rate, n / rate, len(SPECS), seed=4)
    loop = _HostLoop(SPECS, workers=workers, hbm_budget=budget, keep_results=True)
    out = loop.run(reqs)
    return reqs, out, dict(loop.results), loop.evictions


def test_loop_results_and_order_independent_of_workers():
    reqs, out4, r4, _ = _serve(4)
    _, out1, r1, _ = _serve(1)
    assert r4 == r1 and len(r4) == len(reqs)
    assert out4["requests"] == len(reqs) and all(l >= 0 for l in out4["latency_ms"])
    assert out4["p99_ms"] >= out4["p50_ms"] > 0 and out4["throughput_rps"] > 0
    # several requests in flight at once with 4 workers: the pass takes less time
    assert out4["wall_s"] < out1["wall_s"]


def test_loop_per_instance_arrival_order():
    _Stand.log = []
    reqs, _, _, _ = _serve(4)
    for i in range(len(SPECS)):
        assert [s for j, s in _Stand.log if j == i] == list(range(sum(1 for _, k in reqs if k == i)))


def test_loop_evictions_restore_state():
    budget = 3 * max(S.estimate_footprint(s[1], s[2], "mixed") for s in SPECS)
    _, _, r_ev, ev = _serve(2, budget=budget)
    _, _, r_all, ev0 = _serve(2)
    assert ev > 0 and ev0 == 0
    assert r_ev == r_all
