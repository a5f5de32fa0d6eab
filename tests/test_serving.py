"""Serving loop host logic (CPU, with a stand-in runner and clock) and one GPU serving run
whose counts equal the batched runner's own counts."""
import numpy as np
import pytest

from subgraphmatching_amd import serving
from subgraphmatching_amd.serving import admit_slots, poisson_arrivals, serve


class _Res:
    def __init__(self, ans):
        self.ans = ans


class FakeRunner:
    """Answers query q with count q.k; each batch takes `cost` seconds of the fake clock."""

    def __init__(self, slots, clock, cost=0.01):
        self.slots, self.clock, self.cost = slots, clock, cost
        self.batches = []
        self.part = None

    def filter(self, part):
        self.part = part
        self.batches.append([q.k for q in part])
        return [True] * len(part), None, None

    def enumerate(self, orders=None, cap=None, **kw):
        self.clock.t += self.cost
        return [_Res(min(q.k, cap)) for q in self.part]

    def last_ms(self):
        return 1.0, 2.0


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, s):
        self.t += max(s, 1e-6)


class Q:
    def __init__(self, k):
        self.k = k


def test_poisson_arrivals_rate_and_seed():
    a = poisson_arrivals(20000, 500.0, seed=3)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 500.0) < 15.0
    assert np.array_equal(a, poisson_arrivals(20000, 500.0, seed=3))


def test_admission_by_memory():
    slots, stack, per = admit_slots(1 << 22, 169000, mem_budget_bytes=32 << 30, waves=4096)
    assert stack == 4096 * 4 * (1 << 20) and per == 24 * (1 << 22) + (1 << 14) + serving.SPLIT_SLOT_BYTES
    assert slots == min(64, ((32 << 30) - stack) // per)
    with pytest.raises(ValueError):
        admit_slots(1 << 22, 169000, mem_budget_bytes=stack, waves=4096)


def test_serve_fifo_batches_and_latency():
    clock = FakeClock()
    r = FakeRunner(slots=4, clock=clock, cost=0.01)
    qs = [Q(k) for k in range(30)]
    arrivals = np.linspace(0.0, 0.05, 30)
    res = serve(r, qs, arrivals, cap=20, clock=clock, sleep=clock.sleep)
    assert res.completed == 30 and res.counts == [min(k, 20) for k in range(30)]
    flat = [k for b in r.batches for k in b]
    assert flat == list(range(30))                    # first come, first served
    assert all(1 <= len(b) <= 4 for b in r.batches)   # never above the admitted slots
    assert res.batches == len(r.batches)
    assert res.latency_ms["p50"] > 0 and res.latency_ms["p99"] >= res.latency_ms["p50"]
    assert res.device_ms == {"filter": 1.0 * res.batches, "enumerate": 2.0 * res.batches}


def test_serve_waits_for_arrivals():
    clock = FakeClock()
    r = FakeRunner(slots=8, clock=clock, cost=0.001)
    qs = [Q(1) for _ in range(5)]
    arrivals = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    res = serve(r, qs, arrivals, cap=10, clock=clock, sleep=clock.sleep)
    assert [len(b) for b in r.batches] == [1, 1, 1, 1, 1]  # each query alone: the next is not there yet
    assert res.latency_ms["max"] < 10.0


@pytest.mark.gpu
def test_gpu_serving_counts_match_batch_runner():
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd.generators import chung_lu, query_set
    g = chung_lu(3000, 2.3, 8.0, seed=9)
    qs = query_set(g, 40, lambda i, r: int(r.integers(4, 14)), lambda i, r: bool(i % 2), seed=91)
    runner = sgm.Batch(g, slots=8)
    ref = []
    for lo in range(0, len(qs), 8):
        ok, _, _ = runner.filter(qs[lo:lo + 8])
        ref += [r.ans if good else 0 for good, r in zip(ok, runner.enumerate(cap=100000))]
    res = serve(runner, qs, poisson_arrivals(len(qs), 2000.0, seed=4), cap=100000)
    assert res.counts == ref
    assert res.latency_ms["p99"] > 0
