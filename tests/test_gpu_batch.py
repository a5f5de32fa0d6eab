"""GPU parity of the batched runner (sgm_batch_*): graph-replayed filter and enumeration
over several batches (full and partial) against the oracle, on a Chung-Lu graph and a small
R-MAT graph with the 4-32 vertex sparse/dense query mix."""
import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, rmat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def g_cl():
    return chung_lu(2000, 2.3, 8.0, seed=3)


@pytest.fixture(scope="module")
def g_rmat():
    return rmat(13, 60000, seed=5)


def _mix(g, count, seed, lo=4, hi=33):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: i % 2 == 0, seed=seed)


def _check_filter(g, qs, ok, cnt, dsum, kind=L.SGM_FILTER_GQL):
    deg = np.diff(g.off.astype(np.int64))
    for i, q in enumerate(qs):
        ok_r, c_r = (O.gql if kind == L.SGM_FILTER_GQL else O.nlf)(g, q)
        n = q.getVertexCnt()
        assert ok[i] == ok_r, i
        assert [int(x) for x in cnt[i, :n]] == [len(c) for c in c_r], i
        assert [int(x) for x in dsum[i, :n]] == [int(deg[np.asarray(c, dtype=np.int64)].sum()) for c in c_r], i


MAX_DROPPED = 0.25  # at most this fraction of a batch's live queries may pass the oracle's time limit


def _check_enum(g, qs, ok, res, orders, cap, fs=True, limit_ms=20000.0):
    """Compares every live query with the oracle; a query whose oracle run passes limit_ms
    is left out, and the check fails when more than MAX_DROPPED of them are."""
    checked = dropped = 0
    for i, q in enumerate(qs):
        if not ok[i]:
            assert res[i].ans == 0 and res[i].intersections == 0
            continue
        _, c_r = O.gql(g, q)
        order = orders[i] if orders is not None and orders[i] is not None else O.order_ri(q)
        ref = O.enum_cache(g, q, c_r, order, cap=cap, fs=fs, time_limit_ms=limit_ms)
        if ref["timed_out"]:
            say("oracle time limit on query", i, q.getVertexCnt())
            dropped += 1
            continue
        assert res[i].ans == ref["count"], (i, res[i].ans, ref["count"])
        if ref["count"] < cap:
            assert res[i].intersections == ref["intersections"], i
            assert res[i].nodes == ref["nodes"][:q.getVertexCnt()], i
        checked += 1
    assert dropped <= MAX_DROPPED * (checked + dropped), (dropped, checked)
    return checked


def test_batch_replay_matches_oracle(g_cl):
    qs = _mix(g_cl, 36, seed=81, hi=20)
    b = sgm.Batch(g_cl, slots=16)
    for lo in range(0, len(qs), 16):  # 16, 16, 4: the captured graphs are replayed, last batch partial
        part = qs[lo:lo + 16]
        say("batch", lo, len(part))
        ok, cnt, dsum = b.filter(part)
        _check_filter(g_cl, part, ok, cnt, dsum)
        res = b.enumerate(cap=200000)
        assert _check_enum(g_cl, part, ok, res, None, 200000) > 0
    b.close()


def test_batch_nlf_filter_and_plain_cache(g_cl):
    qs = _mix(g_cl, 10, seed=82, hi=12)
    b = sgm.Batch(g_cl, slots=16)
    ok, cnt, dsum = b.filter(qs, kind=L.SGM_FILTER_NLF)
    _check_filter(g_cl, qs, ok, cnt, dsum, kind=L.SGM_FILTER_NLF)
    ok, cnt, dsum = b.filter(qs)  # back to GQL on the same slots
    _check_filter(g_cl, qs, ok, cnt, dsum)
    res = b.enumerate(method=L.SGM_ENUM_CACHE_FILTER, cap=100000)
    assert _check_enum(g_cl, qs, ok, res, None, 100000, fs=False) > 0


def test_batch_gql_and_caller_orders(g_cl):
    qs = _mix(g_cl, 12, seed=83, hi=16)
    b = sgm.Batch(g_cl, slots=12)
    ok, cnt, _ = b.filter(qs)
    orders = []
    for i, q in enumerate(qs):
        o = b.order_gql(i)
        assert o == O.order_gql(g_cl, q, [int(x) for x in cnt[i, :q.getVertexCnt()]])
        orders.append(o)
    orders[1] = None  # RI for this one
    res = b.enumerate(orders=orders, cap=100000)
    assert _check_enum(g_cl, qs, ok, res, orders, 100000) > 0
    with pytest.raises(L.SgmError):
        bad = list(orders)
        bad[0] = [0] * qs[0].getVertexCnt()  # not a permutation
        b.enumerate(orders=bad, cap=1000)


def test_batch_autotune_keeps_counts(g_cl):
    qs = _mix(g_cl, 16, seed=84, hi=14)
    b = sgm.Batch(g_cl, slots=16)
    ok, _, _ = b.filter(qs)
    base = b.enumerate(cap=50000)
    best, ms = b.autotune(cap=50000)
    say("autotune", best, ms)
    assert ms > 0 and best["blocks"] > 0
    assert best["split_chunks"] in (0, 8, 16) and best["stack_words"] >= 0
    tuned = b.enumerate(cap=50000, **best)
    assert [r.ans for r in tuned] == [r.ans for r in base]



def test_batch_rmat_mixed_sizes(g_rmat):
    qs = _mix(g_rmat, 24, seed=85)
    b = sgm.Batch(g_rmat, slots=24)
    ok, cnt, dsum = b.filter(qs)
    _check_filter(g_rmat, qs, ok, cnt, dsum)
    res = b.enumerate(cap=100000)
    assert _check_enum(g_rmat, qs, ok, res, None, 100000) > 0


def test_batch_overlapped_runners(g_cl):
    """Two runners of one context, pipelined: batch k+1 is filtered and launched while batch
    k enumerates (sgm_batch_enumerate_launch / _wait); every batch's results equal the
    oracle's, and a runner with a launch in flight refuses to load another batch."""
    qs = _mix(g_cl, 40, seed=86, hi=18)
    runners = [sgm.Batch(g_cl, slots=10), None]
    runners[1] = sgm.Batch(g_cl, slots=10, ctx=runners[0].ctx)
    parts = [qs[lo:lo + 10] for lo in range(0, len(qs), 10)]
    pending = None
    checked = 0
    for k, part in enumerate(parts):
        r = runners[k % 2]
        say("overlapped batch", k)
        ok, _, _ = r.filter(part)
        r.enumerate_start(cap=100000)
        if k == 0:
            with pytest.raises(L.SgmError):
                r.filter(part)  # launch in flight on this runner
        if pending is not None:
            pr, pok, ppart = pending
            checked += _check_enum(g_cl, ppart, pok, pr.enumerate_wait(), None, 100000)
        pending = (r, ok, part)
    pr, pok, ppart = pending
    checked += _check_enum(g_cl, ppart, pok, pr.enumerate_wait(), None, 100000)
    assert checked > 0
    for r in runners:
        r.close()
