"""GPU parity at the benchmark workloads' own sizes (Chung-Lu 10^4 and 10^5 graphs with
their random-walk queries).

Whole uncapped searches on these inputs run for hours on any hardware (the 8-vertex Chung-Lu 10^4 query
has more than 1.5e13 embeddings; every dense 16-vertex Chung-Lu 10^5 query more than 1e9), so enumeration parity
is shown on root-range slices (tests/slices.py): exact, uncapped sub-searches of the whole
search whose oracle statistics are in tests/golden/config_slices.json (written by
tests/golden/make_config_slices.py), and for one slice per workload recomputed live.  Every
slice compares count, per-depth nodes, intersections and failing-set cut-offs on the cache
search and the aux-table search, with failing sets on and off.  Candidate sets are compared
bit for bit (the Chung-Lu 10^4 query live, the 200 Chung-Lu 10^5 queries by the SHA-256 of their sets)."""
import json
import os
import time

import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from slices import VARIANTS, cand_digest, gpu_stats, oracle_stats
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_slices.json")
METHOD = {"cache_fs": L.SGM_ENUM_CACHE_FILTER_FS, "cache": L.SGM_ENUM_CACHE_FILTER,
          "aux": L.SGM_ENUM_AUX_TABLES, "aux_fs": L.SGM_ENUM_AUX_TABLES_FS}
VARIANTS_ALL = tuple(v[0] for v in VARIANTS)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def fix():
    with open(FIX) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def cl10k_in():
    g = chung_lu(10000, 2.5, 7.4, seed=1)
    return g, query_set(g, 1, 8, False, seed=11)[0]


@pytest.fixture(scope="module")
def cl100k_in():
    g = chung_lu(100_000, 2.3, 16.0, seed=2)
    qs = query_set(g, 200, 16, True, seed=22)
    cans = [sgm.CandidateParam(g, q) for q in qs]
    ok = sgm.filter_batch(cans)
    yield g, qs, cans, ok
    for c in cans:
        c.close()


def _check_ranges(can, order, ranges, tag, variants=None):
    for rec in ranges:
        r = tuple(rec["roots"])
        for name, _, _ in VARIANTS:
            if variants is not None and name not in variants(rec):
                continue
            t0 = time.perf_counter()
            res = sgm.enumerate_query(can, order, METHOD[name], roots=r)
            say(tag, r, name, res.ans, f"{time.perf_counter() - t0:.2f}s")
            assert not res.timed_out and not res.capped
            assert gpu_stats(res) == rec[name], (tag, r, name)


def test_cl10k_filter_order_and_slices(cl10k_in, fix):
    g, q = cl10k_in
    F = fix["cl10k"]
    can = sgm.CandidateParam(g, q)
    assert sgm.Filter_GQL(g, q, can, build_tables=False)
    ok_r, c_r = O.gql(g, q)
    assert ok_r
    for u in range(q.getVertexCnt()):
        assert np.array_equal(can.candidates[u], c_r[u]), u
    assert cand_digest(c_r) == F["cand_digest"]
    order = sgm.Order_RI(q)
    assert order == O.order_ri(q) == F["order"]
    assert len(F["ranges"]) >= 3
    # these roots have fewer than 64 neighbours, so depth-1 splitting leaves each slice's
    # subtree (5e7 to 3e8 embeddings) to one wave (about 6e6 embeddings/s alone): the
    # smallest slice runs every variant, the others one each, to keep the test near a minute
    by_size = sorted(F["ranges"], key=lambda x: x["cache_fs"]["count"])
    plan = {id(by_size[0]): VARIANTS_ALL, id(by_size[1]): ("cache_fs", "aux_fs"), id(by_size[2]): ("cache",)}
    _check_ranges(can, order, F["ranges"], "cl10k", variants=lambda rec: plan[id(rec)])
    # one slice recomputed by the oracle now (the fixture is the oracle's own output)
    r = tuple(F["ranges"][0]["roots"])
    live = oracle_stats(O, g, q, c_r, order, "cache_fs", r)
    assert live == F["ranges"][0]["cache_fs"]
    can.close()


def test_cl100k_candidate_sets_all_queries(cl100k_in, fix):
    """All 200 queries' GQL candidate sets (one filter batch) against the oracle's digests."""
    g, qs, cans, ok = cl100k_in
    F = fix["cl100k"]
    assert ok == F["ok"]
    for i, can in enumerate(cans):
        assert [int(x) for x in can.candidates_count] == F["cand_counts"][i], i
        assert cand_digest(can.candidates) == F["cand_digest"][i], i
    say("cl100k digests", len(cans))
    # two queries recomputed by the oracle now
    for i in (0, 199):
        ok_r, c_r = O.gql(g, qs[i])
        assert cand_digest(c_r) == F["cand_digest"][i]
        assert sgm.Order_RI(qs[i]) == O.order_ri(qs[i])


def test_cl100k_slices(cl100k_in, fix):
    g, qs, cans, ok = cl100k_in
    F = fix["cl100k"]["slices"]
    assert len(F) >= 8
    for key, rec in sorted(F.items(), key=lambda kv: int(kv[0])):
        i = int(key)
        assert rec["order"] == sgm.Order_RI(qs[i])
        assert rec["ranges"], i
        _check_ranges(cans[i], rec["order"], rec["ranges"], f"cl100k q{i}")
    # one slice recomputed by the oracle now, from the GPU's (digest-checked) candidates
    key = sorted(F, key=int)[0]
    i, rec = int(key), F[key]["ranges"][0]
    live = oracle_stats(O, g, qs[i], cans[i].candidates, F[key]["order"], "cache_fs", tuple(rec["roots"]))
    assert live == rec["cache_fs"]


def test_cl100k_slices_in_one_batch(cl100k_in, fix):
    """The same slices from one persistent launch over the eight queries (per-query stats)."""
    g, qs, cans, ok = cl100k_in
    F = fix["cl100k"]["slices"]
    keys = sorted(F, key=int)
    # one range per launch (the range applies to every query of the launch): use each
    # query's first range, one launch per distinct range, all eight queries in each; the
    # other queries' searches of that root may be huge, so each query gets a budget of 3 s
    # of wave time (the slice's own query needs a few milliseconds)
    for key in keys:
        r = tuple(F[key]["ranges"][0]["roots"])
        say("one batch, root", r)
        res = sgm.enumerate_batch([cans[int(k)] for k in keys], [F[k]["order"] for k in keys], roots=r,
                                  time_limit_ms=3_000)
        mine = res[keys.index(key)]
        assert not mine.timed_out, key
        assert gpu_stats(mine) == F[key]["ranges"][0]["cache_fs"], key


def test_cl100k_capped_counts_all_queries(cl100k_in):
    """Every dense 16-vertex Chung-Lu 10^5 query holds more than 10^9 embeddings, so count = min(total, cap) is
    the cap itself for all 200 in one launch; the oracle agrees on two of them."""
    g, qs, cans, ok = cl100k_in
    cap = 1_000_000
    res = sgm.enumerate_batch(cans, None, L.SGM_ENUM_CACHE_FILTER_FS, cap=cap)
    assert all(r.ans == cap and r.capped for r in res)
    for i in (3, 150):
        ref = O.enum_cache(g, qs[i], cans[i].candidates, O.order_ri(qs[i]), cap=cap)
        assert ref["count"] == cap


def test_cl100k_uncapped_budget_times_out(cl100k_in):
    """The dense 16-vertex Chung-Lu 10^5 workload as specified (no cap): with a per-query budget every query times out (their
    counts exceed what the budget can count) and is reported timed out, never as a result;
    a cap ends each query exactly at the cap instead."""
    g, qs, cans, ok = cl100k_in
    sub = cans[:16]
    res = sgm.enumerate_batch(sub, None, L.SGM_ENUM_CACHE_FILTER_FS, time_limit_ms=2000)
    assert all(r.timed_out for r in res)
    assert all(r.busy_ns >= 2e9 for r in res)
    res = sgm.enumerate_batch(sub, None, L.SGM_ENUM_CACHE_FILTER_FS, cap=1_000_000, time_limit_ms=60_000)
    assert all(r.capped and not r.timed_out and r.ans == 1_000_000 for r in res)
