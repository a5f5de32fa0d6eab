"""GPU parity: the gfx950 path (through the C ABI) against the CPU oracle on the same
seeded inputs.  Candidate sets, tables, orders and counts must be bit-exact; with
failing sets on or off, intersection and per-depth node counts must also be exact (one
wave per depth-0 candidate reproduces the reference's search of each subtree)."""
import numpy as np
import pytest

import oracle as O
from brute import count_embeddings, query_from_edges, random_connected_query, random_graph
import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def g1():
    return chung_lu(10000, 2.5, 7.4, seed=1)


@pytest.fixture(scope="module")
def g_small():
    return chung_lu(2000, 2.3, 8.0, seed=3)


def _mixed(g, count, seed, lo=4, hi=17):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


def test_nlf_parity(g1):
    for q in _mixed(g1, 6, 101):
        say("nlf", q.name)
        can = sgm.CandidateParam(g1, q)
        ok = sgm.Filter_NLF(g1, q, can, build_tables=False)
        ok_r, c_r = O.nlf(g1, q)
        assert ok == ok_r
        for u in range(q.getVertexCnt()):
            assert np.array_equal(can.candidates[u], c_r[u])


def test_gql_parity(g1):
    for q in _mixed(g1, 12, 102) + query_set(g1, 3, 8, False, seed=11):
        say("gql", q.name)
        can = sgm.CandidateParam(g1, q)
        ok = sgm.Filter_GQL(g1, q, can, build_tables=False)
        ok_r, c_r = O.gql(g1, q)
        assert ok == ok_r
        assert [int(x) for x in can.candidates_count] == [len(x) for x in c_r]
        for u in range(q.getVertexCnt()):
            assert np.array_equal(can.candidates[u], c_r[u]), (q.name, u)
        ex = can.candidates_exsist
        for u in range(q.getVertexCnt()):
            assert np.array_equal(np.flatnonzero(ex[u]), c_r[u])


def test_gql_batch_equals_single(g1):
    qs = _mixed(g1, 10, 103, 4, 33)
    say("gql batch", len(qs))
    cans = [sgm.CandidateParam(g1, q) for q in qs]
    oks = sgm.filter_batch(cans)
    for q, can, ok in zip(qs, cans, oks):
        ok_r, c_r = O.gql(g1, q)
        assert ok == ok_r
        for u in range(q.getVertexCnt()):
            assert np.array_equal(can.candidates[u], c_r[u])


def test_gql_parity_power_law_100k():
    g = chung_lu(100000, 2.3, 16.0, seed=2)
    for q in query_set(g, 3, 16, True, seed=22):
        say("gql 100k", q.name)
        can = sgm.CandidateParam(g, q)
        ok = sgm.Filter_GQL(g, q, can, build_tables=False)
        ok_r, c_r = O.gql(g, q)
        assert ok == ok_r
        for u in range(q.getVertexCnt()):
            assert np.array_equal(can.candidates[u], c_r[u])


def test_tables_parity(g_small):
    # plus a star and a fan of 20 leaves: a query vertex with 20 out-edges drives every
    # lane of the grouped count pass (one lane per out-edge, up to 31)
    star = query_from_edges(21, [(0, i) for i in range(1, 21)])
    fan = query_from_edges(21, [(0, i) for i in range(1, 21)] + [(i, i + 1) for i in range(1, 20)])
    for q in _mixed(g_small, 4, 104, 4, 9) + [star, fan]:
        say("tables", q.name, q.getVertexCnt())
        can = sgm.CandidateParam(g_small, q)
        if not sgm.Filter_GQL(g_small, q, can):
            assert q.getVertexCnt() < 21, "the star and the fan must pass the filter"
            continue
        c = can.candidates
        for a in range(q.getVertexCnt()):
            for b in q.getNeb(a).tolist():
                off, idx, vv, md = O.table(g_small, c[a], c[b])
                e = can.edge_matrix[a][b]
                assert np.array_equal(e._offset, off)
                assert np.array_equal(e._edge, idx)
                assert np.array_equal(e._edge_v, vv)
                assert e._max_degree == md


def test_orders_parity(g1):
    for q in _mixed(g1, 8, 105):
        can = sgm.CandidateParam(g1, q)
        sgm.Filter_GQL(g1, q, can, build_tables=False)
        assert sgm.Order_RI(q) == O.order_ri(q)
        assert sgm.Order_GQL(g1, q, can) == O.order_gql(g1, q, [int(x) for x in can.candidates_count])


def _compare(g, q, order, cap=None, check_work=True):
    can = sgm.CandidateParam(g, q)
    ok = sgm.Filter_GQL(g, q, can, build_tables=False)
    ok_r, c_r = O.gql(g, q)
    assert ok == ok_r
    elems = {}
    # (failing sets, method, flags, work chunks per split root: 0 = 32)
    for fs, method, flags, chunks in ((True, L.SGM_ENUM_CACHE_FILTER_FS, 0, 0), (False, L.SGM_ENUM_CACHE_FILTER, 0, 0),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_NO_BITMAPS, 0),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SERIAL_LAST, 0),
                                      (False, L.SGM_ENUM_CACHE_FILTER, L.SGM_ENUM_F_SERIAL_LAST, 0),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_NO_SPLIT, 0),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SPLIT_ALL, 0),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SPLIT_ALL, 1),
                                      (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SPLIT_ALL, 8),
                                      (False, L.SGM_ENUM_CACHE_FILTER, L.SGM_ENUM_F_SPLIT_ALL, 0)):
        res = sgm.enumerate_query(can, order, method=method, cap=cap, flags=flags, split_chunks=chunks)
        ref = O.enum_cache(g, q, c_r, order, cap=cap, fs=fs)
        assert res.ans == ref["count"], (q.name, fs, flags, res.ans, ref["count"])
        if check_work and not ref["capped"]:
            assert res.intersections == ref["intersections"], (q.name, fs, flags)
            assert res.nodes == ref["nodes"], (q.name, fs, flags, res.nodes, ref["nodes"])
            assert res.fs_prunes == ref["fs_prunes"], (q.name, fs, flags)
            # the chunked and the one-at-a-time last level read the same elements
            key = (fs, flags & ~L.SGM_ENUM_F_SERIAL_LAST & ~L.SGM_ENUM_F_NO_SPLIT & ~L.SGM_ENUM_F_SPLIT_ALL)
            got = (res.elems_read, res.elems_out, res.elems_scanned)
            if not flags & L.SGM_ENUM_F_NO_BITMAPS:  # binary search instead of bitmaps reads other words
                got = got + (res.probe_words,)
            assert elems.setdefault(key, got) == got, (q.name, fs, flags, elems[key], got)
    for fs, method in ((False, L.SGM_ENUM_AUX_TABLES), (True, L.SGM_ENUM_AUX_TABLES_FS)):
        res_a = sgm.enumerate_query(can, order, method=method, cap=cap)
        ref_a = O.enum_all(g, q, c_r, order, cap=cap, fs=fs)
        assert res_a.ans == ref_a["count"], (q.name, "aux", fs, res_a.ans, ref_a["count"])
        if check_work and not ref_a["capped"]:
            assert res_a.intersections == ref_a["intersections"], (q.name, "aux", fs)
            assert res_a.nodes == ref_a["nodes"], (q.name, "aux", fs, res_a.nodes, ref_a["nodes"])
            assert res_a.fs_prunes == ref_a["fs_prunes"], (q.name, "aux", fs)
    return res


def test_enum_bruteforce_tiny():
    rng = np.random.default_rng(106)
    for i in range(40):
        g = random_graph(int(rng.integers(7, 11)), float(rng.uniform(0.3, 0.6)), rng)
        q = random_connected_query(int(rng.integers(2, 6)), int(rng.integers(0, 3)), rng)
        exp = count_embeddings(g, q)
        say("tiny", i, q.getVertexCnt(), exp)
        can = sgm.CandidateParam(g, q)
        ok = sgm.Filter_GQL(g, q, can, build_tables=False)
        if not ok:
            assert exp == 0
            continue
        for order in (sgm.Order_RI(q), sgm.Order_GQL(g, q, can)):
            _compare(g, q, order)
            assert sgm.enumerate_query(can, order).ans == exp


def test_enum_parity_uncapped(g_small):
    qs = _mixed(g_small, 16, 107, 3, 8)
    for q in qs:
        can = sgm.CandidateParam(g_small, q)
        ok_r, c_r = O.gql(g_small, q)
        if not ok_r:
            continue
        order = O.order_ri(q)
        pre = O.enum_cache(g_small, q, c_r, order, cap=3_000_000)
        if pre["capped"]:
            continue
        say("uncapped", q.name, pre["count"])
        _compare(g_small, q, order)


def test_root_ranges_add_up_on_gpu(g_small):
    """Disjoint root ranges (one launch each) add up to the whole search in count, nodes,
    intersections and failing-set cut-offs, on both searches with and without failing sets,
    and every range equals the oracle's."""
    rng = np.random.default_rng(110)
    V = g_small.getVertexCnt()
    for q in _mixed(g_small, 6, 111, 4, 8):
        can = sgm.CandidateParam(g_small, q)
        ok_r, c_r = O.gql(g_small, q)
        if not sgm.Filter_GQL(g_small, q, can, build_tables=False):
            continue
        order = O.order_ri(q)
        cuts = sorted(int(x) for x in rng.choice(np.arange(1, V), 3, replace=False))
        ranges = list(zip([0] + cuts, cuts + [V]))
        for method, fs, fn in ((L.SGM_ENUM_CACHE_FILTER_FS, True, O.enum_cache),
                               (L.SGM_ENUM_CACHE_FILTER, False, O.enum_cache),
                               (L.SGM_ENUM_AUX_TABLES, False, O.enum_all),
                               (L.SGM_ENUM_AUX_TABLES_FS, True, O.enum_all)):
            whole = sgm.enumerate_query(can, order, method, cap=3_000_000)
            if whole.capped:
                continue
            parts = [sgm.enumerate_query(can, order, method, roots=r) for r in ranges]
            say("ranges", q.name, method, whole.ans, [p.ans for p in parts])
            assert sum(p.ans for p in parts) == whole.ans
            assert sum(p.intersections for p in parts) == whole.intersections
            assert sum(p.fs_prunes for p in parts) == whole.fs_prunes
            assert [sum(x) for x in zip(*[p.nodes for p in parts])] == whole.nodes
            for r, p in zip(ranges, parts):
                ref = fn(g_small, q, c_r, order, fs=fs, roots=r)
                assert (p.ans, p.nodes, p.intersections, p.fs_prunes) == \
                    (ref["count"], ref["nodes"], ref["intersections"], ref["fs_prunes"]), (q.name, method, r)


def test_time_budget_and_host_stop(g1):
    """A heavy uncapped query: a small work budget ends it timed out (never a result); a host
    stop request ends it within a moment, with a larger budget as the backstop."""
    import threading
    import time
    q = query_set(g1, 1, 8, False, seed=11)[0]  # more than 1e13 embeddings
    can = sgm.CandidateParam(g1, q)
    assert sgm.Filter_GQL(g1, q, can, build_tables=False)
    order = sgm.Order_RI(q)
    r = sgm.enumerate_query(can, order, time_limit_ms=500)
    assert r.timed_out and not r.capped and r.busy_ns >= 5e8
    # 16 blocks (64 waves) with an hour of wave time each would run about a minute
    timer = threading.Timer(0.3, can.ctx.stop)
    timer.start()
    t0 = time.time()
    with pytest.raises(L.SgmError) as e:
        sgm.enumerate_query(can, order, blocks=16, time_limit_ms=3_600_000)
    timer.cancel()
    assert e.value.code == L.SGM_ESTOPPED
    assert time.time() - t0 < 10.0
    # the next launch is not affected by the earlier request
    r = sgm.enumerate_query(can, order, cap=1000)
    assert r.ans == 1000 and not r.stopped


def test_enum_capped_cl10k(g1):
    for q in query_set(g1, 2, 8, False, seed=11) + query_set(g1, 2, 16, True, seed=12):
        say("capped", q.name)
        res = _compare(g1, q, O.order_ri(q), cap=1_000_000, check_work=False)
        assert res.ans <= 1_000_000


def test_enum_batch_equals_oracle(g_small):
    """One persistent launch over many queries: per-query counts (capped and uncapped) and,
    for uncapped queries, intersections / nodes / prunes equal the oracle's."""
    qs = _mixed(g_small, 24, 108, 3, 12)
    cans = [sgm.CandidateParam(g_small, q) for q in qs]
    oks = sgm.filter_batch(cans)
    orders = [sgm.Order_RI(q) for q in qs]
    for cap in (None, 50_000):
        say("batch cap", cap)
        refs = []
        for q, can, order in zip(qs, cans, orders):
            ok_r, c_r = O.gql(g_small, q)
            ref = O.enum_cache(g_small, q, c_r, order, cap=cap if cap else 3_000_000)
            refs.append(ref)
        keep = [i for i in range(len(qs)) if cap or not refs[i]["capped"]]
        res = sgm.enumerate_batch([cans[i] for i in keep], [orders[i] for i in keep], cap=cap)
        if not cap:
            # the aux-table search over the same queries in one launch (plain and failing sets)
            for fs, method in ((False, L.SGM_ENUM_AUX_TABLES), (True, L.SGM_ENUM_AUX_TABLES_FS)):
                ra = sgm.enumerate_batch([cans[i] for i in keep], [orders[i] for i in keep], method=method)
                for r, i in zip(ra, keep):
                    ok_r, c_r = O.gql(g_small, qs[i])
                    ref_a = O.enum_all(g_small, qs[i], c_r, orders[i], fs=fs)
                    assert (r.ans, r.nodes, r.intersections, r.fs_prunes) == \
                        (ref_a["count"], ref_a["nodes"], ref_a["intersections"], ref_a["fs_prunes"]), (qs[i].name, fs)
        if not cap:  # orders=None: the library's RI orders, the same search
            res_ri = sgm.enumerate_batch([cans[i] for i in keep], None, cap=cap)
            assert [(r.ans, r.nodes, r.intersections) for r in res_ri] == [(r.ans, r.nodes, r.intersections) for r in res]
        for r, i in zip(res, keep):
            assert r.ans == refs[i]["count"], (qs[i].name, cap, r.ans, refs[i]["count"])
            if not cap:
                assert r.intersections == refs[i]["intersections"], qs[i].name
                assert r.nodes == refs[i]["nodes"], qs[i].name
                assert r.fs_prunes == refs[i]["fs_prunes"], qs[i].name


def test_single_edge_and_star_queries(g_small):
    deg = g_small.degrees().astype(np.int64)
    cases = [(query_from_edges(2, [(0, 1)]), int(deg.sum())),
             (query_from_edges(3, [(0, 1), (0, 2)]), int((deg * (deg - 1)).sum())),
             (query_from_edges(4, [(0, 1), (0, 2), (0, 3)]), int((deg * (deg - 1) * (deg - 2)).sum()))]
    for q, exp in cases:
        say("closed form", q.getVertexCnt())
        can = sgm.CandidateParam(g_small, q)
        assert sgm.Filter_GQL(g_small, q, can, build_tables=False)
        assert sgm.enumerate_query(can, sgm.Order_RI(q)).ans == exp


def test_empty_candidates(g_small):
    """A query vertex whose label no data vertex carries has no candidates: the filter
    returns false (inc/filter.h:87-90) and every enumeration entry point reports 0 without
    searching (src/overall_performance.cpp:54-56), the batched ones next to live queries."""
    q = query_from_edges(3, [(0, 1), (1, 2), (0, 2)], labels=[0, 0, 7])
    live = query_set(g_small, 2, 4, True, seed=109)
    for kind in (L.SGM_FILTER_NLF, L.SGM_FILTER_GQL):
        can = sgm.CandidateParam(g_small, q)
        filt = sgm.Filter_NLF if kind == L.SGM_FILTER_NLF else sgm.Filter_GQL
        assert not filt(g_small, q, can, build_tables=False)
        assert not (O.nlf if kind == L.SGM_FILTER_NLF else O.gql)(g_small, q)[0]
        order = sgm.Order_RI(q)
        for method in (L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_CACHE_FILTER, L.SGM_ENUM_AUX_TABLES,
                       L.SGM_ENUM_AUX_TABLES_FS):
            assert sgm.enumerate_query(can, order, method=method).ans == 0
        cans = [sgm.CandidateParam(g_small, x) for x in live]
        oks = sgm.filter_batch(cans + [can], kind)
        assert oks[-1] is False and all(oks[:-1])
        for method in (L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_AUX_TABLES, L.SGM_ENUM_AUX_TABLES_FS):
            res = sgm.enumerate_batch(cans + [can], None, method, cap=10000)
            assert res[-1].ans == 0
            assert all(r.ans == 10000 for r in res[:-1])  # the live queries have many embeddings
    b = sgm.Batch(g_small, slots=4)
    ok, cnt, _ = b.filter(live + [q])
    assert ok == [True, True, False]
    res = b.enumerate(cap=10000)
    assert res[-1].ans == 0 and all(r.ans == 10000 for r in res[:-1])


def test_order_must_be_connected(g_small):
    q = query_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    can = sgm.CandidateParam(g_small, q)
    sgm.Filter_GQL(g_small, q, can, build_tables=False)
    with pytest.raises(L.SgmError) as e:
        sgm.enumerate_query(can, [0, 2, 1, 3])
    assert e.value.code == L.SGM_EORDER
