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
    for q in _mixed(g_small, 4, 104, 4, 9):
        say("tables", q.name)
        can = sgm.CandidateParam(g_small, q)
        if not sgm.Filter_GQL(g_small, q, can):
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
    for fs, method, flags in ((True, L.SGM_ENUM_CACHE_FILTER_FS, 0), (False, L.SGM_ENUM_CACHE_FILTER, 0),
                              (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_NO_BITMAPS),
                              (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SERIAL_LAST),
                              (False, L.SGM_ENUM_CACHE_FILTER, L.SGM_ENUM_F_SERIAL_LAST),
                              (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_NO_SPLIT),
                              (True, L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_F_SPLIT_ALL),
                              (False, L.SGM_ENUM_CACHE_FILTER, L.SGM_ENUM_F_SPLIT_ALL)):
        res = sgm.enumerate_query(can, order, method=method, cap=cap, flags=flags)
        ref = O.enum_cache(g, q, c_r, order, cap=cap, fs=fs)
        assert res.ans == ref["count"], (q.name, fs, flags, res.ans, ref["count"])
        if check_work and not ref["capped"]:
            assert res.intersections == ref["intersections"], (q.name, fs, flags)
            assert res.nodes == ref["nodes"], (q.name, fs, flags, res.nodes, ref["nodes"])
            assert res.fs_prunes == ref["fs_prunes"], (q.name, fs, flags)
            # the chunked and the one-at-a-time last level read the same elements
            key = (fs, flags & ~L.SGM_ENUM_F_SERIAL_LAST & ~L.SGM_ENUM_F_NO_BITMAPS & ~L.SGM_ENUM_F_NO_SPLIT
                   & ~L.SGM_ENUM_F_SPLIT_ALL)
            got = (res.elems_read, res.elems_out, res.elems_scanned)
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


def test_enum_capped_config1(g1):
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
    k = g_small.getMaxDgree() + 1
    if k > 31:
        pytest.skip("max degree too large for a star query")
    q = query_from_edges(k + 1, [(0, i) for i in range(1, k + 1)])
    can = sgm.CandidateParam(g_small, q)
    assert not sgm.Filter_GQL(g_small, q, can, build_tables=False)
    assert sgm.enumerate_query(can, sgm.Order_RI(q)).ans == 0


def test_order_must_be_connected(g_small):
    q = query_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    can = sgm.CandidateParam(g_small, q)
    sgm.Filter_GQL(g_small, q, can, build_tables=False)
    with pytest.raises(L.SgmError) as e:
        sgm.enumerate_query(can, [0, 2, 1, 3])
    assert e.value.code == L.SGM_EORDER
