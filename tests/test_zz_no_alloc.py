"""The no-allocation entry point (sgm_enumerate_noalloc, through the C ABI) against the
allocating one (sgm_enumerate): identical counts, per-depth nodes, intersections,
failing-set cut-offs and element counters, on two calls in a row of which the second has
the smaller input (the static scratch array keeps the first call's contents), with and
without failing sets, capped and uncapped, a root range and depth-1 splitting; and its
rejections (aux methods, a stack request larger than the array)."""
import ctypes as C

import pytest

import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.api import _opts
from subgraphmatching_amd.generators import chung_lu, query_set

pytestmark = pytest.mark.gpu

FIELDS = ("count", "capped", "timed_out", "intersections", "fs_prunes", "elems_read", "elems_out", "elems_scanned")


def say(*a):
    print(*a, flush=True)


def _call(fn, can, order, opts):
    st = L.Stats()
    o = L.u32(order)
    rc = fn(can._q, L.ptr(o), C.byref(opts), C.byref(st))
    assert rc == L.SGM_OK, (rc, L.lib().sgm_ctx_last_error(can.ctx.handle))
    n = can.query.getVertexCnt()
    return {**{f: int(getattr(st, f)) for f in FIELDS}, "nodes": [int(x) for x in st.nodes[:n]]}


def _mixed(g, count, lo, hi, seed):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


def _inputs():
    """Queries whose whole searches finish quickly (a few 10^5 to 10^6 embeddings; the small
    graph's query 11 has failing-set cut-offs): first on a 2,000-vertex graph, then on a
    300-vertex one."""
    big_g = chung_lu(2000, 2.3, 8.0, seed=3)
    small_g = chung_lu(300, 2.3, 5.0, seed=19)
    big = [q for i, q in enumerate(_mixed(big_g, 12, 3, 8, 107)) if i in (0, 7, 11)]
    small = [q for i, q in enumerate(_mixed(small_g, 14, 5, 9, 20)) if i in (1, 3, 11, 13)]
    return (big_g, big), (small_g, small)


def test_noalloc_matches_enumerate():
    (gb, qb), (gs, qs) = _inputs()
    cases = []
    for g, queries in ((gb, qb), (gs, qs)):
        for q in queries:
            can = sgm.CandidateParam(g, q)
            assert sgm.Filter_GQL(g, q, can, build_tables=False)
            cases.append((can, sgm.Order_RI(q)))
    assert len(cases) == 7
    V = gb.getVertexCnt()
    settings = [dict(method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None), dict(method=L.SGM_ENUM_CACHE_FILTER, cap=None),
                dict(method=L.SGM_ENUM_CACHE_FILTER_FS, cap=50_000),
                dict(method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, flags=L.SGM_ENUM_F_SPLIT_ALL),
                dict(method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, roots=(V // 3, 2 * V // 3))]
    for kw in settings:
        opts = _opts(kw["method"], kw["cap"], flags=kw.get("flags", 0), roots=kw.get("roots"))
        # references from the allocating entry point first, then the no-allocation one on
        # every input in a row: big inputs, then the smaller ones
        ref = [_call(L.lib().sgm_enumerate, can, order, opts) for can, order in cases]
        got = [_call(L.lib().sgm_enumerate_noalloc, can, order, opts) for can, order in cases]
        for (can, _), r, x in zip(cases, ref, got):
            say("noalloc", can.query.name, kw, r["count"], x["count"])
            if kw["cap"] is None:
                assert x == r, can.query.name
            else:  # capped: the count is min(total, cap); the rest depends on the schedule
                assert (x["count"], x["capped"]) == (r["count"], r["capped"]), can.query.name
    for can, _ in cases:
        can.close()


def test_noalloc_rejections():
    g = chung_lu(300, 2.5, 5.0, seed=4)
    q = query_set(g, 1, 4, True, seed=303)[0]  # 2012 embeddings
    can = sgm.CandidateParam(g, q)
    assert sgm.Filter_GQL(g, q, can, build_tables=False)
    o = L.u32(sgm.Order_RI(q))
    st = L.Stats()
    for opts in (_opts(L.SGM_ENUM_AUX_TABLES, None), _opts(L.SGM_ENUM_CACHE_FILTER_FS, None, stack_words=1 << 28),
                 _opts(L.SGM_ENUM_CACHE_FILTER_FS, None, hub_min_degree=1024)):
        assert L.lib().sgm_enumerate_noalloc(can._q, L.ptr(o), C.byref(opts), C.byref(st)) == L.SGM_EINVAL
    # and a valid call afterwards still works
    ok = _opts(L.SGM_ENUM_CACHE_FILTER_FS, None)
    assert L.lib().sgm_enumerate_noalloc(can._q, L.ptr(o), C.byref(ok), C.byref(st)) == L.SGM_OK
    ref = sgm.enumerate_query(can, sgm.Order_RI(q))
    assert int(st.count) == ref.ans
    can.close()
