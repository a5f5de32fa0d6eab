"""GPU parity of the dynamic-order searches (explore_Intersection_DynamicOrder_DAF / _VEQ:
k_enum_dyn) and of the DAF weight array (computeWeightArray: k_weight_rows) against the
oracle on the same seeded inputs.  Uncapped searches match the oracle's count,
intersections and extensions per depth exactly (the chosen vertex at every node, ties
included, decides the last two); capped searches stop at the cap.  Start orders are the
reference driver's: BFS from DAFStartVertex."""
import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from brute import random_connected_query, random_graph
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels

pytestmark = pytest.mark.gpu

METHOD = {"DAF": L.SGM_ENUM_DYN_DAF, "VEQ": L.SGM_ENUM_DYN_VEQ}
ORACLE_CAP = 3_000_000  # searches the oracle finishes in about a second


def say(*a):
    print(*a, flush=True)


def mixed(g, count, seed, lo=4, hi=13):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


@pytest.fixture(scope="module")
def g2k():
    """Labelled power-law graph: most searches of 4..12-vertex queries end below a few
    million embeddings, so the oracle finishes them."""
    return with_labels(chung_lu(2000, 2.4, 6.0, seed=71), 4, seed=75)


@pytest.fixture(scope="module")
def plain():
    return chung_lu(2000, 2.4, 6.0, seed=71)


def _daf(g, q):
    can = sgm.CandidateParam(g, q)
    ok = sgm.Filter_DAF(g, q, can)
    ok_r, c_r = O.filter_tree("DAF", g, q)
    assert ok == ok_r, q.name
    return can, ok, c_r


def _oracle(g, q, c_r, kind, cap=None):
    bfs, parent, _ = O.dyn_bfs(g, q)
    w = O.daf_weights(g, q, c_r, bfs, parent) if kind == "DAF" else None
    return bfs, O.enum_dynamic(g, q, c_r, kind, cap=cap, bfs_order=bfs, weights=w)


def _same(r, ref, tag):
    assert (r.ans, r.intersections, r.nodes) == (ref["count"], ref["intersections"], ref["nodes"]), tag


def test_start_vertex_and_weights(g2k):
    """DAFStartVertex, the BFS order computeWeightArray returns and every weight."""
    seen = 0
    for q in mixed(g2k, 12, 711, 3, 17):
        can, ok, c_r = _daf(g2k, q)
        if ok:
            bfs_r, parent, _ = O.dyn_bfs(g2k, q)
            assert sgm.DAFStartVertex(g2k, q, can) == O.start_vertex("DAF", g2k, q), q.name
            bfs = []
            w = sgm.computeWeightArray(g2k, q, can, bfs)
            assert bfs == [int(x) for x in bfs_r], q.name
            w_r = O.daf_weights(g2k, q, c_r, bfs_r, parent)
            for u in range(q.getVertexCnt()):
                assert np.array_equal(w[u], w_r[u]), (q.name, u)
            seen += 1
        can.close()
    assert seen >= 6


@pytest.mark.parametrize("kind", ["DAF", "VEQ"])
def test_uncapped_parity(g2k, kind):
    checked = 0
    for q in mixed(g2k, 16, 720 + len(kind), 4, 11):
        can, ok, c_r = _daf(g2k, q)
        if not ok:
            can.close()
            continue
        bfs, ref = _oracle(g2k, q, c_r, kind, cap=ORACLE_CAP)
        if ref["capped"]:
            can.close()
            continue
        say(kind, q.name, ref["count"])
        r = sgm.enumerate_query(can, bfs, METHOD[kind])
        _same(r, ref, (kind, q.name))
        assert not r.capped and r.retries == 0
        # hub rows intersected by binary search instead of their adjacency bitmaps
        _same(sgm.enumerate_query(can, bfs, METHOD[kind], flags=L.SGM_ENUM_F_NO_BITMAPS), ref, (kind, q.name, "nb"))
        checked += 1
        can.close()
    assert checked >= 6


@pytest.mark.parametrize("kind", ["DAF", "VEQ"])
def test_labelled_and_small_graphs(kind):
    """Labelled power-law graphs and dense random graphs (many ties in the queue)."""
    checked = 0
    g = with_labels(chung_lu(2000, 2.4, 8.0, seed=72), 8, seed=73)
    inputs = [(g, q) for q in mixed(g, 8, 730, 4, 9)]
    rng = np.random.default_rng(74)
    for _ in range(30):
        gg = random_graph(int(rng.integers(10, 16)), float(rng.uniform(0.3, 0.6)), rng)
        inputs.append((gg, random_connected_query(int(rng.integers(3, 7)), int(rng.integers(0, 4)), rng)))
    for gg, q in inputs:
        can, ok, c_r = _daf(gg, q)
        if ok:
            bfs, ref = _oracle(gg, q, c_r, kind, cap=ORACLE_CAP)
            if not ref["capped"]:
                r = sgm.enumerate_query(can, bfs, METHOD[kind])
                _same(r, ref, (kind, q.name))
                checked += 1
        can.close()
    say(kind, "checked", checked)
    assert checked >= 20


@pytest.mark.parametrize("kind", ["DAF", "VEQ"])
def test_capped(plain, kind):
    """A cap ends the search with count == cap (the reference stops at ans == output_limit)."""
    seen = 0
    for q in mixed(plain, 16, 740, 5, 13):
        can, ok, c_r = _daf(plain, q)
        if ok:
            bfs, ref = _oracle(plain, q, c_r, kind, cap=1000)
            r = sgm.enumerate_query(can, bfs, METHOD[kind], cap=1000)
            assert r.ans == ref["count"] and r.capped == ref["capped"], q.name
            seen += ref["capped"]
        can.close()
    assert seen >= 3


def test_batch_default_orders_and_root_ranges(g2k):
    """Many queries in one launch with the default start orders (orders None), and disjoint
    root ranges that add up to the whole search."""
    qs = mixed(g2k, 20, 750, 4, 10)
    cans, refs = [], []
    for q in qs:
        can, ok, c_r = _daf(g2k, q)
        if not ok:
            can.close()
            continue
        _, ref = _oracle(g2k, q, c_r, "DAF", cap=ORACLE_CAP)
        if ref["capped"]:
            can.close()
            continue
        cans.append(can)
        refs.append(ref)
    say("batch", len(cans))
    assert len(cans) >= 8
    out = sgm.enumerate_batch(cans, None, L.SGM_ENUM_DYN_DAF)
    for can, r, ref in zip(cans, out, refs):
        _same(r, ref, can.query.name)
    V = g2k.getVertexCnt()
    cuts = [0, V // 7, V // 3, V // 2, V]
    parts = [sgm.enumerate_batch(cans, None, L.SGM_ENUM_DYN_VEQ, roots=(a, b)) for a, b in zip(cuts, cuts[1:])]
    whole = sgm.enumerate_batch(cans, None, L.SGM_ENUM_DYN_VEQ)
    for i in range(len(cans)):
        assert sum(p[i].ans for p in parts) == whole[i].ans
        assert sum(p[i].intersections for p in parts) == whole[i].intersections
        assert [sum(p[i].nodes[d] for p in parts) for d in range(len(whole[i].nodes))] == whole[i].nodes
    for can in cans:
        can.close()


def test_small_stack_reruns(g2k):
    """A stack too small for the intersections is re-run deeper, with the same result."""
    for q in mixed(g2k, 10, 760, 7, 12):
        can, ok, c_r = _daf(g2k, q)
        if ok:
            bfs, ref = _oracle(g2k, q, c_r, "VEQ", cap=ORACLE_CAP)
            if not ref["capped"] and ref["intersections"]:
                r = sgm.enumerate_query(can, bfs, L.SGM_ENUM_DYN_VEQ, stack_words=4)
                _same(r, ref, q.name)
                say("retries", r.retries)
                can.close()
                return
        can.close()
    pytest.fail("no query with intersections")


def test_dropin_calls(g2k):
    """The reference's driver sequence (src/order_ablations.cpp:131-172): computeWeightArray,
    then explore_Intersection_DynamicOrder_DAF with the weight array it was handed; VEQ over
    the DAFStartVertex BFS.  match_order comes back empty, Profiler counts intersections."""
    for q in mixed(g2k, 10, 770, 4, 9):
        can, ok, c_r = _daf(g2k, q)
        if not ok:
            can.close()
            continue
        bfs_r, ref = _oracle(g2k, q, c_r, "DAF", cap=ORACLE_CAP)
        if ref["capped"]:
            can.close()
            continue
        cfg = sgm.Config(useFilter=True, can=can)
        bfs = []
        w = sgm.computeWeightArray(g2k, q, can, bfs)
        prof = sgm.Profiler.getInst()
        before = prof.total_intersection
        match = [9, 9]
        r = sgm.explore_Intersection_DynamicOrder_DAF(g2k, q, bfs, match, cfg, w)
        _same(r, ref, q.name)
        assert match == [] and prof.total_intersection - before == ref["intersections"]
        # weights the caller changed are the ones used: all-equal weights leave the ranking
        # to degree-1 and the heap, as the oracle does with the same weights
        flat = [np.ones_like(x) for x in w]
        r2 = sgm.explore_Intersection_DynamicOrder_DAF(g2k, q, bfs, [], cfg, flat)
        ref2 = O.enum_dynamic(g2k, q, c_r, "DAF", bfs_order=bfs_r, weights=flat)
        _same(r2, ref2, q.name)
        _, ref3 = _oracle(g2k, q, c_r, "VEQ")
        r3 = sgm.explore_Intersection_DynamicOrder_VEQ(g2k, q, bfs, [], cfg)
        _same(r3, ref3, q.name)
        can.close()
        return
    pytest.fail("no query finished in the oracle")
