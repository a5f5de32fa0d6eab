"""GPU parity of the BFS-tree filters (CFL, CECI, DAF: k_refine steps on the device) and the
QSI, CFL, CECI and GuP orders (host rules over device NLF counts, aux tables and row sums)
against the oracle, on the same seeded inputs: candidate sets bit for bit, orders exactly,
and the reference's candidate-space rows built on them (counts, intersections, per-depth
nodes) exactly."""
import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels

pytestmark = pytest.mark.gpu

FILTERS = {"CFL": sgm.Filter_CFL, "CECI": sgm.Filter_CECI, "DAF": sgm.Filter_DAF}
KIND = {"CFL": L.SGM_FILTER_CFL, "CECI": L.SGM_FILTER_CECI, "DAF": L.SGM_FILTER_DAF}
ORDERS = {"QSI": sgm.Order_QSI, "CFL": sgm.Order_CFL, "CECI": sgm.Order_CECI, "GuP": sgm.Order_GuP}


def say(*a):
    print(*a, flush=True)


def mixed(g, count, seed, lo=4, hi=17):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


@pytest.fixture(scope="module")
def g3k():
    return chung_lu(3000, 2.4, 7.0, seed=61)


def _same_sets(can, c_r, tag):
    assert [int(x) for x in can.candidates_count] == [len(x) for x in c_r], tag
    for u, ref in enumerate(c_r):
        assert np.array_equal(can.candidates[u], ref), (tag, u)


@pytest.mark.parametrize("kind", ["CFL", "CECI", "DAF"])
def test_tree_filter_parity(g3k, kind):
    for q in mixed(g3k, 10, 611, 4, 33):
        say(kind, q.name)
        can = sgm.CandidateParam(g3k, q)
        ok = FILTERS[kind](g3k, q, can, build_tables=False)
        ok_r, c_r = O.filter_tree(kind, g3k, q)
        assert ok == ok_r, q.name
        if ok:
            _same_sets(can, c_r, (kind, q.name))
        can.close()


@pytest.mark.parametrize("labels", [3, 8])
def test_tree_filter_parity_labelled(labels):
    g = with_labels(chung_lu(3000, 2.4, 7.0, seed=62), labels, seed=63)
    for q in mixed(g, 8, 620 + labels, 4, 13):
        for kind in FILTERS:
            say(kind, labels, q.name)
            can = sgm.CandidateParam(g, q)
            ok = FILTERS[kind](g, q, can, build_tables=False)
            ok_r, c_r = O.filter_tree(kind, g, q)
            assert ok == ok_r, (kind, q.name)
            if ok:
                _same_sets(can, c_r, (kind, q.name))
            can.close()


@pytest.mark.parametrize("kind", ["CFL", "CECI", "DAF"])
def test_tree_filter_batch_equals_oracle(g3k, kind):
    """Many queries of uneven size in one set of launches (one refinement launch per step
    index for the whole batch)."""
    qs = mixed(g3k, 24, 630 + KIND[kind], 3, 33)
    cans = [sgm.CandidateParam(g3k, q) for q in qs]
    oks = sgm.filter_batch(cans, KIND[kind])
    say(kind, "batch", len(qs), sum(oks))
    for q, can, ok in zip(qs, cans, oks):
        ok_r, c_r = O.filter_tree(kind, g3k, q)
        assert ok == ok_r, q.name
        if ok:
            _same_sets(can, c_r, (kind, q.name))
        can.close()


def test_tree_filters_power_law_100k():
    """The Chung-Lu 10^5 graph (Chung-Lu 10^5, exponent 2.3, average degree 16) and four of its
    dense 16-vertex queries."""
    g = chung_lu(100000, 2.3, 16.0, seed=2)
    qs = query_set(g, 4, 16, True, seed=22)
    for kind in FILTERS:
        cans = [sgm.CandidateParam(g, q) for q in qs]
        oks = sgm.filter_batch(cans, KIND[kind])
        for q, can, ok in zip(qs, cans, oks):
            say(kind, "100k", q.name)
            ok_r, c_r = O.filter_tree(kind, g, q)
            assert ok == ok_r
            if ok:
                assert O.cand_digest(can.candidates) == O.cand_digest(c_r), (kind, q.name)
            can.close()


@pytest.mark.parametrize("filt", ["DAF", "GQL"])
def test_orders_parity(g3k, filt):
    seen_gup = 0
    for q in mixed(g3k, 14, 640, 4, 33):
        can = sgm.CandidateParam(g3k, q)
        if filt == "GQL":
            ok = sgm.Filter_GQL(g3k, q, can)
            ok_r, c_r = O.gql(g3k, q)
        else:
            ok = sgm.Filter_DAF(g3k, q, can)
            ok_r, c_r = O.filter_tree("DAF", g3k, q)
        assert ok == ok_r
        if not ok:
            continue
        for name, fn in ORDERS.items():
            say(filt, name, q.name)
            ref = O.order_tree(name, g3k, q, c_r)
            if ref is None:
                with pytest.raises(L.SgmError):
                    fn(g3k, q, can)
                continue
            assert fn(g3k, q, can) == ref, (name, q.name)
            seen_gup += name == "GuP"
        can.close()
    assert seen_gup > 3


def test_order_ceci_without_candidates(g3k):
    """Order_CECI reads NLF counts only, so it runs on an unfiltered query, as in the
    reference (inc/order.h:599-607 takes no candidate space)."""
    for q in mixed(g3k, 6, 650, 4, 20):
        ok_r, c_r = O.nlf(g3k, q)
        assert sgm.Order_CECI(g3k, q) == O.order_tree("CECI", g3k, q, c_r)


def test_tree_rows_enumerate_like_oracle(g3k):
    """The reference harness's QSI / CFL / CECI / DAF rows (src/overall_performance.cpp:66-127,
    158-172): each filter and order, then the candidate-space search (DAF with failing
    sets); counts, intersections, per-depth nodes and cut-offs equal the oracle's."""
    cap = 200_000
    for q in mixed(g3k, 8, 660, 4, 11):
        for row in ("QSI", "CFL", "CECI", "DAF"):
            can = sgm.CandidateParam(g3k, q)
            if row == "QSI":
                ok = sgm.Filter_NLF(g3k, q, can)
                ok_r, c_r = O.nlf(g3k, q)
            else:
                ok = FILTERS[row](g3k, q, can)
                ok_r, c_r = O.filter_tree(row, g3k, q)
            assert ok == ok_r
            if not ok:
                can.close()
                continue
            order = sgm.Order_RI(q) if row == "DAF" else ORDERS[row](g3k, q, can)
            ref_order = O.order_ri(q) if row == "DAF" else O.order_tree(row, g3k, q, c_r)
            assert order == ref_order, (row, q.name)
            method = L.SGM_ENUM_AUX_TABLES_FS if row == "DAF" else L.SGM_ENUM_AUX_TABLES
            res = sgm.enumerate_query(can, order, method, cap=cap)
            ref = O.enum_all(g3k, q, c_r, order, cap=cap, fs=row == "DAF")
            say(row, q.name, res.ans, ref["count"])
            assert res.ans == ref["count"], (row, q.name)
            if ref["count"] < cap:
                assert res.intersections == ref["intersections"], (row, q.name)
                assert list(res.nodes[:q.getVertexCnt()]) == ref["nodes"], (row, q.name)
                assert res.fs_prunes == ref["fs_prunes"], (row, q.name)
            can.close()
