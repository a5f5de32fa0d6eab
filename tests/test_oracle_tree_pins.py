"""Pins the oracle's BFS-tree filters (CFL, CECI, DAF) and its QSI / CFL / CECI / GuP orders
(oracle/sgm_oracle_tree.c) to independent restatements: core numbers from their
definition, std::priority_queue itself for the one libstdc++ tie order, set-level
restatements of every filter schedule and order (tests/tree_sets.py), soundness against
brute-force embeddings, and hand-worked answers.  Checks that a plausible slip fails: a
different DAF pass count, CFL without its same-level reverse pruning, a fixpoint, the wrong
end of QSI's start edge.
"""
import numpy as np
import pytest

import oracle as O
import tree_sets as T
from brute import embeddings, nlf_by_definition, query_from_edges, random_connected_query, random_graph
from subgraphmatching_amd.generators import chung_lu, query_set


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def cases(seed, count, V=(8, 14), p=(0.25, 0.55), n=(3, 6), labels=0):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        g = random_graph(int(rng.integers(*V)), float(rng.uniform(*p)), rng)
        k = int(rng.integers(n[0], n[1] + 1))
        q = random_connected_query(k, int(rng.integers(0, 4)), rng)
        if labels:
            g.set_labels(rng.integers(0, labels, g.getVertexCnt()))
            q.set_labels(rng.integers(0, labels, k))
        yield g, q


def mid_cases():
    """Power-law graphs of a few hundred vertices with random-walk queries (dense and sparse)."""
    out = []
    for s in range(3):
        g = chung_lu(400, 2.4, 6.0, seed=100 + s)
        out += [(g, q) for q in query_set(g, 4, 7, True, seed=200 + s) + query_set(g, 4, 6, False, seed=300 + s)]
    return out


def test_heap_tie_order_is_std_priority_queue():
    rng = np.random.default_rng(5)
    for _ in range(1500):
        m = int(rng.integers(1, 33))
        keys = rng.integers(0, 4, m).astype(float)
        keep = int(rng.choice([1, 3, 5, m]))
        assert O.heap_pop_order(keys, keep) == O.heap_pop_order(keys, keep, use_std=True), (keys, keep)


def test_kcore_by_definition():
    rng = np.random.default_rng(7)
    for _ in range(300):
        k = int(rng.integers(2, 14))
        q = random_connected_query(k, int(rng.integers(0, 2 * k)), rng)
        assert O.kcore(q) == T.core_numbers(q)
    # known answers: K4 plus a pendant vertex; a cycle; a path
    assert O.kcore(query_from_edges(5, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4)])) == [3, 3, 3, 3, 1]
    assert O.kcore(query_from_edges(4, [(0, 1), (1, 2), (2, 3), (0, 3)])) == [2, 2, 2, 2]
    assert O.kcore(query_from_edges(3, [(0, 1), (1, 2)])) == [1, 1, 1]


def test_bfs_tree():
    rng = np.random.default_rng(8)
    for _ in range(200):
        k = int(rng.integers(2, 12))
        q = random_connected_query(k, int(rng.integers(0, k)), rng)
        s = int(rng.integers(k))
        order, parent, level = O.bfs(q, s)
        o2, p2, l2, _ = T.bfs_tree(q, s)
        assert order == o2
        assert all(parent[v] == p2[v] and level[v] == l2[v] for v in range(k))
    # children in adjacency order: from 0 of the star 0-{3,1,2} the order is 0,1,2,3
    assert O.bfs(query_from_edges(4, [(0, 3), (0, 1), (0, 2)]), 0)[0] == [0, 1, 2, 3]


def test_start_vertices():
    for labels in (0, 3):
        for g, q in cases(11 + labels, 150, labels=labels):
            assert O.start_vertex("CFL", g, q) == T.cfl_start(g, q)
            assert O.start_vertex("CECI", g, q) == T.ceci_daf_start(g, q, False)
            assert O.start_vertex("DAF", g, q) == T.ceci_daf_start(g, q, True)


def _same(ok, c, ok2, c2):
    assert ok == ok2
    if ok:
        assert [x.tolist() for x in c] == c2


@pytest.mark.parametrize("labels", [0, 2, 4])
def test_filters_match_set_restatements(labels):
    for g, q in cases(21 + labels, 200, labels=labels):
        _same(*O.filter_tree("CFL", g, q), *T.cfl_sets(g, q))
        _same(*O.filter_tree("CECI", g, q), *T.ceci_sets(g, q))
        _same(*O.filter_tree("DAF", g, q), *T.daf_sets(g, q))


def test_filters_match_set_restatements_power_law():
    for g, q in mid_cases():
        _same(*O.filter_tree("CFL", g, q), *T.cfl_sets(g, q))
        _same(*O.filter_tree("CECI", g, q), *T.ceci_sets(g, q))
        _same(*O.filter_tree("DAF", g, q), *T.daf_sets(g, q))


def test_filters_sound_and_within_nlf():
    """Every data vertex used by some embedding survives every filter (the filters remove
    only vertices that cannot be matched), and every set lies inside the NLF set."""
    for labels in (0, 2):
        for g, q in cases(31 + labels, 120, V=(6, 9), n=(3, 4), labels=labels):
            used = [set() for _ in range(q.getVertexCnt())]
            for m in embeddings(g, q):
                for u, v in enumerate(m):
                    used[u].add(v)
            nlf = nlf_by_definition(g, q)
            for kind in ("CFL", "CECI", "DAF"):
                ok, c = O.filter_tree(kind, g, q)
                if any(used):
                    assert ok, kind
                if ok:
                    for u in range(q.getVertexCnt()):
                        assert used[u] <= set(c[u].tolist()), (kind, u)
                        assert set(c[u].tolist()) <= set(nlf[u]), (kind, u)


# Witnesses found by a seeded search over random graphs: on each, the stated schedule and a
# plausible alternative give different sets, so the equality checks above would catch that
# slip.  (graph CSR offsets, adjacency, query CSR offsets, adjacency)
DAF_WITNESS = ([0, 1, 5, 7, 10, 12, 12, 17, 20, 22, 23, 24, 25, 26, 28, 30, 32],
               [6, 3, 7, 14, 15, 6, 7, 1, 11, 14, 6, 7, 0, 2, 4, 12, 15, 1, 2, 4, 9, 13, 8, 13, 3, 6, 8, 10, 1, 3,
                1, 6],
               [0, 4, 7, 8, 10, 12], [1, 2, 3, 4, 0, 3, 4, 0, 0, 1, 0, 1])
CFL_WITNESS = ([0, 3, 6, 9, 13, 14, 18, 21, 23, 27, 29, 30, 34, 36],
               [1, 5, 11, 0, 8, 9, 5, 8, 10, 6, 8, 9, 11, 5, 0, 2, 4, 12, 3, 8, 11, 11, 12, 1, 2, 3, 6, 1, 3, 2, 0,
                3, 6, 7, 5, 7],
               [0, 2, 6, 8, 11, 12, 14], [1, 3, 0, 2, 3, 4, 1, 5, 0, 1, 5, 1, 2, 3])
CECI_WITNESS = ([0, 0, 3, 7, 9, 11, 13, 15, 16, 19, 19, 20, 21, 23, 24],
                [3, 6, 11, 3, 7, 8, 12, 1, 2, 8, 13, 8, 10, 1, 12, 2, 2, 4, 5, 5, 1, 2, 6, 4],
                [0, 3, 4, 6, 8, 11, 12, 14], [1, 2, 3, 0, 0, 6, 0, 4, 3, 5, 6, 4, 2, 4])


def _witness(w):
    from subgraphmatching_amd.graph import Graph, Query
    return (Graph(np.array(w[0], dtype=np.uint32), np.array(w[1], dtype=np.uint32)),
            Query(np.array(w[2], dtype=np.uint32), np.array(w[3], dtype=np.uint32)))


def test_schedules_matter():
    """DAF's three passes are not its fixpoint; CFL's same-level reverse pruning and CECI's
    non-tree-edge invalidation each change a set."""
    from brute import adjsets
    g, q = _witness(DAF_WITNESS)
    ok, c = O.filter_tree("DAF", g, q)
    assert ok and [x.tolist() for x in c] == T.daf_sets(g, q)[1] != T.daf_sets(g, q, passes=7)[1]

    g, q = _witness(CFL_WITNESS)
    ok, c = O.filter_tree("CFL", g, q)
    A, nlf = adjsets(g), [set(x) for x in nlf_by_definition(g, q)]
    order, parent, level, children = T.bfs_tree(q, T.cfl_start(g, q))
    prev, fol = T.prev_follow(q, order)
    C = {order[0]: set(nlf[order[0]])}
    for u in order[1:]:  # generation and bottom-up pruning, but no same-level pruning
        C[u] = T._keep(A, nlf[u], prev[u], C)
    for lv in range(max(level.values()) - 1, -1, -1):
        for u in [x for x in order if level[x] == lv]:
            deeper = [w for w in fol[u] if level[w] > lv]
            if deeper:
                C[u] = T._keep(A, C[u], deeper, C)
    assert ok and [x.tolist() for x in c] == T.cfl_sets(g, q)[1] != [sorted(C[u]) for u in range(q.getVertexCnt())]

    g, q = _witness(CECI_WITNESS)
    ok, c = O.filter_tree("CECI", g, q)
    A, nlf = adjsets(g), [set(x) for x in nlf_by_definition(g, q)]
    order, parent, level, children = T.bfs_tree(q, T.ceci_daf_start(g, q, False))
    C = {order[0]: set(nlf[order[0]])}
    for u in order[1:]:  # tree-edge construction only
        p = parent[u]
        C[u] = {v for v in nlf[u] if A[v] & C[p]}
        C[p] = {v for v in C[p] if A[v] & nlf[u]}
    assert ok and [x.tolist() for x in c] == T.ceci_sets(g, q)[1] != [sorted(C[u]) for u in range(q.getVertexCnt())]


def test_ceci_root_and_exit_cases():
    # a label absent from the data graph: the root (or a later vertex) has no candidate
    g = random_graph(8, 0.5, np.random.default_rng(3))
    g.set_labels(np.zeros(8, dtype=np.int64))
    q = query_from_edges(3, [(0, 1), (1, 2)], labels=[0, 0, 7])
    for kind in ("CFL", "CECI", "DAF"):
        ok, _ = O.filter_tree(kind, g, q)
        assert not ok, kind


# ------------------------------------------------------------------------------ orders
def test_qsi_hand_worked():
    """Triangle 0-1-2 with pendant 3 on 2, matched against itself (unlabeled).  NLF sets:
    C(0) = C(1) = {0,1,2}, C(2) = {2}, C(3) = {0,1,2,3}; table entries e(0,1) = 6,
    e(0,2) = e(1,2) = 2, e(2,3) = 3.  The start edge is (0,2) (fewest entries, first seen),
    taken from its higher-degree end: 2, 0.  Then 1 through e(2,1) = 2, then 3."""
    q = query_from_edges(4, [(0, 1), (0, 2), (1, 2), (2, 3)])
    from subgraphmatching_amd.graph import Graph
    g = Graph(q.off, q.adj)
    ok, c = O.nlf(g, q)
    assert ok and [x.tolist() for x in c] == [[0, 1, 2], [0, 1, 2], [2], [0, 1, 2, 3]]
    assert O.order_tree("QSI", g, q, c) == [2, 0, 1, 3]


def _connected_prefixes(q, order):
    N = T.qnbrs(q)
    return all(any(w in order[:i] for w in N[order[i]]) for i in range(1, len(order)))


@pytest.mark.parametrize("labels", [0, 3])
def test_orders_match_restatements(labels):
    seen_gup = 0
    for g, q in list(cases(51 + labels, 150, labels=labels)) + (mid_cases() if not labels else []):
        ok, c = O.filter_tree("DAF", g, q)
        if not ok:
            continue
        cl = [x.tolist() for x in c]
        n = q.getVertexCnt()
        qsi = O.order_tree("QSI", g, q, c)
        assert qsi == T.qsi_order(g, q, cl)
        cfl = O.order_tree("CFL", g, q, c)
        assert cfl == T.cfl_order(g, q, cl)
        ceci = O.order_tree("CECI", g, q, c)
        assert ceci == T.bfs_tree(q, T.ceci_daf_start(g, q, False))[0]
        gup = O.order_tree("GuP", g, q, c)
        assert gup == T.gup_order(g, q, cl)
        for o in (qsi, cfl, ceci, gup):
            if o is None:
                continue
            assert sorted(o) == list(range(n)) and _connected_prefixes(q, o)
        if gup is not None:
            seen_gup += 1
    assert seen_gup > 10


def test_gup_refuses_paths_and_cycles():
    """Order_GuP asserts that some query vertex has degree > 2 (inc/order.h:484-492)."""
    from subgraphmatching_amd.graph import Graph
    g = chung_lu(200, 2.5, 5.0, seed=1)
    for es, n in (([(0, 1), (1, 2), (2, 3)], 4), ([(0, 1), (1, 2), (2, 3), (3, 0)], 4)):
        q = query_from_edges(n, es)
        ok, c = O.filter_tree("DAF", g, q)
        assert ok
        assert O.order_tree("GuP", g, q, c) is None
    assert isinstance(Graph, type)
