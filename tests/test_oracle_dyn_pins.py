"""Pins the oracle's dynamic-order searches (explore_Intersection_DynamicOrder_DAF / _VEQ) and
the DAF weight array (oracle/sgm_oracle_dyn.c): counts against the plain candidate-space
search and brute force (an order never changes a count), intersections and per-depth
extensions against the same search built on std::priority_queue itself
(oracle/sort_probe.cpp, so the restated heap moves are checked inside the search), and the
weights against a Python restatement of computeWeightArray's recursion."""
import numpy as np
import pytest

import oracle as O
import tree_sets as T
from brute import count_embeddings, random_connected_query, random_graph
from subgraphmatching_amd.generators import chung_lu, query_set

M32 = 1 << 32


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def cases(seed, count, V=(8, 13), p=(0.3, 0.6), n=(3, 6)):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        g = random_graph(int(rng.integers(*V)), float(rng.uniform(*p)), rng)
        k = int(rng.integers(n[0], n[1] + 1))
        yield g, random_connected_query(k, int(rng.integers(0, 4)), rng)


def weights_py(g, q, cands, bfs, parent):
    """computeWeightArray (inc/order.h:980-1030) over sets and dicts."""
    Tb = T.tables(g, q, cands)
    prev, _ = T.prev_follow(q, bfs)
    _, _, _, children = T.bfs_tree(q, bfs[0])
    w = [[M32 - 1] * len(c) for c in cands]
    for x in reversed(bfs):
        one = True
        for ch in children[x]:
            if len(prev[ch]) == 1:
                one = False
                for k in range(len(cands[x])):
                    w[x][k] = min(w[x][k], sum(w[ch][i] for i in Tb[x, ch][k]) % M32)
        if one:
            w[x] = [1] * len(cands[x])
    return w


def test_weights_match_restatement():
    for g, q in list(cases(71, 150)):
        ok, c = O.filter_tree("DAF", g, q)
        if not ok:
            continue
        bfs, parent, _ = O.dyn_bfs(g, q)
        cl = [x.tolist() for x in c]
        assert [x.tolist() for x in O.daf_weights(g, q, c, bfs, parent)] == weights_py(g, q, cl, bfs, parent)


def test_counts_equal_plain_search_and_brute_force():
    for g, q in cases(72, 120, V=(6, 9), n=(3, 4)):
        ok, c = O.filter_tree("DAF", g, q)
        bf = count_embeddings(g, q)
        if not ok:
            assert bf == 0
            continue
        plain = O.enum_all(g, q, c, O.order_ri(q))["count"]
        assert plain == bf
        for kind in ("DAF", "VEQ"):
            assert O.enum_dynamic(g, q, c, kind)["count"] == bf, kind


@pytest.mark.parametrize("kind", ["DAF", "VEQ"])
def test_search_equals_std_priority_queue_restatement(kind):
    """Uncapped on random graphs, capped on power-law graphs: count, intersections and
    per-depth extensions equal the std::priority_queue restatement (tie order included)."""
    checked = 0
    inputs = [(g, q, None) for g, q in cases(73, 150)]
    for s in range(2):
        g = chung_lu(1500, 2.4, 7.0, seed=80 + s)
        inputs += [(g, q, 100_000) for q in query_set(g, 6, lambda i, r: int(r.integers(4, 11)),
                                                       lambda i, r: bool(i % 2), seed=90 + s)]
    for g, q, cap in inputs:
        ok, c = O.filter_tree("DAF", g, q)
        if not ok:
            continue
        bfs, parent, _ = O.dyn_bfs(g, q)
        w = O.daf_weights(g, q, c, bfs, parent) if kind == "DAF" else None
        d = O.enum_dynamic(g, q, c, kind, cap=cap)
        p = O.probe_dynamic(g, q, c, bfs, w, cap=cap)
        assert (d["count"], d["intersections"], d["nodes"]) == p, q.name
        checked += 1
    assert checked > 60


def test_ranking_changes_the_search():
    """The two rankings (weights vs candidate counts) pick different orders on some inputs:
    intersections differ while counts agree, so a ranking slip would show."""
    diff = 0
    g = chung_lu(1500, 2.4, 7.0, seed=81)
    for q in query_set(g, 12, lambda i, r: int(r.integers(5, 11)), lambda i, r: True, seed=95):
        ok, c = O.filter_tree("DAF", g, q)
        if not ok:
            continue
        a = O.enum_dynamic(g, q, c, "DAF", cap=50_000)
        b = O.enum_dynamic(g, q, c, "VEQ", cap=50_000)
        assert a["count"] == b["count"]
        diff += a["intersections"] != b["intersections"]
    assert diff > 0
