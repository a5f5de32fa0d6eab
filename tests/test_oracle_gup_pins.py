"""Pins the oracle's GuP conflict detection (oracle/sgm_oracle_gup.c): guards equal a second
restatement built on std::partition, std::unordered_set and std::sort themselves
(oracle/sort_probe.cpp), which checks the restated element orders; counts with the check
equal counts without it and brute force (a guard may only remove candidates that complete
no embedding); the check prunes on these inputs; guards are small, sorted, and absent at
depth 0 and at the last depth, where nothing can be reserved."""
import numpy as np
import pytest

import oracle as O
from brute import count_embeddings, random_connected_query, random_graph
from subgraphmatching_amd.generators import chung_lu, query_set


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def cases(seed, count, V=(8, 15), p=(0.25, 0.7), n=(3, 7)):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        g = random_graph(int(rng.integers(*V)), float(rng.uniform(*p)), rng)
        yield g, random_connected_query(int(rng.integers(n[0], n[1] + 1)), int(rng.integers(0, 4)), rng)


def test_guards_equal_std_restatement():
    seen = 0
    inputs = list(cases(81, 150))
    g = chung_lu(600, 2.4, 6.0, seed=82)
    inputs += [(g, q) for q in query_set(g, 10, lambda i, r: int(r.integers(4, 9)), lambda i, r: bool(i % 2),
                                         seed=83)]
    for g, q in inputs:
        for filt in (O.gql, lambda g, q: O.filter_tree("DAF", g, q)):
            ok, c = filt(g, q)
            if not ok:
                continue
            order = O.order_ri(q)
            a = O.gup_guards(g, q, c, order)
            b = O.probe_gup_guards(g, q, c, order)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), q.name
            seen += int((a[1][0::4] > 0).sum())
    assert seen > 100


def test_counts_equal_plain_search_and_brute_force():
    prunes = 0
    for g, q in cases(84, 200, V=(7, 12), n=(3, 6)):
        ok, c = O.gql(g, q)
        bf = count_embeddings(g, q)
        if not ok:
            assert bf == 0
            continue
        order = O.order_ri(q)
        gd = O.gup_guards(g, q, c, order)
        for fs in (False, True):
            plain = O.enum_all(g, q, c, order, fs=fs)
            cd = O.enum_all(g, q, c, order, fs=fs, guards=gd)
            assert plain["count"] == cd["count"] == bf, (q.name, fs)
            # a pruned candidate is not extended: no depth gains nodes
            assert all(a <= b for a, b in zip(cd["nodes"], plain["nodes"]))
            if not fs:
                prunes += cd["fs_prunes"]
    assert prunes > 50


def test_guard_shape():
    for g, q in cases(85, 80):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        goff, w = O.gup_guards(g, q, c, order)
        n = q.getVertexCnt()
        rec = w.reshape(-1, 4)
        assert (rec[:, 0] <= 3).all()
        for k in range(len(rec)):
            s = int(rec[k, 0])
            assert list(rec[k, 1:1 + s]) == sorted(set(rec[k, 1:1 + s].tolist()))
        assert (rec[goff[0]:goff[1], 0] == 0).all() and (rec[goff[n - 1]:goff[n], 0] == 0).all()
