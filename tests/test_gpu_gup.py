"""GPU parity of GuP's conflict detection (explore_Intersection_All with
useGUPConflictDetection: SGM_ENUM_AUX_TABLES_GUPCD, with or without DAF failing sets)
against the oracle on the same seeded inputs: the library's reservation guards equal the
oracle's record for record, and the search's count, intersections, extensions per depth and
prunes equal the oracle's exactly when uncapped."""
import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from brute import random_connected_query, random_graph
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels

pytestmark = pytest.mark.gpu

CAP = 2_000_000


def say(*a):
    print(*a, flush=True)


def mixed(g, count, seed, lo=4, hi=11):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


def inputs():
    g = with_labels(chung_lu(1500, 2.4, 6.0, seed=91), 4, seed=92)
    out = [(g, q) for q in mixed(g, 10, 93)]
    rng = np.random.default_rng(94)
    for _ in range(25):
        gg = random_graph(int(rng.integers(9, 15)), float(rng.uniform(0.3, 0.6)), rng)
        out.append((gg, random_connected_query(int(rng.integers(3, 7)), int(rng.integers(0, 4)), rng)))
    return out


@pytest.fixture(scope="module")
def cases():
    return inputs()


def _filtered(g, q, kind):
    can = sgm.CandidateParam(g, q)
    if kind == "GQL":
        ok = sgm.Filter_GQL(g, q, can)
        ok_r, c_r = O.gql(g, q)
    else:
        ok = sgm.Filter_DAF(g, q, can)
        ok_r, c_r = O.filter_tree("DAF", g, q)
    assert ok == ok_r, q.name
    return can, ok, c_r


def test_guards_equal_oracle(cases):
    seen = 0
    for g, q in cases:
        for kind in ("GQL", "DAF"):
            can, ok, c_r = _filtered(g, q, kind)
            if ok:
                order = O.order_ri(q)
                goff, words = sgm.gup_guards(can, order)
                goff_r, words_r = O.gup_guards(g, q, c_r, order)
                assert np.array_equal(goff, goff_r) and np.array_equal(words, words_r), (q.name, kind)
                seen += int((words_r[0::4] > 0).sum())
            can.close()
    say("guards", seen)
    assert seen > 50


@pytest.mark.parametrize("fs", [False, True])
def test_search_equals_oracle(cases, fs):
    method = L.SGM_ENUM_AUX_TABLES_FS_GUPCD if fs else L.SGM_ENUM_AUX_TABLES_GUPCD
    checked = prunes = 0
    for g, q in cases:
        can, ok, c_r = _filtered(g, q, "DAF")
        if not ok:
            can.close()
            continue
        order = O.order_ri(q)
        gd = O.gup_guards(g, q, c_r, order)
        ref = O.enum_all(g, q, c_r, order, cap=CAP, fs=fs, guards=gd)
        r = sgm.enumerate_query(can, order, method, cap=CAP)
        assert r.ans == ref["count"], (q.name, fs)
        if not ref["capped"]:
            assert (r.intersections, r.nodes, r.fs_prunes) == (ref["intersections"], ref["nodes"],
                                                                 ref["fs_prunes"]), (q.name, fs)
            checked += 1
            prunes += ref["fs_prunes"]
        can.close()
    say("checked", checked, "prunes", prunes)
    assert checked >= 20 and prunes > 0


def test_batch_and_dropin(cases):
    """Many queries in one launch with RI orders from the library, and the drop-in call with
    the template flags as keywords."""
    cans, refs = [], []
    for g, q in cases[:10]:
        can, ok, c_r = _filtered(g, q, "GQL")
        if not ok:
            can.close()
            continue
        order = O.order_ri(q)
        cans.append(can)
        refs.append(O.enum_all(g, q, c_r, order, cap=CAP, guards=O.gup_guards(g, q, c_r, order)))
    assert len(cans) >= 5
    out = sgm.enumerate_batch(cans, None, L.SGM_ENUM_AUX_TABLES_GUPCD, cap=CAP)
    for can, r, ref in zip(cans, out, refs):
        assert r.ans == ref["count"], can.query.name
        if not ref["capped"]:
            assert (r.intersections, r.nodes, r.fs_prunes) == (ref["intersections"], ref["nodes"], ref["fs_prunes"])
    can = cans[0]
    g, q = can.data, can.query
    prof = sgm.Profiler.getInst()
    before = prof.gup_cd_cnt
    r = sgm.explore_Intersection_All(g, q, O.order_ri(q), sgm.Config(useFilter=True, can=can), output_limit=CAP,
                                     GUPConflictDetection=True)
    assert r.ans == refs[0]["count"] and prof.gup_cd_cnt - before == r.fs_prunes
    with pytest.raises(NotImplementedError):
        sgm.explore_Intersection_All(g, q, O.order_ri(q), sgm.Config(useFilter=True, can=can), GuPFailingPrune=True)
    for c in cans:
        c.close()


@pytest.mark.parametrize("fs", [False, True])
def test_aux_last_level_hub_bitmaps(fs):
    """The aux-table search's last level probes a sibling's row through the sibling's
    adjacency bitmap when it is a hub (degree >= 64): the same results as the binary
    searches of SGM_ENUM_F_NO_BITMAPS, and the oracle's, uncapped."""
    g = with_labels(chung_lu(4000, 2.2, 12.0, seed=95), 3, seed=96)
    assert int(np.diff(g.off).max()) >= 64
    method = L.SGM_ENUM_AUX_TABLES_FS if fs else L.SGM_ENUM_AUX_TABLES
    checked = 0
    cap = 10_000_000
    for q in mixed(g, 14, 97, 4, 9):
        can = sgm.CandidateParam(g, q)
        ok = sgm.Filter_GQL(g, q, can)
        ok_r, c_r = O.gql(g, q)
        assert ok == ok_r
        if ok:
            order = O.order_ri(q)
            ref = O.enum_all(g, q, c_r, order, cap=cap, fs=fs)
            r = sgm.enumerate_query(can, order, method, cap=cap)
            rb = sgm.enumerate_query(can, order, method, cap=cap, flags=L.SGM_ENUM_F_NO_BITMAPS)
            assert r.ans == rb.ans == ref["count"], q.name
            if not ref["capped"]:
                for x in (r, rb):
                    assert (x.intersections, x.nodes, x.fs_prunes) == (ref["intersections"], ref["nodes"],
                                                                         ref["fs_prunes"]), q.name
                checked += 1
        can.close()
    say("checked", checked)
    assert checked >= 4
