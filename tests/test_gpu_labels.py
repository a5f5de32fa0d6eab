"""GPU parity for labelled graphs and queries: labelled NLF and GQL candidate sets, and
the searches that follow from them (cache + filter with and without failing sets, the
aux-table search, the batched launch), bit-exact against the oracle; counts against brute
force on tiny inputs.  Labels reach the device filters through sgm_graph_set_labels /
sgm_query_set_labels (the reference's NLF maps, inc/graph.h:76-112, and NLFByVertex,
inc/filter.h:24-58)."""
import numpy as np
import pytest

import oracle as O
from brute import count_embeddings, query_from_edges, random_connected_query, random_graph
import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels
from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gl():
    return with_labels(chung_lu(10000, 2.5, 7.4, seed=1), 4, seed=41, skew=0.5)


@pytest.fixture(scope="module")
def gl_small():
    return with_labels(chung_lu(2000, 2.3, 8.0, seed=3), 3, seed=42)


def _mixed(g, count, seed, lo=4, hi=17):
    return query_set(g, count, lambda i, r: int(r.integers(lo, hi)), lambda i, r: bool(i % 2), seed=seed)


def _same_cands(can, q, c_r):
    assert [int(x) for x in can.candidates_count] == [len(x) for x in c_r]
    for u in range(q.getVertexCnt()):
        assert np.array_equal(can.candidates[u], c_r[u]), (q.name, u)


def test_labelled_nlf_parity(gl):
    for q in _mixed(gl, 8, 201):
        say("nlf lab", q.name)
        can = sgm.CandidateParam(gl, q)
        ok = sgm.Filter_NLF(gl, q, can, build_tables=False)
        ok_r, c_r = O.nlf(gl, q)
        assert ok == ok_r
        if ok:
            _same_cands(can, q, c_r)


def test_labelled_gql_parity(gl):
    qs = _mixed(gl, 12, 202) + query_set(gl, 3, 8, False, seed=203)
    for q in qs:
        say("gql lab", q.name)
        can = sgm.CandidateParam(gl, q)
        ok = sgm.Filter_GQL(gl, q, can, build_tables=False)
        ok_r, c_r = O.gql(gl, q)
        assert ok == ok_r
        if ok:
            _same_cands(can, q, c_r)
            ex = can.candidates_exsist
            for u in range(q.getVertexCnt()):
                assert np.array_equal(np.flatnonzero(ex[u]), c_r[u])
    # one batched filter launch over labelled and unlabeled queries of the same graph
    mixed = qs[:6] + [sgm.Query(q.off, q.adj, name=q.name + "_u") for q in qs[6:10]]
    cans = [sgm.CandidateParam(gl, q) for q in mixed]
    oks = sgm.filter_batch(cans)
    for q, can, ok in zip(mixed, cans, oks):
        ok_r, c_r = O.gql(gl, q)
        assert ok == ok_r, q.name
        if ok:
            _same_cands(can, q, c_r)


def test_labelled_enum_parity(gl_small):
    done = 0
    for q in _mixed(gl_small, 20, 204, 3, 9):
        ok_r, c_r = O.gql(gl_small, q)
        if not ok_r:
            continue
        order = O.order_ri(q)
        pre = O.enum_cache(gl_small, q, c_r, order, cap=3_000_000)
        if pre["capped"]:
            continue
        say("enum lab", q.name, pre["count"])
        _compare(gl_small, q, order)
        can = sgm.CandidateParam(gl_small, q)
        sgm.Filter_GQL(gl_small, q, can, build_tables=False)
        _compare(gl_small, q, sgm.Order_GQL(gl_small, q, can))
        done += 1
    assert done >= 8


def test_labelled_enum_batch(gl_small):
    qs = _mixed(gl_small, 16, 205, 3, 10)
    cans = [sgm.CandidateParam(gl_small, q) for q in qs]
    sgm.filter_batch(cans)
    refs, keep = [], []
    for i, q in enumerate(qs):
        ok_r, c_r = O.gql(gl_small, q)
        ref = O.enum_cache(gl_small, q, c_r, O.order_ri(q), cap=3_000_000)
        if not ref["capped"]:
            keep.append(i)
            refs.append(ref)
    say("batch lab", len(keep))
    res = sgm.enumerate_batch([cans[i] for i in keep], None)
    for r, ref, i in zip(res, refs, keep):
        assert (r.ans, r.nodes, r.intersections, r.fs_prunes) == \
            (ref["count"], ref["nodes"], ref["intersections"], ref["fs_prunes"]), qs[i].name


def test_labelled_tiny_bruteforce():
    rng = np.random.default_rng(206)
    seen = 0
    for i in range(40):
        g = random_graph(int(rng.integers(7, 11)), float(rng.uniform(0.3, 0.7)), rng)
        g.set_labels(rng.integers(0, 2, g.getVertexCnt()))
        k = int(rng.integers(2, 6))
        q = random_connected_query(k, int(rng.integers(0, 3)), rng)
        q.set_labels(rng.integers(0, 2, k))
        exp = count_embeddings(g, q)
        say("tiny lab", i, k, exp)
        can = sgm.CandidateParam(g, q)
        if not sgm.Filter_GQL(g, q, can, build_tables=False):
            assert exp == 0
            continue
        seen += exp > 0
        for order in (sgm.Order_RI(q), sgm.Order_GQL(g, q, can)):
            _compare(g, q, order)
            assert sgm.enumerate_query(can, order).ans == exp
    assert seen >= 5


def test_label_mismatch_between_graph_and_query(gl_small):
    """An unlabeled query on a labelled graph matches label-0 vertices only; a labelled
    query on an unlabeled graph has no candidates for its non-zero labels; all-zero labels
    are the unlabeled case."""
    g0 = chung_lu(2000, 2.3, 8.0, seed=3)
    tri = query_from_edges(3, [(0, 1), (1, 2), (0, 2)])
    for g, q in ((gl_small, tri), (g0, query_from_edges(3, [(0, 1), (1, 2), (0, 2)], labels=[0, 1, 0])),
                 (g0, query_from_edges(3, [(0, 1), (1, 2), (0, 2)], labels=[0, 0, 0]))):
        can = sgm.CandidateParam(g, q)
        ok = sgm.Filter_GQL(g, q, can, build_tables=False)
        ok_r, c_r = O.gql(g, q)
        assert ok == ok_r
        if ok:
            _same_cands(can, q, c_r)
            ref = O.enum_cache(g, q, c_r, O.order_ri(q))
            assert sgm.enumerate_query(can, O.order_ri(q)).ans == ref["count"]
    lab0 = np.flatnonzero(gl_small.label_array() == 0)
    can = sgm.CandidateParam(gl_small, tri)
    assert sgm.Filter_NLF(gl_small, tri, can, build_tables=False)
    assert set(can.candidates[0].tolist()) <= set(lab0.tolist())


def test_query_labels_reset_filter_state(gl_small):
    q = query_set(gl_small, 1, 5, True, seed=207)[0]
    can = sgm.CandidateParam(gl_small, q)
    assert sgm.Filter_GQL(gl_small, q, can, build_tables=False)
    other = (q.label_array() + 1) % 3
    lab = L.u32(other)
    assert L.lib().sgm_query_set_labels(can._q, L.ptr(lab), q.getVertexCnt()) == L.SGM_OK
    with pytest.raises(L.SgmError):
        sgm.enumerate_query(can, sgm.Order_RI(q))


def test_labelled_batch_runner(gl_small):
    """The replayed-graph runner with labels: per-slot candidate counts, counts and (uncapped)
    nodes / intersections / prunes equal the oracle's; a batch mixing labelled and unlabeled
    queries on an unlabeled graph, and unlabeled queries on the labelled graph, too."""
    g0 = chung_lu(2000, 2.3, 8.0, seed=3)
    qs = _mixed(gl_small, 12, 209, 3, 10)
    cases = [(gl_small, qs), (gl_small, [sgm.Query(q.off, q.adj, name=q.name + "_u") for q in qs[:6]]),
             (g0, qs[:6] + query_set(g0, 6, lambda i, r: int(r.integers(3, 8)), True, seed=210))]
    for g, batch in cases:
        b = sgm.Batch(g, slots=16)
        for rep in range(2):  # the second filter replays the captured graph
            ok, cnt, _ = b.filter(batch)
            res = b.enumerate(cap=3_000_000)
            say("batch lab", g.name, len(batch), rep)
            for q, good, c, r in zip(batch, ok, cnt, res):
                ok_r, c_r = O.gql(g, q)
                assert bool(good) == ok_r, q.name
                if not ok_r:
                    assert r.ans == 0
                    continue
                assert [int(x) for x in c[:q.getVertexCnt()]] == [len(x) for x in c_r], q.name
                ref = O.enum_cache(g, q, c_r, O.order_ri(q), cap=3_000_000)
                assert r.ans == ref["count"], q.name
                if not ref["capped"]:
                    assert (r.nodes, r.intersections, r.fs_prunes) == \
                        (ref["nodes"], ref["intersections"], ref["fs_prunes"]), q.name


def test_labels_attached_after_upload():
    """Labels set on a Graph after it went to the device reach the device filters."""
    g = chung_lu(2000, 2.3, 8.0, seed=11)
    q = query_set(g, 1, 5, True, seed=212)[0]
    can = sgm.CandidateParam(g, q)
    assert sgm.Filter_GQL(g, q, can, build_tables=False)
    lab = np.random.default_rng(213).integers(0, 2, g.getVertexCnt())
    g.set_labels(lab)
    q.set_labels(lab[:q.getVertexCnt()])
    can2 = sgm.CandidateParam(g, q)
    ok = sgm.Filter_GQL(g, q, can2, build_tables=False)
    ok_r, c_r = O.gql(g, q)
    assert ok == ok_r
    if ok:
        _same_cands(can2, q, c_r)


def test_batch_recaptures_after_relabel():
    """A batch runner's captured labelled filter reads the graph's label tables; relabelling
    the graph so that its label-run table grows must recapture it (the old graph would read
    freed tables), and the next filter must match the oracle under the new labels."""
    g = chung_lu(2000, 2.3, 8.0, seed=3)
    rng = np.random.default_rng(214)
    g.set_labels(rng.integers(0, 2, g.getVertexCnt()))
    b = sgm.Batch(g, slots=8)
    for labels in (2, 7):  # 7 labels: more (label, count) runs per vertex than 2
        if labels != 2:
            g.set_labels(rng.integers(0, labels, g.getVertexCnt()))
            b.ctx.upload(g)  # the new labels go to the device graph
        qs = query_set(g, 8, lambda i, r: int(r.integers(3, 7)), lambda i, r: bool(i % 2), seed=215 + labels)
        for rep in range(2):
            ok, cnt, _ = b.filter(qs)
            say("relabel batch", labels, rep)
            for q, good, c in zip(qs, ok, cnt):
                ok_r, c_r = O.gql(g, q)
                assert bool(good) == ok_r, q.name
                if ok_r:
                    assert [int(x) for x in c[:q.getVertexCnt()]] == [len(x) for x in c_r], q.name


def test_relabel_makes_filtered_queries_stale():
    """Candidates filtered under the graph's old labels are refused by the enumeration
    (SGM_ESTATE) until the query is filtered again."""
    g = chung_lu(2000, 2.3, 8.0, seed=11)
    q = query_set(g, 1, 5, True, seed=216)[0]
    can = sgm.CandidateParam(g, q)
    assert sgm.Filter_GQL(g, q, can, build_tables=False)
    order = sgm.Order_RI(q)
    before = sgm.enumerate_query(can, order, cap=100000).ans
    g.set_labels(np.random.default_rng(217).integers(0, 2, g.getVertexCnt()))
    can.ctx.upload(g)
    for method in (L.SGM_ENUM_CACHE_FILTER_FS, L.SGM_ENUM_AUX_TABLES):
        with pytest.raises(L.SgmError) as e:
            sgm.enumerate_query(can, order, method=method)
        assert e.value.code == L.SGM_ESTATE
    g.set_labels(None)
    can.ctx.upload(g)
    assert sgm.Filter_GQL(g, q, can, build_tables=False)  # filtered again: usable
    assert sgm.enumerate_query(can, order, cap=100000).ans == before
