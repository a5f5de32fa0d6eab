"""Pins the CPU oracle (oracle/sgm_oracle.c) to brute force, closed forms and hand-worked
known answers.  The reference ships no tests or golden vectors and does not build in this
image (it needs Eigen3 and a cmake-generated config.h), so these checks are what make the
oracle trustworthy.  Every check here is chosen so that a plausible restatement mistake
(wrong tie-break, Jacobi instead of Gauss-Seidel GQL updates, a fixpoint instead of two
rounds, a dropped backward neighbour, off-by-one depth) fails one of them.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from brute import (count_embeddings, embeddings, gql_by_hall, nlf_by_definition, prefix_nodes, query_from_edges,
                   random_connected_query, random_graph, adjsets)
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def small_cases(seed, count, V=(7, 10), p=(0.3, 0.6), n=(3, 5), labels=0):
    """Random small (graph, query) pairs; labels > 0 draws labels 0..labels-1 for both."""
    rng = np.random.default_rng(seed)
    for _ in range(count):
        g = random_graph(int(rng.integers(*V)), float(rng.uniform(*p)), rng)
        k = int(rng.integers(n[0], n[1] + 1))
        q = random_connected_query(k, int(rng.integers(0, 3)), rng)
        if labels:
            g.set_labels(rng.integers(0, labels, g.getVertexCnt()))
            q.set_labels(rng.integers(0, labels, k))
        yield g, q


# ------------------------------------------------- set intersection: against set semantics
def test_intersection_matches_sets():
    """The oracle's intersection (AVX2 when built with it, as the reference's build is:
    inc/common_ops.h:18-150, 248-254) equals both the scalar merge and Python's set
    intersection on sorted lists of every size shape: empty, single, tails shorter than a
    vector, skewed (ratio > 2: broadcast path) and balanced (the 2 x 4 path)."""
    rng = np.random.default_rng(41)
    assert O.isect_is_avx2()
    sizes = [0, 1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 64, 100, 257, 1000]
    for na in sizes:
        for nb in sizes:
            for dens in (0.05, 0.5, 0.95):
                universe = max(8, int((na + nb) / dens))
                a = np.sort(rng.choice(universe, min(na, universe), replace=False)).astype(np.uint32)
                b = np.sort(rng.choice(universe, min(nb, universe), replace=False)).astype(np.uint32)
                same, got = O.isect_check(a, b)
                assert same, (na, nb, dens)
                assert got.tolist() == sorted(set(a.tolist()) & set(b.tolist())), (na, nb, dens)


# ---------------------------------------------------------------- NLF: closed form
def test_nlf_is_degree_filter():
    g = chung_lu(500, 2.5, 6.0, seed=3)
    for q in query_set(g, 5, 6, False, seed=4) + query_set(g, 5, 7, True, seed=5):
        ok, c = O.nlf(g, q)
        deg = g.degrees()
        assert ok
        for u in range(q.getVertexCnt()):
            exp = np.flatnonzero(deg >= q.getVertexDegree(u))
            assert np.array_equal(c[u], exp)


def test_nlf_empty_set_fails():
    g = random_graph(6, 0.2, np.random.default_rng(0))
    q = query_from_edges(5, [(0, 1), (0, 2), (0, 3), (0, 4)])  # degree 4 > max data degree
    if g.getMaxDgree() < 4:
        ok, _ = O.nlf(g, q)
        assert not ok


def test_labelled_nlf_by_definition():
    """Labelled NLF (inc/filter.h:24-58): label, degree and neighbour-label counts, from
    the definition; the plausible slips (no label test, degree only, counts ignored) are
    each caught by some case."""
    g = with_labels(chung_lu(400, 2.5, 6.0, seed=31), 3, seed=32, skew=1.0)
    cases = [(g, q) for q in query_set(g, 8, 6, False, seed=33) + query_set(g, 8, 7, True, seed=34)]
    cases += list(small_cases(35, 120, labels=2))
    slips = {"no_label": 0, "no_counts": 0}
    for gg, q in cases:
        ok, c = O.nlf(gg, q)
        exp = nlf_by_definition(gg, q)
        assert ok == all(len(x) > 0 for x in exp)
        if ok:
            assert [x.tolist() for x in c] == exp
        deg = gg.degrees()
        gl, ql = gg.label_array(), q.label_array()
        for u in range(q.getVertexCnt()):
            only_deg = np.flatnonzero(deg >= q.getVertexDegree(u)).tolist()
            lab_deg = np.flatnonzero((deg >= q.getVertexDegree(u)) & (gl == ql[u])).tolist()
            slips["no_label"] += lab_deg != only_deg
            slips["no_counts"] += lab_deg != exp[u]
    assert slips["no_label"] > 0 and slips["no_counts"] > 0


# ---------------------------------------------------------------- GQL: Hall's condition
@pytest.mark.parametrize("labels", [0, 2])
def test_gql_matches_hall_restatement(labels):
    diff_jacobi = diff_rounds = 0
    for g, q in small_cases(1, 150, labels=labels):
        ok, c = O.gql(g, q)
        ok_b, c_b = gql_by_hall(g, q)
        assert ok == ok_b
        if ok:
            assert [x.tolist() for x in c] == c_b
        _, cj = gql_by_hall(g, q, jacobi=True)
        _, c3 = gql_by_hall(g, q, rounds=3)
        diff_jacobi += cj != c_b
        diff_rounds += c3 != c_b
    # the suite must be able to tell the schedule apart from its plausible mistakes
    assert diff_jacobi > 0 or diff_rounds > 0


def test_gql_schedule_distinguishing_cases():
    """Search seeded small cases until both schedule mistakes are exposed, and check the
    oracle agrees with the reference schedule on them."""
    seen_j = seen_r = False
    for g, q in small_cases(7, 2000, V=(6, 9), p=(0.25, 0.5), n=(4, 5)):
        ok_b, c_b = gql_by_hall(g, q)
        _, cj = gql_by_hall(g, q, jacobi=True)
        _, c3 = gql_by_hall(g, q, rounds=3)
        if cj != c_b or c3 != c_b:
            ok, c = O.gql(g, q)
            assert ok == ok_b
            if ok:
                assert [x.tolist() for x in c] == c_b
            seen_j |= cj != c_b
            seen_r |= c3 != c_b
        if seen_j and seen_r:
            break
    assert seen_j and seen_r


def test_gql_sound_and_subset_of_nlf():
    for g, q in small_cases(2, 60):
        ok, c = O.gql(g, q)
        _, cn = O.nlf(g, q)
        emb = list(embeddings(g, q))
        if not ok:
            assert not emb
            continue
        for u in range(q.getVertexCnt()):
            assert set(c[u].tolist()) <= set(cn[u].tolist())
            for m in emb:
                assert m[u] in set(c[u].tolist())


# ---------------------------------------------------------------- tables: definition
def test_tables_by_definition():
    for g, q in small_cases(3, 30):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        A = adjsets(g)
        for a in range(q.getVertexCnt()):
            for b in q.getNeb(a).tolist():
                off, idx, vv, md = O.table(g, c[a], c[b])
                lens = []
                for i, va in enumerate(c[a].tolist()):
                    exp = [k for k, vb in enumerate(c[b].tolist()) if vb in A[va]]
                    got = idx[off[i]:off[i + 1]].tolist()
                    assert got == exp
                    assert vv[off[i]:off[i + 1]].tolist() == [int(c[b][k]) for k in exp]
                    lens.append(len(exp))
                assert md == max(lens + [0])


def test_tables_transpose_symmetry():
    g = chung_lu(400, 2.5, 6.0, seed=9)
    for q in query_set(g, 4, 6, True, seed=10):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        for a in range(q.getVertexCnt()):
            for b in q.getNeb(a).tolist():
                o1, i1, _, _ = O.table(g, c[a], c[b])
                o2, i2, _, _ = O.table(g, c[b], c[a])
                fwd = {(i, int(k)) for i in range(len(c[a])) for k in i1[o1[i]:o1[i + 1]]}
                bwd = {(int(k), j) for j in range(len(c[b])) for k in i2[o2[j]:o2[j + 1]]}
                assert fwd == bwd


# ---------------------------------------------------------------- orders: known answers
PATH4 = [(0, 1), (1, 2), (2, 3)]
TRI_TAIL = [(0, 1), (1, 2), (0, 2), (2, 3)]
STAR3 = [(0, 1), (0, 2), (0, 3)]


def test_order_ri_known_answers():
    # root: max degree, lowest id on ties (1 over 2); step 1 tie {0,2} broken by
    # tie-break 2 (vertex 2 has an unvisited neighbour adjacent to no ordered vertex)
    assert O.order_ri(query_from_edges(4, PATH4)) == [1, 2, 0, 3]
    # root 2 (degree 3); step 1 ties {0,1,3} on backward neighbours, tie-break 1 drops 3,
    # tie-break 2 leaves {0,1} -> lowest id 0; then 1 (two backward neighbours), then 3
    assert O.order_ri(query_from_edges(4, TRI_TAIL)) == [2, 0, 1, 3]
    # star: centre first, then leaves in id order
    assert O.order_ri(query_from_edges(4, STAR3)) == [0, 1, 2, 3]
    # square with one diagonal 1-3: roots tie at degree 3 -> 1; next 3 (tie {0,2,3} on
    # one backward neighbour; tie-break 1: 0 and 2 see {3} adjacent to 1 -> count 1 each,
    # 3 has unvisited nbrs {0,2}, both adjacent to 1 -> count 1; tie-break 2: all 0 -> 0)
    sq = [(0, 1), (1, 2), (2, 3), (0, 3), (1, 3)]
    assert O.order_ri(query_from_edges(4, sq)) == [1, 0, 3, 2]


def test_order_gql_known_answers():
    assert O.order_gql(_g(10), query_from_edges(4, PATH4), [5, 3, 3, 7]) == [1, 2, 0, 3]
    assert O.order_gql(_g(10), query_from_edges(4, STAR3), [4, 4, 2, 2]) == [2, 0, 3, 1]
    assert O.order_gql(_g(10), query_from_edges(4, TRI_TAIL), [5, 5, 5, 5]) == [2, 0, 1, 3]


def _g(V):
    return random_graph(V, 0.5, np.random.default_rng(0))


def test_orders_are_connected_permutations():
    g = chung_lu(800, 2.5, 6.0, seed=12)
    for q in query_set(g, 10, lambda i, r: int(r.integers(4, 17)), lambda i, r: bool(i % 2), seed=13):
        ok, c = O.gql(g, q)
        for o in (O.order_ri(q), O.order_gql(g, q, [len(x) for x in c])):
            assert sorted(o) == list(range(q.getVertexCnt()))
            for i in range(1, len(o)):
                assert any(q.hasEdge(o[i], o[j]) for j in range(i))


# ---------------------------------------------------------------- PlanIR: known answers
def _ops(p):
    return [(o["depth"], o["edges"], o["parent"], o["vdepth"]) for o in p["ops"]]


def test_plan_known_answers():
    tri = query_from_edges(3, [(0, 1), (1, 2), (0, 2)])
    p = O.plan(tri, [0, 1, 2])
    assert _ops(p) == [(0, [0], None, 1), (1, [0, 1], 0, 2)]
    assert p["iter_set"] == [None, 0, 1]
    p = O.plan(query_from_edges(4, PATH4), [1, 2, 0, 3])
    assert _ops(p) == [(0, [0], None, 1), (1, [1], None, 3)]
    assert p["iter_set"] == [None, 0, 0, 1]
    k4 = query_from_edges(4, [(a, b) for a in range(4) for b in range(a + 1, 4)])
    p = O.plan(k4, [0, 1, 2, 3])
    assert _ops(p) == [(0, [0], None, 1), (1, [0, 1], 0, 2), (2, [0, 1, 2], 1, 3)]
    assert p["iter_set"] == [None, 0, 1, 2]


def test_plan_labelled_known_answers():
    """Ops are keyed by (label of the follow vertex, edge list) (inc/code_gen.h:640-653):
    equal edge lists with different labels stay apart, sort by label before size, and a
    parent must carry the same label (:777-787)."""
    star = query_from_edges(3, [(0, 1), (0, 2)])
    assert _ops(O.plan(star, [0, 1, 2])) == [(0, [0], None, 1)]
    lab = query_from_edges(3, [(0, 1), (0, 2)], labels=[0, 5, 3])
    p = O.plan(lab, [0, 1, 2])
    assert _ops(p) == [(0, [0], None, 2), (0, [0], None, 1)]
    assert p["iter_set"] == [None, 1, 0]
    # triangle 0-1-2 plus pendant 3 on 0; vertex 2's label differs from 1's and 3's
    es = [(0, 1), (0, 2), (1, 2), (0, 3)]
    p = O.plan(query_from_edges(4, es, labels=[0, 1, 2, 1]), [0, 1, 2, 3])
    assert _ops(p) == [(0, [0], None, 1), (0, [0], None, 2), (1, [0, 1], 1, 2)]
    assert p["iter_set"] == [None, 0, 2, 0]
    p = O.plan(query_from_edges(4, es), [0, 1, 2, 3])
    assert _ops(p) == [(0, [0], None, 1), (1, [0, 1], 0, 2)]
    assert p["iter_set"] == [None, 0, 1, 0]


def test_plan_invariants():
    g = chung_lu(800, 2.5, 6.0, seed=14)
    for q in query_set(g, 12, lambda i, r: int(r.integers(4, 33)), lambda i, r: bool(i % 2), seed=15):
        order = O.order_ri(q)
        p = O.plan(q, order)
        pos = {v: i for i, v in enumerate(order)}
        ops = p["ops"]
        # unique edge sets; depth = max edge index; parents are the prefix
        assert len({tuple(o["edges"]) for o in ops}) == len(ops)
        for o in ops:
            assert o["depth"] == max(o["edges"])
            if o["parent"] is None:
                assert len(o["edges"]) == 1
            else:
                assert ops[o["parent"]]["edges"] == o["edges"][:-1]
        for d in range(1, q.getVertexCnt()):
            back = sorted(pos[x] for x in q.getNeb(order[d]).tolist() if pos[x] < d)
            assert ops[p["iter_set"][d]]["edges"] == back


def test_op_sort_matches_libstdcxx():
    """The oracle restates std::sort's exact steps; equal keys must land in the same
    relative order as the library routine (this decides the kept _v_depth)."""
    rng = np.random.default_rng(16)
    for trial in range(300):
        n = int(rng.integers(1, 700)) if trial % 3 else int(rng.integers(1, 40))
        depth = rng.integers(0, 4, n)
        size = rng.integers(1, 4, n)
        elems = np.zeros((n, 32), dtype=np.uint32)
        for i in range(n):
            elems[i, :size[i]] = np.sort(rng.choice(5, size[i], replace=False))
        a = O.sort_ops_perm(depth, size, elems)
        b = O.sort_ops_perm(depth, size, elems, use_std=True)
        assert a == b


# ---------------------------------------------------------------- enumeration: counts
@pytest.mark.parametrize("labels", [0, 2])
def test_counts_match_brute_force(labels):
    rng = np.random.default_rng(17)
    for g, q in small_cases(4, 80, labels=labels):
        exp = count_embeddings(g, q)
        ok, c = O.gql(g, q)
        if not ok:
            assert exp == 0
            continue
        for order in (O.order_ri(q), O.order_gql(g, q, [len(x) for x in c])):
            assert O.enum_cache(g, q, c, order)["count"] == exp
            assert O.enum_cache(g, q, c, order, fs=False)["count"] == exp
            assert O.enum_all(g, q, c, order)["count"] == exp
            assert O.enum_all(g, q, c, order, fs=True)["count"] == exp
            if exp > 1:
                cap = int(rng.integers(1, exp))
                assert O.enum_cache(g, q, c, order, cap=cap)["count"] == cap
                assert O.enum_all(g, q, c, order, cap=cap)["count"] == cap


def test_root_range_counts_match_brute_force():
    """A root range [lo, hi) counts exactly the embeddings whose first order vertex maps into
    it (brute force over injective maps), on both searches, failing sets on and off."""
    rng = np.random.default_rng(23)
    for g, q in small_cases(24, 40):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        V = g.getVertexCnt()
        lo = int(rng.integers(0, V - 1))
        hi = int(rng.integers(lo + 1, V + 1))
        exp = sum(1 for m in embeddings(g, q) if lo <= m[order[0]] < hi)
        for fs in (True, False):
            assert O.enum_cache(g, q, c, order, fs=fs, roots=(lo, hi))["count"] == exp
            assert O.enum_all(g, q, c, order, fs=fs, roots=(lo, hi))["count"] == exp


def test_root_ranges_add_up():
    """Disjoint root ranges add up to the whole search in every statistic (count, per-depth
    nodes, intersections, failing-set cut-offs): the depth-0 loop never prunes, so each
    subtree is searched alike whatever else runs."""
    g = chung_lu(300, 2.3, 5.0, seed=19)  # inputs on which failing sets cut (see below)
    V = g.getVertexCnt()
    rng = np.random.default_rng(31)
    qs = query_set(g, 16, lambda i, r: int(r.integers(5, 9)), lambda i, r: bool(i % 2), seed=20)
    cut = 0
    for q in qs:
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        cuts = sorted(int(x) for x in rng.choice(np.arange(1, V), 3, replace=False))
        ranges = list(zip([0] + cuts, cuts + [V]))
        for fn, fs in ((O.enum_cache, True), (O.enum_cache, False), (O.enum_all, True), (O.enum_all, False)):
            whole = fn(g, q, c, order, fs=fs, cap=3_000_000)
            if whole["capped"]:
                continue
            parts = [fn(g, q, c, order, fs=fs, roots=r) for r in ranges]
            assert sum(p["count"] for p in parts) == whole["count"]
            assert sum(p["intersections"] for p in parts) == whole["intersections"]
            assert sum(p["fs_prunes"] for p in parts) == whole["fs_prunes"]
            assert [sum(x) for x in zip(*[p["nodes"] for p in parts])] == whole["nodes"]
            cut += whole["fs_prunes"] if fs else 0
    assert cut > 0  # the failing sets did cut inside some subtree


def test_counts_closed_forms():
    g = chung_lu(300, 2.5, 8.0, seed=18)
    deg = g.degrees().astype(np.int64)
    A = adjsets(g)
    tri = sum(len(A[a] & A[b]) for a in range(300) for b in A[a] if a < b) // 3
    cases = [
        (query_from_edges(3, [(0, 1), (1, 2), (0, 2)]), 6 * tri),
        (query_from_edges(3, [(0, 1), (1, 2)]), int((deg * (deg - 1)).sum())),
        (query_from_edges(4, STAR3), int((deg * (deg - 1) * (deg - 2)).sum())),
    ]
    for q, exp in cases:
        ok, c = O.gql(g, q)
        order = O.order_ri(q)
        assert O.enum_cache(g, q, c, order)["count"] == exp
        assert O.enum_all(g, q, c, order)["count"] == exp


# ---------------------------------------------------------------- node / intersection counts
def _op_sets(g, p, emb_by_depth, depth):
    A = adjsets(g)
    res = []
    for o in p["ops"]:
        if o["depth"] != depth:
            continue
        s = None
        for e in o["edges"]:
            s = set(A[emb_by_depth[e]]) if s is None else s & A[emb_by_depth[e]]
        res.append((o, s))
    return res


def test_aux_path_nodes_and_intersections_brute():
    for g, q in small_cases(5, 60):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        st = O.enum_all(g, q, c, order)
        nodes = prefix_nodes(g, q, [x.tolist() for x in c], order)
        assert st["nodes"] == nodes
        pos = {v: i for i, v in enumerate(order)}
        nprev = [sum(1 for x in q.getNeb(order[d]).tolist() if pos[x] < d) for d in range(len(order))]
        assert st["intersections"] == sum(nodes[d - 1] * (nprev[d] - 1) for d in range(1, len(order)))


@pytest.mark.parametrize("labels", [0, 2])
def test_cache_path_nodes_and_intersections_brute(labels):
    """Cache + filter with failing sets off: a depth-d node exists iff the prefix is a
    valid partial embedding and every op of every shallower depth was non-empty; each
    visit of depth d evaluates depth d-1's ops in plan order up to the first empty one.
    Ops intersect whole neighbour lists also for labelled queries (inc/explore.h:214)."""
    for g, q in small_cases(6, 60, labels=labels):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        p = O.plan(q, order)
        n = len(order)
        A, QA = adjsets(g), adjsets(q)
        cs = [set(x.tolist()) for x in c]
        nodes = [0] * n
        inter = 0

        def rec(d, emb):
            nonlocal inter
            if d > 0:
                for o, s in _op_sets(g, p, emb, d - 1):
                    if o["parent"] is not None:
                        inter += 1
                    if not s:
                        return
            u = order[d]
            pool = cs[u] if d == 0 else None
            for v in (sorted(pool) if d == 0 else sorted(set.intersection(*[A[emb[j]] for j in range(d) if order[j] in QA[u]]))):
                if v not in cs[u] or v in emb:
                    continue
                nodes[d] += 1
                if d + 1 < n:
                    rec(d + 1, emb + [v])
        rec(0, [])
        st = O.enum_cache(g, q, c, order, fs=False)
        assert st["nodes"] == nodes
        assert st["intersections"] == inter


def test_failing_sets_only_prune():
    """Failing sets never change counts and never add nodes; on some inputs they cut."""
    g = chung_lu(300, 2.3, 5.0, seed=19)
    cut = acut = 0
    for q in query_set(g, 16, lambda i, r: int(r.integers(5, 9)), lambda i, r: bool(i % 2), seed=20):
        ok, c = O.gql(g, q)
        if not ok:
            continue
        order = O.order_ri(q)
        on = O.enum_cache(g, q, c, order, cap=2000000)
        off = O.enum_cache(g, q, c, order, fs=False, cap=2000000)
        if on["capped"] or off["capped"]:
            continue
        assert on["count"] == off["count"]
        assert all(a <= b for a, b in zip(on["nodes"], off["nodes"]))
        cut += on["fs_prunes"]
        # the same on the aux-table search (explore_Intersection_All<true,...>)
        aon = O.enum_all(g, q, c, order, fs=True, cap=2000000)
        aoff = O.enum_all(g, q, c, order, cap=2000000)
        if not (aon["capped"] or aoff["capped"]):
            assert aon["count"] == aoff["count"] == off["count"]
            assert all(a <= b for a, b in zip(aon["nodes"], aoff["nodes"]))
            assert aon["intersections"] <= aoff["intersections"]
            acut += aon["fs_prunes"]
    assert cut > 0 and acut > 0


def test_aux_daf_hand_worked():
    """Failing sets on the aux-table search with hand-chosen candidate sets.  Data: star
    0-{1,2,3} plus edge 4-5; query: star u0-{u1,u2}; order (u0, u1, u2); anc = {0}, {0,1},
    {0,2}.  (a) C(u2) = {4}: the first u1 candidate's child has no candidates, fail[1] =
    anc[2] = {0,2} lacks depth 1, so the other two u1 candidates are cut (nodes 1,1,0; one
    prune; without failing sets 1,3,0).  (b) C(u2) = {1}: the child of u1 = 1 meets an
    injectivity conflict, fail[2] = anc[2] | anc[1] holds depth 1, no cut; u1 = 2 then
    matches (nodes 1,2,1; count 1)."""
    g = query_from_edges(6, [(0, 1), (0, 2), (0, 3), (4, 5)])
    q = query_from_edges(3, [(0, 1), (0, 2)])
    arr = lambda *x: np.array(x, dtype=np.uint32)  # noqa: E731
    st = O.enum_all(g, q, [arr(0), arr(1, 2, 3), arr(4)], [0, 1, 2], fs=True)
    assert (st["count"], st["nodes"], st["fs_prunes"]) == (0, [1, 1, 0], 1)
    st = O.enum_all(g, q, [arr(0), arr(1, 2, 3), arr(4)], [0, 1, 2])
    assert (st["count"], st["nodes"], st["fs_prunes"]) == (0, [1, 3, 0], 0)
    st = O.enum_all(g, q, [arr(0), arr(1, 2), arr(1)], [0, 1, 2], fs=True)
    assert (st["count"], st["nodes"], st["fs_prunes"]) == (1, [1, 2, 1], 0)


def test_daf_hand_worked():
    """Two hand-worked failing-set cut-offs with hand-chosen candidate sets (the
    enumerator accepts any candidate sets; they define the search space).

    Case 1 (all-filtered rule, inc/explore.h:1068-1070): query c-a, c-b (c=0,a=1,b=2),
    order [0,1,2], data star 0-{1,2,3} plus edge 1-4, C = [{0},{1,2,3},{4}].  At depth 2
    every element of N(M0)={1,2,3} fails the filter, so fail[1] = anc(2) = {0,2}; it lacks
    depth 1, so the two remaining depth-1 siblings are cut: nodes [1,1,0] vs [1,3,0].

    Case 2 (empty-set rule with conf_dep, inc/explore.h:225-240): query edges 0-1, 0-2,
    0-3, 2-3, order [0,1,2,3]; plan ops {0}@0 and {0,2}@2 (v_depth 3); data star
    0-{1,2,3}; C = [{0},{1,2,3},{2,3},{1,2,3}].  With M1=1 both depth-2 children make
    {0,2} empty, each setting fail[2] = anc(3) = {0,2,3}; the union lacks depth 1, so
    M1=2,3 are cut: nodes [1,1,2,0] and 2 intersections (fs off: [1,3,4,0], 4).
    """
    g = query_from_edges(5, [(0, 1), (0, 2), (0, 3), (1, 4)])
    q = query_from_edges(3, [(0, 1), (0, 2)])
    c = [np.array([0]), np.array([1, 2, 3]), np.array([4])]
    on = O.enum_cache(g, q, c, [0, 1, 2])
    off = O.enum_cache(g, q, c, [0, 1, 2], fs=False)
    assert on["count"] == off["count"] == 0
    assert on["nodes"] == [1, 1, 0] and on["fs_prunes"] == 1
    assert off["nodes"] == [1, 3, 0]

    g = query_from_edges(4, [(0, 1), (0, 2), (0, 3)])
    q = query_from_edges(4, [(0, 1), (0, 2), (0, 3), (2, 3)])
    p = O.plan(q, [0, 1, 2, 3])
    assert _ops(p) == [(0, [0], None, 1), (2, [0, 2], 0, 3)]
    c = [np.array([0]), np.array([1, 2, 3]), np.array([2, 3]), np.array([1, 2, 3])]
    on = O.enum_cache(g, q, c, [0, 1, 2, 3])
    off = O.enum_cache(g, q, c, [0, 1, 2, 3], fs=False)
    assert on["count"] == off["count"] == 0
    assert on["nodes"] == [1, 1, 2, 0] and on["intersections"] == 2 and on["fs_prunes"] == 1
    assert off["nodes"] == [1, 3, 4, 0] and off["intersections"] == 4
