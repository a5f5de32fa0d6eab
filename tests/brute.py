"""Brute-force checkers used to pin the oracle (pure Python, tiny inputs only).

These are independent of the oracle's algorithms: embeddings by exhaustive injective
maps (labels must agree), NLF by its definition (label, degree and neighbour-label
counts), the GQL local matching by Hall's condition over every subset of the left side,
tables by an adjacency test on every pair.
"""
from __future__ import annotations

import itertools
from collections import Counter

import numpy as np

from subgraphmatching_amd.graph import Graph, Query


def adjsets(g):
    return [set(g.getNeb(v).tolist()) for v in range(g.getVertexCnt())]


def count_embeddings(g, q):
    """Number of injective, label-preserving maps q -> g preserving every query edge
    (non-induced)."""
    return sum(1 for _ in embeddings(g, q))


def embeddings(g, q):
    A = adjsets(g)
    QA = adjsets(q)
    n, V = q.getVertexCnt(), g.getVertexCnt()
    gl, ql = g.label_array().tolist(), q.label_array().tolist()
    for m in itertools.permutations(range(V), n):
        if all(gl[m[a]] == ql[a] for a in range(n)) and \
                all(m[b] in A[m[a]] for a in range(n) for b in QA[a] if a < b):
            yield m


def nlf_by_definition(g, q):
    """NLF candidate sets from the definition: same label, deg(v) >= deg_q(u), and for each
    label the neighbours of v carry at least as often as u's do."""
    A = adjsets(g)
    gl, ql = g.label_array().tolist(), q.label_array().tolist()
    out = []
    for u in range(q.getVertexCnt()):
        need = Counter(ql[w] for w in q.getNeb(u).tolist())
        c = []
        for v in range(g.getVertexCnt()):
            if gl[v] != ql[u] or len(A[v]) < q.getVertexDegree(u):
                continue
            have = Counter(gl[w] for w in A[v])
            if all(have[l] >= k for l, k in need.items()):
                c.append(v)
        out.append(c)
    return out


def hall_ok(left_nbrs):
    """True iff a matching saturating the left side exists (Hall's condition)."""
    L = len(left_nbrs)
    for r in range(1, L + 1):
        for S in itertools.combinations(range(L), r):
            if len(set().union(*[left_nbrs[i] for i in S])) < r:
                return False
    return True


def gql_by_hall(g, q, jacobi=False, rounds=2):
    """The reference's GQL schedule (NLF, 2 rounds, u ascending, removals visible at
    once) with the local matching decided by Hall's condition.  ``jacobi`` and
    ``rounds`` exist only so tests can show the schedule matters."""
    A = adjsets(g)
    n = q.getVertexCnt()
    cands = nlf_by_definition(g, q)
    if any(len(c) == 0 for c in cands):
        return False, cands
    valid = [set(c) for c in cands]
    for _ in range(rounds):
        snapshot = [set(s) for s in valid]
        for u in range(n):
            qn = q.getNeb(u).tolist()
            src = snapshot if jacobi else valid
            for v in list(cands[u]):
                if v not in valid[u]:
                    continue
                dn = sorted(A[v])
                left = [set(j for j, w in enumerate(dn) if w in src[un]) for un in qn]
                if not hall_ok(left):
                    valid[u].discard(v)
    out = [sorted(valid[u]) for u in range(n)]
    return all(len(c) > 0 for c in out), out


def prefix_nodes(g, q, cands, order):
    """Per-depth node counts of the plain (no pruning) candidate-space search: maps of
    order[0..d] into the candidate sets, injective, preserving query edges among them."""
    A = adjsets(g)
    QA = adjsets(q)
    n = q.getVertexCnt()
    cs = [set(c) for c in cands]
    nodes = [0] * n

    def rec(d, m):
        u = order[d]
        for v in cands[u]:
            if v in m.values():
                continue
            if any(m[order[j]] not in A[v] for j in range(d) if order[j] in QA[u]):
                continue
            nodes[d] += 1
            if d + 1 < n:
                m[u] = v
                rec(d + 1, m)
                del m[u]
    rec(0, {})
    return nodes


def random_graph(V, p, rng):
    es = [(a, b) for a in range(V) for b in range(a + 1, V) if rng.random() < p]
    return Graph.from_edges(V, np.array(es, dtype=np.int64).reshape(-1, 2))


def random_connected_query(n, extra, rng):
    """Random spanning tree on n vertices plus ``extra`` random additional edges."""
    es = set()
    for v in range(1, n):
        es.add((int(rng.integers(v)), v))
    for _ in range(extra):
        a, b = sorted(rng.choice(n, 2, replace=False).tolist())
        es.add((a, b))
    return query_from_edges(n, sorted(es))


def query_from_edges(n, es, labels=None):
    g = Graph.from_edges(n, np.array(es, dtype=np.int64).reshape(-1, 2))
    return Query(g.off, g.adj, labels=labels)
