"""Seeded synthetic inputs: Chung-Lu and R-MAT data graphs, random-walk queries.

Every generator takes an explicit integer seed and uses numpy's PCG64, so the same seed
gives the same graph on every host.  Graphs are simple and undirected, returned as sorted
CSR ``Graph`` objects, unlabeled unless ``with_labels`` adds labels.  The reference ships no generator; these
definitions are this project's own (the workload sizes follow the benchmark
configurations: Chung-Lu 10^4 / 10^5 vertices, R-MAT scale 20 / 22).
"""
from __future__ import annotations

import numpy as np

from .graph import Graph, Query, csr_from_edges

__all__ = ["chung_lu", "rmat", "random_walk_query", "query_set", "with_labels"]


def _dedup_keep_first(keys: np.ndarray) -> np.ndarray:
    _, idx = np.unique(keys, return_index=True)
    return keys[np.sort(idx)]


def _grow_edge_keys(V: int, target_edges: int, sampler, rng, tol: float = 0.01):
    """Sample endpoint pairs until ``target_edges`` distinct undirected edges exist
    (within ``tol``), keeping first occurrences in sampling order, then trim."""
    keys = np.zeros(0, dtype=np.int64)
    rounds = 0
    while keys.size < target_edges * (1.0 - tol):
        need = target_edges - keys.size
        m = int(need * 1.25) + 1024
        a, b = sampler(m)
        lo = np.minimum(a, b)
        hi = np.maximum(a, b)
        ok = lo != hi
        new = lo[ok] * V + hi[ok]
        keys = _dedup_keep_first(np.concatenate([keys, new]))
        rounds += 1
        if rounds > 200:
            raise RuntimeError("edge sampler saturated before reaching the target edge count")
    return keys[:target_edges]


def _finish(V: int, keys: np.ndarray, rng, permute: bool, name: str) -> Graph:
    a = keys // V
    b = keys % V
    if permute:
        perm = rng.permutation(V).astype(np.int64)
        a = perm[a]
        b = perm[b]
    off, adj = csr_from_edges(V, np.stack([a, b], axis=1))
    return Graph(off, adj, name)


def chung_lu(num_vertices: int, gamma: float, avg_degree: float, seed: int,
             permute: bool = True) -> Graph:
    """Chung-Lu power-law graph.

    Weights w_i = (i+1)^(-1/(gamma-1)); endpoints are drawn independently with
    probability proportional to w; self loops are dropped and duplicates merged; sampling
    continues until the number of distinct edges reaches V*avg_degree/2 (the average
    degree is then exact).  Vertex ids are shuffled with the same seed so hubs are not the
    lowest ids.
    """
    V = int(num_vertices)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = (np.arange(V, dtype=np.float64) + 1.0) ** (-1.0 / (gamma - 1.0))
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    target = int(round(V * avg_degree / 2.0))

    def sampler(m):
        a = np.searchsorted(cdf, rng.random(m), side="right")
        b = np.searchsorted(cdf, rng.random(m), side="right")
        return np.minimum(a, V - 1).astype(np.int64), np.minimum(b, V - 1).astype(np.int64)

    keys = _grow_edge_keys(V, target, sampler, rng, tol=0.0)
    return _finish(V, keys, rng, permute, f"chunglu_{V}_{gamma}_{avg_degree}_s{seed}")


def rmat(scale: int, num_edges: int, seed: int, a=0.57, b=0.19, c=0.19, d=0.05,
         permute: bool = True, batch: int = 1 << 22) -> Graph:
    """R-MAT graph with 2^scale vertices and ``num_edges`` distinct undirected edges.

    Each directed pair picks one of four quadrants per bit level with probabilities
    (a, b, c, d); pairs are symmetrised, self loops dropped, duplicates merged, and
    sampling continues until ``num_edges`` distinct edges exist.  Vertex ids are shuffled
    with the same seed.
    """
    V = 1 << int(scale)
    rng = np.random.Generator(np.random.PCG64(seed))
    t_ab = a + b
    t_abc = a + b + c

    def sampler(m):
        srcs, dsts = [], []
        left = m
        while left > 0:
            k = min(left, batch)
            s = np.zeros(k, dtype=np.int64)
            t = np.zeros(k, dtype=np.int64)
            for _ in range(int(scale)):
                r = rng.random(k, dtype=np.float32)
                down = r >= t_ab                       # quadrant c or d: source bit 1
                right = ((r >= a) & (r < t_ab)) | (r >= t_abc)  # quadrant b or d: target bit 1
                s = (s << 1) | down
                t = (t << 1) | right
            srcs.append(s)
            dsts.append(t)
            left -= k
        return np.concatenate(srcs), np.concatenate(dsts)

    keys = _grow_edge_keys(V, int(num_edges), sampler, rng, tol=0.0)
    return _finish(V, keys, rng, permute, f"rmat_{scale}_{num_edges}_s{seed}")


def with_labels(g: Graph, num_labels: int, seed: int, skew: float = 0.0) -> Graph:
    """The same CSR with vertex labels 0..num_labels-1 drawn independently per vertex,
    label l with probability proportional to (l+1)^-skew (0 = uniform)."""
    rng = np.random.Generator(np.random.PCG64([seed, 0x1ABE1]))
    p = np.arange(1, num_labels + 1, dtype=np.float64) ** -float(skew)
    lab = rng.choice(num_labels, size=g.getVertexCnt(), p=p / p.sum())
    return Graph(g.off, g.adj, name=f"{g.name}_L{num_labels}", labels=lab)


def _is_dense(k: int, e: int) -> bool:
    """Dense = average degree > 3.  Below 5 vertices that is impossible (K4 has average
    degree 3), so there dense means at least k+1 edges (a diamond or K4 at k=4; the
    triangle at k=3)."""
    if k >= 5:
        return 2.0 * e / k > 3.0
    return e >= k + 1 if k == 4 else e >= k


def _walk_arrays(g: Graph):
    off = g.off.astype(np.int64)
    deg = np.diff(off)
    return off, g.adj, deg, np.flatnonzero(deg > 0)


def random_walk_query(g: Graph, k: int, dense: bool, rng: np.random.Generator,
                      max_attempts: int = 100000, _arrays=None) -> Query:
    """Extract a connected k-vertex query by a random walk.

    Start at a random vertex of degree >= 1 and step to uniform random neighbours until k
    distinct vertices are seen (restart after 100*k steps).  ``dense``: the induced
    subgraph on the walk's vertices, accepted iff its average degree is > 3 (see
    ``_is_dense`` for k < 5).  Sparse: only the traversed walk edges, accepted iff the
    average degree is <= 3 and the query is not dense.  Vertices are
    relabelled 0..k-1 in visit order, so the walk itself is always one embedding.
    """
    off, adj, deg, nz = _arrays if _arrays is not None else _walk_arrays(g)
    if nz.size == 0:
        raise ValueError("graph has no edges")
    for _ in range(max_attempts):
        start = int(nz[rng.integers(nz.size)])
        seen = {start: 0}
        order = [start]
        walk_edges = set()
        cur = start
        steps = 0
        while len(order) < k and steps < 100 * k:
            nxt = int(adj[off[cur] + rng.integers(deg[cur])])
            walk_edges.add((min(cur, nxt), max(cur, nxt)))
            if nxt not in seen:
                seen[nxt] = len(order)
                order.append(nxt)
            cur = nxt
            steps += 1
        if len(order) < k:
            continue
        if dense:
            es = []
            for i, x in enumerate(order):
                nb = adj[off[x]:off[x + 1]]
                for j in range(i + 1, k):
                    y = order[j]
                    p = np.searchsorted(nb, y)
                    if p < nb.size and nb[p] == y:
                        es.append((i, j))
            if not _is_dense(k, len(es)):
                continue
        else:
            es = [(seen[x], seen[y]) for (x, y) in sorted(walk_edges)]
            if _is_dense(k, len(es)) or 2.0 * len(es) / k > 3.0:
                continue
        qoff, qadj = csr_from_edges(k, np.array(es, dtype=np.int64).reshape(-1, 2))
        # labelled graphs: query vertex i takes the label of the walk's i-th vertex
        labels = None if g.labels is None else g.labels[np.array(order, dtype=np.int64)]
        return Query(qoff, qadj, name=f"rw{k}{'d' if dense else 's'}", labels=labels)
    raise RuntimeError(f"no {'dense' if dense else 'sparse'} {k}-vertex query found")


def query_set(g: Graph, count: int, sizes, dense, seed: int):
    """``count`` queries; ``sizes`` and ``dense`` are either scalars or callables of the
    query index.  Each query i uses its own child seed (seed, i), so subsets reproduce."""
    out = []
    arrays = _walk_arrays(g)  # shared by every walk (the same values each call would compute)
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64([seed, i]))
        k = sizes(i, rng) if callable(sizes) else int(sizes)
        dn = dense(i, rng) if callable(dense) else bool(dense)
        q = random_walk_query(g, k, dn, rng, _arrays=arrays)
        q.name = f"q{i}_{q.name}"
        out.append(q)
    return out
