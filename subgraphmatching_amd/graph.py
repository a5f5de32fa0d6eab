"""Host-side graph containers and the reference text format.

``Graph`` / ``Query`` mirror the reference's ``Graph`` / ``Query`` classes
(inc/graph.h:51-378, :380-495): an undirected simple graph held as a CSR with each
adjacency list sorted ascending (inc/graph.h:259-260), plus optional vertex labels
(``labels`` is None when every label is 0).  Labels are u32 values; the reference numbers
them 0..L-1 (inc/graph.h:93-112).

Text format (README.md:64-84, parse rules inc/graph.h:190-266)::

    t <V> <E>
    v <id> <label> <degree>      one line per vertex, ids 0..V-1 in order
    e <u> <v>                    one line per undirected edge

The reference fills adjacency slots from the ``v`` lines' degree fields, so those must be
exact; ``load_text`` checks that instead of trusting it.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

__all__ = ["Graph", "Query", "csr_from_edges", "load_cached", "load_csr", "load_text", "save_csr", "save_text"]


def csr_from_edges(num_vertices: int, edges: np.ndarray):
    """Build a sorted, deduplicated, self-loop-free CSR from an (m, 2) edge array.

    Returns ``(off, adj)`` as uint32 arrays (``off`` has V+1 entries, ``adj`` 2|E|).
    """
    V = int(num_vertices)
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    if e.size:
        if e.min() < 0 or e.max() >= V:
            raise ValueError("edge endpoint out of range")
    a = np.minimum(e[:, 0], e[:, 1])
    b = np.maximum(e[:, 0], e[:, 1])
    keep = a != b
    key = np.unique(a[keep] * V + b[keep])
    a = key // V
    b = key % V
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    order = np.lexsort((dst, src))
    src = src[order]
    dst = dst[order]
    deg = np.bincount(src, minlength=V)
    off = np.zeros(V + 1, dtype=np.int64)
    np.cumsum(deg, out=off[1:])
    if off[-1] >= 2**32:
        raise ValueError("graph too large for 32-bit offsets")
    return off.astype(np.uint32), dst.astype(np.uint32)


class Graph:
    """Undirected simple graph in sorted CSR form (inc/graph.h:51-378)."""

    def __init__(self, off: np.ndarray, adj: np.ndarray, name: str = "", labels=None):
        self.off = np.ascontiguousarray(off, dtype=np.uint32)
        self.adj = np.ascontiguousarray(adj, dtype=np.uint32)
        self.name = name
        if self.off.ndim != 1 or self.off.size < 1:
            raise ValueError("offsets must be a 1-D array of V+1 entries")
        if int(self.off[-1]) != self.adj.size:
            raise ValueError("offsets do not cover the adjacency array")
        self.labels = None
        self.set_labels(labels)

    def set_labels(self, labels):
        """Attach vertex labels (V values; None or all zero = unlabeled)."""
        if labels is None:
            self.labels = None
            return self
        lab = np.ascontiguousarray(labels, dtype=np.int64)
        if lab.shape != (self.getVertexCnt(),):
            raise ValueError(f"expected {self.getVertexCnt()} labels, got shape {lab.shape}")
        if lab.size and (lab.min() < 0 or lab.max() > 0xFFFFFFFE):
            raise ValueError("labels must be in [0, 2^32 - 1)")
        self.labels = lab.astype(np.uint32) if np.any(lab) else None
        return self

    def is_labelled(self) -> bool:
        return self.labels is not None

    @classmethod
    def from_edges(cls, num_vertices: int, edges, name: str = "", labels=None):
        off, adj = csr_from_edges(num_vertices, np.asarray(edges))
        return cls(off, adj, name, labels=labels)

    # reference accessors -------------------------------------------------------------
    def getVertexCnt(self) -> int:  # inc/graph.h:291
        return int(self.off.size - 1)

    def getEdgeCnt(self) -> int:  # inc/graph.h:299
        return int(self.adj.size // 2)

    def getVertexDegree(self, v: int) -> int:  # inc/graph.h:295
        return int(self.off[v + 1] - self.off[v])

    def getMaxDgree(self) -> int:  # inc/graph.h:307
        d = self.degrees()
        return int(d.max()) if d.size else 0

    def getNeb(self, v: int) -> np.ndarray:  # inc/graph.h:347-350
        return self.adj[self.off[v]:self.off[v + 1]]

    def hasEdge(self, u: int, v: int) -> bool:  # inc/graph.h:376-378
        nb = self.getNeb(u)
        i = int(np.searchsorted(nb, v))
        return i < nb.size and int(nb[i]) == v

    def getVertexLabel(self, v: int) -> int:  # inc/graph.h:359
        return 0 if self.labels is None else int(self.labels[v])

    def label_array(self) -> np.ndarray:
        """Labels of every vertex (zeros when unlabeled)."""
        if self.labels is None:
            return np.zeros(self.getVertexCnt(), dtype=np.uint32)
        return self.labels

    # helpers ---------------------------------------------------------------------------
    def degrees(self) -> np.ndarray:
        return np.diff(self.off.astype(np.int64))

    def edges(self) -> np.ndarray:
        """Undirected edge list (u < v), sorted."""
        src = np.repeat(np.arange(self.getVertexCnt(), dtype=np.int64), self.degrees())
        dst = self.adj.astype(np.int64)
        m = src < dst
        return np.stack([src[m], dst[m]], axis=1)

    def digest(self) -> str:
        """SHA-256 of V, the CSR and (for labelled graphs) the labels."""
        h = hashlib.sha256()
        h.update(np.uint64(self.getVertexCnt()).tobytes())
        h.update(self.off.tobytes())
        h.update(self.adj.tobytes())
        if self.labels is not None:
            h.update(self.labels.tobytes())
        return h.hexdigest()

    def Load(self, path: str):  # inc/graph.h:190-266
        g = load_text(path)
        self.off, self.adj, self.name, self.labels = g.off, g.adj, g.name, g.labels
        return self


class Query(Graph):
    """Query graph (inc/graph.h:380-495).  At most 32 vertices on the accelerated path."""

    MAX_VERTICES = 32

    def __init__(self, off, adj, name: str = "", labels=None):
        super().__init__(off, adj, name, labels=labels)
        if self.getVertexCnt() > self.MAX_VERTICES:
            raise ValueError(f"query has {self.getVertexCnt()} vertices; at most 32 are supported")

    def is_connected(self) -> bool:
        n = self.getVertexCnt()
        if n == 0:
            return False
        seen = {0}
        stack = [0]
        while stack:
            x = stack.pop()
            for y in self.getNeb(x).tolist():
                if y not in seen:
                    seen.add(y)
                    stack.append(y)
        return len(seen) == n


def load_text(path: str, cls=Graph):
    """Read the reference text format (inc/graph.h:190-266), labels included."""
    V = E = None
    decl_deg = None
    labels = None
    edges = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "t":
                V, E = int(parts[1]), int(parts[2])
                decl_deg = np.zeros(V, dtype=np.int64)
                labels = np.zeros(V, dtype=np.int64)
            elif parts[0] == "v":
                vid, lab, d = int(parts[1]), int(parts[2]), int(parts[3])
                if not 0 <= vid < V:
                    raise ValueError(f"{path}: vertex id {vid} out of range")
                labels[vid] = lab
                decl_deg[vid] = d
            elif parts[0] == "e":
                edges.append((int(parts[1]), int(parts[2])))
    if V is None:
        raise ValueError(f"{path}: missing 't' header")
    g = cls.from_edges(V, np.array(edges, dtype=np.int64).reshape(-1, 2),
                       name=os.path.splitext(os.path.basename(path))[0], labels=labels)
    if g.getEdgeCnt() != E or not np.array_equal(g.degrees(), decl_deg):
        raise ValueError(f"{path}: header/degree fields disagree with the edge list "
                         "(the reference requires exact degrees and a simple graph)")
    return g


def save_text(g: Graph, path: str):
    """Write the reference text format."""
    V = g.getVertexCnt()
    deg = g.degrees()
    lab = g.label_array()
    e = g.edges()
    with open(path, "w") as f:
        f.write(f"t {V} {e.shape[0]}\n")
        f.write("".join(f"v {i} {int(lab[i])} {int(deg[i])}\n" for i in range(V)))
        f.write("".join(f"e {int(a)} {int(b)}\n" for a, b in e))


# ---------------------------------------------------------------------------------------
# Binary CSR cache.  Text graphs parse slowly (the R-MAT 2^22 graph has 2^26 edge lines); the
# binary form is the CSR itself, little-endian:
#   "SGMCSR01" | u64 V | u64 A (adjacency length, 2E) | u64 flags (bit 0: labels follow)
#   | u32 off[V+1] | u32 adj[A] | u32 labels[V] (if flagged)
# with a JSON manifest beside it (<path>.json): format, V, E, the SHA-256 of the CSR
# (Graph.digest) and whatever the writer records about the input (generator, parameters,
# seed, source file).
CSR_MAGIC = b"SGMCSR01"


def save_csr(g: Graph, path: str, meta: dict | None = None) -> dict:
    """Write g as a binary CSR plus its manifest; returns the manifest."""
    V = g.getVertexCnt()
    off = np.ascontiguousarray(g.off, dtype=np.uint32)
    adj = np.ascontiguousarray(g.adj, dtype=np.uint32)
    with open(path, "wb") as f:
        f.write(CSR_MAGIC)
        f.write(np.array([V, adj.size, 1 if g.labels is not None else 0], dtype="<u8").tobytes())
        f.write(off.astype("<u4").tobytes())
        f.write(adj.astype("<u4").tobytes())
        if g.labels is not None:
            f.write(g.labels.astype("<u4").tobytes())
    manifest = {"format": "sgm-csr-1", "vertices": int(V), "edges": int(g.getEdgeCnt()), "sha256": g.digest(),
                "name": g.name}
    manifest.update(meta or {})
    with open(path + ".json", "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    return manifest


def load_csr(path: str, cls=Graph, verify: bool = True):
    """Read a binary CSR written by save_csr.  With ``verify``: offsets monotone from 0 to the
    adjacency length, neighbour ids in range, every list strictly ascending, an even
    adjacency length, and the SHA-256 against the manifest (when present)."""
    with open(path, "rb") as f:
        head = f.read(32)
        if len(head) != 32 or head[:8] != CSR_MAGIC:
            raise ValueError(f"{path}: not an sgm binary CSR")
        V, A, flags = (int(x) for x in np.frombuffer(head[8:], dtype="<u8"))
        off = np.fromfile(f, dtype="<u4", count=V + 1)
        adj = np.fromfile(f, dtype="<u4", count=A)
        if off.size != V + 1 or adj.size != A:
            raise ValueError(f"{path}: truncated")
        labels = None
        if flags & 1:
            labels = np.fromfile(f, dtype="<u4", count=V)
            if labels.size != V:
                raise ValueError(f"{path}: truncated labels")
            if labels.size and labels.max() == 0xFFFFFFFF:
                raise ValueError(f"{path}: label out of range")
    g = cls(off.astype(np.uint32), adj.astype(np.uint32), name=os.path.splitext(os.path.basename(path))[0],
            labels=labels)
    if verify:
        if off[0] != 0 or off[-1] != A or np.any(np.diff(off.astype(np.int64)) < 0):
            raise ValueError(f"{path}: offsets are not a CSR")
        if A and (int(adj.max()) >= V):
            raise ValueError(f"{path}: neighbour id out of range")
        if A % 2:
            raise ValueError(f"{path}: odd adjacency length (the graph must be undirected)")
        if A > 1:
            # adj[p] < adj[p+1] unless p+1 starts a list
            up = np.diff(adj.astype(np.int64)) > 0
            starts = off[1:-1].astype(np.int64) - 1
            starts = starts[(starts >= 0) & (starts < A - 1)]
            up[starts] = True
            if not up.all():
                raise ValueError(f"{path}: neighbour lists must be strictly ascending")
        mpath = path + ".json"
        if os.path.exists(mpath):
            with open(mpath) as mf:
                manifest = json.load(mf)
            if manifest.get("sha256") != g.digest():
                raise ValueError(f"{path}: SHA-256 differs from {mpath}")
    return g


def _file_sha256(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for block in iter(lambda: f.read(1 << 22), b""):
            h.update(block)
    return h.hexdigest()


def load_cached(text_path: str, cls=Graph):
    """The reference text format through a binary cache: <text_path>.csr is written on the
    first load, with the source's name, size and SHA-256 in its manifest, and read
    afterwards only while the text file still has that size and digest (hashing the text
    is much cheaper than parsing it; modification times are not trusted, since copies
    can carry old ones)."""
    csr = text_path + ".csr"
    src_bytes = os.path.getsize(text_path)
    src_sha = _file_sha256(text_path)
    mpath = csr + ".json"
    if os.path.exists(csr) and os.path.exists(mpath):
        try:
            with open(mpath) as mf:
                manifest = json.load(mf)
            if manifest.get("source_bytes") == src_bytes and manifest.get("source_sha256") == src_sha:
                return load_csr(csr, cls=cls)
        except ValueError:
            pass  # damaged: rebuild below
    g = load_text(text_path, cls=cls)
    save_csr(g, csr, {"source": os.path.basename(text_path), "source_bytes": src_bytes, "source_sha256": src_sha})
    return g

