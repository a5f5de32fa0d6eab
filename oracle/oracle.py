"""ctypes front-end of the CPU oracle (liboracle.so, built from sgm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INVALID = 0xFFFFFFFF
MAXQ = 32


def build(force: bool = False):
    """Compile liboracle.so / libsortprobe.so in place with the committed Makefile."""
    args = ["make", "-C", HERE]
    if force:
        args += ["-B"]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


_lib = None


class Stats(C.Structure):
    _fields_ = [("count", C.c_uint64), ("intersections", C.c_uint64),
                ("nodes", C.c_uint64 * MAXQ), ("fs_prunes", C.c_uint64),
                ("enum_ms", C.c_double), ("timed_out", C.c_int32), ("capped", C.c_int32)]

    def as_dict(self, n):
        return {"count": int(self.count), "intersections": int(self.intersections),
                "nodes": [int(x) for x in self.nodes[:n]], "fs_prunes": int(self.fs_prunes),
                "enum_ms": float(self.enum_ms), "timed_out": bool(self.timed_out),
                "capped": bool(self.capped)}


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        if _lib.orc_abi_version() != 4:  # a stale build from before the current arguments
            build(force=True)
            raise RuntimeError("liboracle.so was stale and has been rebuilt; restart the process")
        assert _lib.orc_stats_size() == C.sizeof(Stats)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _lab(g):
    """Vertex labels of a graph as a pointer, or None (every label 0)."""
    lab = getattr(g, "labels", None)
    return None if lab is None else _p(_u32(lab))


def nlf(g, q):
    V, n = g.getVertexCnt(), q.getVertexCnt()
    cand = np.zeros((n, V), dtype=np.uint32)
    cnt = np.zeros(n, dtype=np.uint32)
    ok = lib().orc_nlf(C.c_uint32(V), _p(g.off), _p(g.adj), _lab(g), C.c_uint32(n), _p(q.off), _p(q.adj),
                       _lab(q), _p(cand), _p(cnt))
    return bool(ok), [cand[u, :cnt[u]].copy() for u in range(n)]


def gql(g, q, return_stats=False):
    V, n = g.getVertexCnt(), q.getVertexCnt()
    cand = np.zeros((n, V), dtype=np.uint32)
    cnt = np.zeros(n, dtype=np.uint32)
    stats = np.zeros(4, dtype=np.uint64)
    ok = lib().orc_gql(C.c_uint32(V), _p(g.off), _p(g.adj), _lab(g), C.c_uint32(n), _p(q.off), _p(q.adj),
                       _lab(q), _p(cand), _p(cnt), _p(stats))
    res = (bool(ok), [cand[u, :cnt[u]].copy() for u in range(n)])
    if return_stats:
        return res + ({"checks": int(stats[0]), "removed_r0": int(stats[1]),
                       "removed_r1": int(stats[2])},)
    return res


def table(g, cand_a, cand_b):
    """Aux table for one directed query edge: (offsets, idx, vertex ids, max_degree)."""
    V = g.getVertexCnt()
    ca, cb = _u32(cand_a), _u32(cand_b)
    scratch = np.zeros(max(V, 1), dtype=np.int32)
    L = lib()
    L.orc_table.restype = C.c_uint64
    e = L.orc_table(C.c_uint32(V), _p(g.off), _p(g.adj), _p(ca), C.c_uint32(ca.size), _p(cb),
                    C.c_uint32(cb.size), None, None, None, None, _p(scratch))
    off = np.zeros(ca.size + 1, dtype=np.uint32)
    idx = np.zeros(max(e, 1), dtype=np.uint32)
    vv = np.zeros(max(e, 1), dtype=np.uint32)
    md = C.c_uint32(0)
    L.orc_table(C.c_uint32(V), _p(g.off), _p(g.adj), _p(ca), C.c_uint32(ca.size), _p(cb),
                C.c_uint32(cb.size), _p(off), _p(idx), _p(vv), C.byref(md), _p(scratch))
    return off, idx[:e], vv[:e], int(md.value)


def order_ri(q):
    n = q.getVertexCnt()
    o = np.zeros(n, dtype=np.uint32)
    lib().orc_order_ri(C.c_uint32(n), _p(q.off), _p(q.adj), _p(o))
    return [int(x) for x in o]


def order_gql(g, q, counts):
    n = q.getVertexCnt()
    o = np.zeros(n, dtype=np.uint32)
    cnt = _u32(counts)
    lib().orc_order_gql(C.c_uint32(g.getVertexCnt()), C.c_uint32(n), _p(q.off), _p(q.adj),
                        _p(cnt), _p(o))
    return [int(x) for x in o]


def plan(q, order):
    """PlanIR as a dict: ops (depth, label, parent, vdepth, edge depth list) and iter_set."""
    n = q.getVertexCnt()
    o = _u32(order)
    nops = C.c_uint32(0)
    arrs = [np.zeros(512, dtype=np.uint32) for _ in range(4)]
    it = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_plan_export(C.c_uint32(n), _p(q.off), _p(q.adj), _lab(q), _p(o), C.byref(nops),
                               *[_p(a) for a in arrs], _p(it))
    if rc:
        raise ValueError("order is not connected (empty iteration set)")
    k = nops.value
    depth, parent, vdepth, emask = (a[:k] for a in arrs)
    ops = []
    for i in range(k):
        m = int(emask[i])
        ops.append({"depth": int(depth[i]),
                    "parent": None if parent[i] == INVALID else int(parent[i]),
                    "vdepth": int(vdepth[i]),
                    "edges": [b for b in range(32) if (m >> b) & 1]})
    return {"ops": ops, "iter_set": [None if x == INVALID else int(x) for x in it]}


def sort_ops_perm(depth, size, elems, use_std=False):
    """Permutation produced by the oracle's restated sort (or std::sort itself)."""
    n = len(depth)
    d, s = _u32(depth), _u32(size)
    e = _u32(elems).reshape(n, 32)
    perm = np.zeros(max(n, 1), dtype=np.uint32)
    if use_std:
        path = os.path.join(HERE, "libsortprobe.so")
        if not os.path.exists(path):
            build()
        C.CDLL(path).probe_std_sort(C.c_long(n), _p(d), _p(s), _p(e), _p(perm))
    else:
        lib().orc_sort_ops_test(C.c_long(n), _p(d), _p(s), _p(e), _p(perm))
    return [int(x) for x in perm[:n]]


def _cand_matrix(cands, V):
    n = len(cands)
    stride = max([len(c) for c in cands] + [1])
    m = np.zeros((n, stride), dtype=np.uint32)
    cnt = np.zeros(n, dtype=np.uint32)
    for u, c in enumerate(cands):
        m[u, :len(c)] = c
        cnt[u] = len(c)
    return m, cnt, stride


def enum_cache(g, q, cands, order, cap=None, fs=True, time_limit_ms=0.0, roots=None):
    """explore_Intersection_Cache_DAF_Filter restated; returns the stats dict.  ``roots``
    = (lo, hi): search only the depth-0 candidates with lo <= id < hi."""
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    o = _u32(order)
    st = Stats()
    rc = lib().orc_enum_cache(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off),
                              _p(q.adj), _lab(q), _p(m), _p(cnt), C.c_uint32(stride), _p(o),
                              C.c_uint64(cap if cap is not None else 2**64 - 1),
                              C.c_int(1 if fs else 0), C.c_double(time_limit_ms), *_roots(roots), C.byref(st))
    if rc:
        raise ValueError(f"orc_enum_cache failed ({rc})")
    return st.as_dict(n)


def cand_digest(cands) -> str:
    """SHA-256 over candidate sets in query-vertex order: per set its size, then its ids
    ascending, all as little-endian u32 (the form parity checks compare)."""
    import hashlib
    h = hashlib.sha256()
    for c in cands:
        a = np.ascontiguousarray(c, dtype="<u4")
        h.update(np.uint32(a.size).astype("<u4").tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def isect_check(a, b):
    """(same, result): the oracle's intersection of sorted u32 arrays a and b, and whether it
    equals the scalar merge's."""
    a, b = _u32(a), _u32(b)
    o1 = np.zeros(max(1, min(a.size, b.size)), dtype=np.uint32)
    o2 = np.zeros_like(o1)
    n = C.c_uint32(0)
    same = lib().orc_isect_test(_p(a), C.c_uint32(a.size), _p(b), C.c_uint32(b.size), _p(o1), _p(o2), C.byref(n))
    return bool(same), o1[:n.value].copy()


def isect_is_avx2() -> bool:
    return bool(lib().orc_isect_avx2())


def _roots(roots):
    lo, hi = (0, 0) if roots is None else roots
    return C.c_uint32(int(lo)), C.c_uint32(int(hi))


def enum_all(g, q, cands, order, cap=None, time_limit_ms=0.0, fs=False, roots=None, guards=None):
    """explore_Intersection_All<fs,false,cd,false,false> restated; returns the stats dict.
    ``roots`` as in enum_cache.  ``guards``: (goff, guard words) from gup_guards for GuP's
    conflict detection (cd); its prunes are counted in fs_prunes."""
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    o = _u32(order)
    st = Stats()
    goff = gw = None
    if guards is not None:
        goff, gw = _u32(guards[0]), _u32(guards[1])
    rc = lib().orc_enum_all_cd(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off),
                               _p(q.adj), _p(m), _p(cnt), C.c_uint32(stride), _p(o),
                               C.c_uint64(cap if cap is not None else 2**64 - 1), C.c_int(1 if fs else 0),
                               C.c_double(time_limit_ms), *_roots(roots),
                               _p(goff) if goff is not None else None, _p(gw) if gw is not None else None,
                               C.byref(st))
    if rc:
        raise ValueError(f"orc_enum_all failed ({rc})")
    return st.as_dict(n)


def gup_guards(g, q, cands, order, r=3):
    """GUPConflictDetection::Init's reservation guards for ``order``: (goff[n+1], words) with 4
    words per (depth, candidate position) in depth order: size (0 = none), then the guard's
    data vertices ascending."""
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    o = _u32(order)
    goff = np.zeros(n + 1, dtype=np.uint32)
    total = int(sum(int(cnt[u]) for u in order))
    words = np.zeros(4 * max(total, 1), dtype=np.uint32)
    rc = lib().orc_gup_guards(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off), _p(q.adj), _p(m),
                              _p(cnt), C.c_uint32(stride), _p(o), C.c_uint32(r), _p(goff), _p(words))
    if rc:
        raise ValueError(f"orc_gup_guards failed ({rc})")
    return goff, words[:4 * total]


# ---- BFS-tree filters and orders (sgm_oracle_tree.c) -----------------------------------
FILTER_KIND = {"CFL": 3, "CECI": 4, "DAF": 5}   # the C ABI's SGM_FILTER_* numbers
ORDER_KIND = {"QSI": 1, "CFL": 2, "CECI": 3, "GuP": 4}


def filter_tree(kind, g, q):
    """_CFL_Filter / _CECI_Filter / _DAF_Filter + FilterParamConstruct's sort: (ok, sets).
    ``kind`` is "CFL", "CECI" or "DAF"."""
    V, n = g.getVertexCnt(), q.getVertexCnt()
    cand = np.zeros((n, max(V, 1)), dtype=np.uint32)
    cnt = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_filter_tree(C.c_int(FILTER_KIND[kind]), C.c_uint32(V), _p(g.off), _p(g.adj), _lab(g),
                               C.c_uint32(n), _p(q.off), _p(q.adj), _lab(q), _p(cand), _p(cnt))
    if rc < 0:
        raise ValueError(f"orc_filter_tree({kind}) failed ({rc})")
    return bool(rc), [cand[u, :cnt[u]].copy() for u in range(n)]


def start_vertex(kind, g, q):
    """CFLStartVertex / CECIStartVertex / DAFStartVertex."""
    L = lib()
    L.orc_start_vertex.restype = C.c_uint32
    return int(L.orc_start_vertex(C.c_int(FILTER_KIND[kind]), C.c_uint32(g.getVertexCnt()), _p(g.off), _p(g.adj),
                                  _lab(g), C.c_uint32(q.getVertexCnt()), _p(q.off), _p(q.adj), _lab(q)))


def kcore(q):
    """Query::KCore core numbers."""
    n = q.getVertexCnt()
    core = np.zeros(n, dtype=np.int32)
    lib().orc_kcore(C.c_uint32(n), _p(q.off), _p(q.adj), _p(core))
    return [int(x) for x in core]


def bfs(q, start):
    """BFS tree: (order, parent, level)."""
    n = q.getVertexCnt()
    o, par, lev = (np.zeros(n, dtype=np.uint32) for _ in range(3))
    L = lib()
    L.orc_bfs.restype = C.c_uint32
    k = L.orc_bfs(C.c_uint32(n), _p(q.off), _p(q.adj), C.c_uint32(start), _p(o), _p(par), _p(lev))
    return [int(x) for x in o[:k]], [int(x) for x in par], [int(x) for x in lev]


def order_tree(kind, g, q, cands):
    """Order_QSI / Order_CFL / Order_CECI / Order_GuP over the given (filtered) candidate
    sets; None where the reference asserts (GuP on a query of maximum degree <= 2)."""
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    o = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_order_tree(C.c_int(ORDER_KIND[kind]), C.c_uint32(V), _p(g.off), _p(g.adj), _lab(g),
                              C.c_uint32(n), _p(q.off), _p(q.adj), _lab(q), _p(m), _p(cnt), C.c_uint32(stride),
                              _p(o))
    return None if rc else [int(x) for x in o]


def heap_pop_order(keys, keep, use_std=False):
    """Vertices in the order CFLStartVertex's max-heap visits them after trimming it to
    ``keep`` entries: the oracle's restatement, or std::priority_queue itself."""
    k = np.ascontiguousarray(keys, dtype=np.float64)
    out = np.zeros(max(k.size, 1), dtype=np.uint32)
    if use_std:
        path = os.path.join(HERE, "libsortprobe.so")
        if not os.path.exists(path):
            build()
        C.CDLL(path).probe_std_heap(C.c_uint(k.size), _p(k), C.c_uint(keep), _p(out))
    else:
        lib().orc_heap_test(C.c_uint32(k.size), _p(k), C.c_uint32(keep), _p(out))
    return [int(x) for x in out[:min(keep, k.size)]]


# ---- dynamic-order searches (sgm_oracle_dyn.c) ------------------------------------------
class DynStats(C.Structure):
    _fields_ = [("count", C.c_uint64), ("intersections", C.c_uint64), ("nodes", C.c_uint64 * MAXQ),
                ("enum_ms", C.c_double), ("timed_out", C.c_int32), ("capped", C.c_int32)]


def _weights_ptrs(weights):
    if weights is None:
        return None, None
    arrs = [np.ascontiguousarray(w, dtype=np.uint32) for w in weights]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return ptrs, arrs


def dyn_bfs(g, q):
    """The BFS order the dynamic searches start from: BFS from DAFStartVertex
    (computeWeightArray inc/order.h:980-994; src/order_ablations.cpp:158-160)."""
    return bfs(q, start_vertex("DAF", g, q))


def daf_weights(g, q, cands, bfs_order=None, parent=None):
    """computeWeightArray's per-candidate weights, one uint32 array per query vertex."""
    if bfs_order is None:
        bfs_order, parent, _ = dyn_bfs(g, q)
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    out = [np.zeros(max(int(c), 1), dtype=np.uint32) for c in cnt]
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in out])
    lib().orc_daf_weights(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off), _p(q.adj), _p(m), _p(cnt),
                          C.c_uint32(stride), _p(_u32(bfs_order)), _p(_u32(parent)), ptrs)
    return [out[u][:int(cnt[u])] for u in range(n)]


def enum_dynamic(g, q, cands, kind="DAF", cap=None, time_limit_ms=0.0, bfs_order=None, weights=None):
    """explore_Intersection_DynamicOrder_DAF ("DAF", ranked by computeWeightArray weights) or
    _VEQ ("VEQ", ranked by candidate counts); returns count, intersections, per-depth
    extensions (the last not a reference counter), timed_out, capped."""
    assert lib().orc_dyn_stats_size() == C.sizeof(DynStats)
    if bfs_order is None:
        bfs_order, parent, _ = dyn_bfs(g, q)
        if kind == "DAF" and weights is None:
            weights = daf_weights(g, q, cands, bfs_order, parent)
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    ptrs, keep = _weights_ptrs(weights if kind == "DAF" else None)
    st = DynStats()
    rc = lib().orc_enum_dynamic(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off), _p(q.adj), _p(m),
                                _p(cnt), C.c_uint32(stride), _p(_u32(bfs_order)), ptrs,
                                C.c_uint64(cap if cap is not None else 2**64 - 1), C.c_double(time_limit_ms),
                                C.byref(st))
    if rc:
        raise ValueError(f"orc_enum_dynamic failed ({rc})")
    return {"count": int(st.count), "intersections": int(st.intersections), "nodes": [int(x) for x in st.nodes[:n]],
            "timed_out": bool(st.timed_out), "capped": bool(st.capped), "enum_ms": float(st.enum_ms)}


def probe_gup_guards(g, q, cands, order, r=3):
    """The guards built with std::partition / std::unordered_set themselves (libsortprobe.so)."""
    path = os.path.join(HERE, "libsortprobe.so")
    if not os.path.exists(path):
        build()
    P = C.CDLL(path)
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    goff = np.zeros(n + 1, dtype=np.uint32)
    total = int(sum(int(cnt[u]) for u in order))
    words = np.zeros(4 * max(total, 1), dtype=np.uint32)
    P.probe_gup_guards(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off), _p(q.adj), _p(m), _p(cnt),
                       C.c_uint32(stride), _p(_u32(order)), C.c_uint32(r), _p(goff), _p(words))
    return goff, words[:4 * total]


def probe_dynamic(g, q, cands, bfs_order, weights, cap=None):
    """The same search on std::priority_queue itself (libsortprobe.so): (count, intersections,
    per-depth extensions)."""
    path = os.path.join(HERE, "libsortprobe.so")
    if not os.path.exists(path):
        build()
    P = C.CDLL(path)
    V, n = g.getVertexCnt(), q.getVertexCnt()
    m, cnt, stride = _cand_matrix(cands, V)
    ptrs, keep = _weights_ptrs(weights)
    c_, i_ = C.c_uint64(), C.c_uint64()
    nodes = np.zeros(MAXQ, dtype=np.uint64)
    P.probe_dyn_enum(C.c_uint32(V), _p(g.off), _p(g.adj), C.c_uint32(n), _p(q.off), _p(q.adj), _p(m), _p(cnt),
                     C.c_uint32(stride), _p(_u32(bfs_order)), ptrs, C.c_uint64(cap if cap is not None else 2**64 - 1),
                     C.byref(c_), C.byref(i_), _p(nodes))
    return int(c_.value), int(i_.value), [int(x) for x in nodes[:n]]
