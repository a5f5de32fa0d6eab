"""Python mirror of the reference's filter / order / enumerate call surface, backed by the
gfx950 library through its C ABI.

Reference surface (inc/api.h, inc/filter.h, inc/order.h, inc/explore.h):

    CandidateParam can(data, query);            -> can = CandidateParam(data, query)
    Filter_NLF(data, query, can) / Filter_GQL     -> same names, return bool
    Order_RI(query, order) / Order_GQL(...)      -> return the order list
    ReuseParam reuse(query, order)               -> ReuseParam(query, order)
    Config cfg; cfg.useFilter = ...              -> Config(...)
    explore_Intersection_Cache_DAF_Filter(data, query, order, cfg, time, ans, stop, limit)
                                                 -> returns ExploreResult(time, ans, ...)
    explore_Intersection_All<false x5>(...)      -> explore_Intersection_All(...)

Profiler counters (inc/common_type.h:150-226) are mirrored by ``Profiler.getInst()``.
Everything runs on the GPU; a missing library or device raises instead of falling back.
"""
from __future__ import annotations

import atexit
import ctypes as C
import threading
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .graph import Graph, Query

__all__ = ["Context", "default_context", "DeviceGraph", "CandidateParam", "Edges", "Filter_NLF",
           "Filter_GQL", "Filter_CFL", "Filter_CECI", "Filter_DAF", "Order_RI", "Order_GQL", "Order_QSI",
           "Order_CFL", "Order_CECI", "Order_GuP", "query_kcore", "ReuseParam", "Config", "ExploreResult",
           "explore_Intersection_Cache_DAF_Filter", "explore_Intersection_All", "Profiler",
           "filter_batch", "enumerate_query", "enumerate_batch", "plan_export", "Batch", "DAFStartVertex",
           "computeWeightArray", "explore_Intersection_DynamicOrder_DAF", "explore_Intersection_DynamicOrder_VEQ",
           "gup_guards"]


def _check(code, ctx=None):
    if code != L.SGM_OK:
        msg = L.lib().sgm_ctx_last_error(ctx.handle if ctx else None)
        raise L.SgmError(code, msg.decode() if msg else "")


class Context:
    """One device context (stream, scratch buffers) on one GPU."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        code = L.lib().sgm_ctx_create(device, C.byref(h))
        if code != L.SGM_OK:
            raise L.SgmError(code, f"cannot create a context on device {device}")
        self.handle = h
        self.device = device
        self._graphs = {}
        self._children = weakref.WeakSet()  # batches and cost models on this context
        _live_contexts.add(self)

    def adopt(self, obj):
        """Register an object whose library handle needs this context (closed first)."""
        self._children.add(obj)

    def close(self):
        if self.handle:
            for ch in list(self._children):
                ch.close()
            for dg in list(self._graphs.values()):
                dg.close()
            self._graphs.clear()
            L.lib().sgm_ctx_destroy(self.handle)
            self.handle = None

    def filter_ms(self):
        """(nlf, gql, candidate-count) device milliseconds of the last filter call."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(L.lib().sgm_ctx_filter_ms(self.handle, C.byref(a), C.byref(b), C.byref(c)), self)
        return a.value, b.value, c.value

    def stop(self):
        """Cooperative stop of the running enumeration (the reference's stop_flag)."""
        _check(L.lib().sgm_stop(self.handle), self)

    def upload(self, g: Graph) -> "DeviceGraph":
        key = id(g)
        dg = self._graphs.get(key)
        if dg is None or dg.graph is not g:
            dg = DeviceGraph(self, g)
            self._graphs[key] = dg
        elif dg.labels is not g.labels:  # labels attached or replaced since the upload
            dg.set_labels(g.labels)
        return dg

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None
_default_lock = threading.Lock()
_live_contexts = weakref.WeakSet()


@atexit.register
def _close_all_contexts():
    """Release every device object in dependency order while the interpreter is intact
    (finalizers at exit run in no fixed order)."""
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


class DeviceGraph:
    """The data graph resident in HBM."""

    def __init__(self, ctx: Context, g: Graph):
        self.ctx = ctx
        self.graph = g
        h = C.c_void_p()
        off, adj = L.u32(g.off), L.u32(g.adj)
        _check(L.lib().sgm_graph_upload(ctx.handle, L.ptr(off), L.ptr(adj), g.getVertexCnt(), adj.size,
                                        C.byref(h)), ctx)
        self.handle = h
        self.max_degree = int(L.lib().sgm_graph_max_degree(h))
        self.labels = None
        if g.labels is not None:
            self.set_labels(g.labels)

    def set_labels(self, labels):
        """Vertex labels of the resident graph (None: unlabeled); set them before filtering."""
        lab = None if labels is None else L.u32(labels)
        _check(L.lib().sgm_graph_set_labels(self.handle, None if lab is None else L.ptr(lab),
                                            self.graph.getVertexCnt()), self.ctx)
        self.labels = labels

    def close(self):
        if self.handle:
            L.lib().sgm_graph_destroy(self.handle)
            self.handle = None


class Edges:
    """One auxiliary table edge_matrix[a][b] (inc/graph.h:13-49), downloaded lazily."""

    def __init__(self, can: "CandidateParam", a: int, b: int):
        e = C.c_uint64()
        md = C.c_uint32()
        _check(L.lib().sgm_table_info(can._q, a, b, C.byref(e), C.byref(md)), can.ctx)
        self._e_cnt = int(e.value)
        self._max_degree = int(md.value)
        self._v_cnt = int(can.candidates_count[a])
        self._offset = np.zeros(self._v_cnt + 1, dtype=np.uint32)
        self._edge = np.zeros(max(self._e_cnt, 1), dtype=np.uint32)
        self._edge_v = np.zeros(max(self._e_cnt, 1), dtype=np.uint32)
        _check(L.lib().sgm_table_download(can._q, a, b, L.ptr(self._offset), L.ptr(self._edge),
                                          L.ptr(self._edge_v)), can.ctx)
        self._edge = self._edge[:self._e_cnt]
        self._edge_v = self._edge_v[:self._e_cnt]

    def getNeb(self, i: int) -> np.ndarray:
        return self._edge[self._offset[i]:self._offset[i + 1]]

    def getNeb_V(self, i: int) -> np.ndarray:
        return self._edge_v[self._offset[i]:self._offset[i + 1]]


class _EdgeMatrix:
    def __init__(self, can):
        self._can = can
        self._cache = {}

    def __getitem__(self, a):
        can = self._can

        class _Row:
            def __getitem__(_, b):
                if not can.query.hasEdge(a, b):
                    return None
                key = (a, b)
                if key not in self._cache:
                    self._cache[key] = Edges(can, a, b)
                return self._cache[key]
        return _Row()


class CandidateParam:
    """Candidate space of one query on the device (inc/api.h:9-164).

    ``candidates[u]`` / ``candidates_count[u]`` / ``candidates_exsist[u][v]`` /
    ``edge_matrix[a][b]`` are materialised on the host on first access.
    """

    def __init__(self, data: Graph, query: Query, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.data = data
        self.query = query
        self.dg = self.ctx.upload(data)
        h = C.c_void_p()
        qoff, qadj = L.u32(query.off), L.u32(query.adj)
        _check(L.lib().sgm_query_create(self.dg.handle, L.ptr(qoff), L.ptr(qadj), query.getVertexCnt(),
                                        C.byref(h)), self.ctx)
        self._q = h
        if query.labels is not None:
            qlab = L.u32(query.labels)
            _check(L.lib().sgm_query_set_labels(h, L.ptr(qlab), query.getVertexCnt()), self.ctx)
        self._filtered = False
        self._counts = None
        self._cands = None
        self._exist = None
        self.edge_matrix = _EdgeMatrix(self)

    def close(self):
        if self._q:
            L.lib().sgm_query_destroy(self._q)
            self._q = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_tables(self):
        """FilterParamConstruct's aux-edge tables (inc/table.h:7-137) on the device."""
        _check(L.lib().sgm_build_tables(self._q), self.ctx)

    def _after_filter(self):
        # host copies (counts, lists, membership, tables) are fetched on first access
        self._filtered = True
        self._counts = None
        self._cands = None
        self._exist = None
        self.edge_matrix = _EdgeMatrix(self)

    @property
    def candidates_count(self) -> np.ndarray:
        if self._counts is None:
            if not self._filtered:
                raise RuntimeError("filter the candidates first")
            cnt = np.zeros(self.query.getVertexCnt(), dtype=np.uint32)
            _check(L.lib().sgm_cand_counts(self._q, L.ptr(cnt)), self.ctx)
            self._counts = cnt
        return self._counts

    @property
    def candidates(self):
        if self._cands is None:
            out = []
            for u in range(self.query.getVertexCnt()):
                c = int(self.candidates_count[u])
                buf = np.zeros(max(c, 1), dtype=np.uint32)
                _check(L.lib().sgm_cand_download(self._q, u, L.ptr(buf), max(c, 1)), self.ctx)
                out.append(buf[:c])
            self._cands = out
        return self._cands

    @property
    def candidates_exsist(self) -> np.ndarray:
        """n x V boolean matrix, as the reference's candidates_exsist."""
        if self._exist is None:
            V = self.data.getVertexCnt()
            words = np.zeros(max(V, 1), dtype=np.uint32)
            _check(L.lib().sgm_cand_exist_download(self._q, L.ptr(words), V), self.ctx)
            n = self.query.getVertexCnt()
            self._exist = ((words[None, :V] >> np.arange(n, dtype=np.uint32)[:, None]) & 1).astype(bool)
        return self._exist


def _filter(data, query, can, kind, build_tables):
    code = L.lib().sgm_filter(can._q, kind)
    if code == L.SGM_EEMPTY:
        can._after_filter()
        return False
    _check(code, can.ctx)
    can._after_filter()
    if build_tables:
        _check(L.lib().sgm_build_tables(can._q), can.ctx)
    return True


def Filter_NLF(data: Graph, query: Query, can: CandidateParam, build_tables: bool = True) -> bool:
    """inc/filter.h:1049-1055: NLF candidates, then FilterParamConstruct (tables)."""
    return _filter(data, query, can, L.SGM_FILTER_NLF, build_tables)


def Filter_GQL(data: Graph, query: Query, can: CandidateParam, build_tables: bool = True) -> bool:
    """inc/filter.h:1057-1063: GQL candidates, then FilterParamConstruct (tables)."""
    return _filter(data, query, can, L.SGM_FILTER_GQL, build_tables)


def Filter_CFL(data: Graph, query: Query, can: CandidateParam, build_tables: bool = True) -> bool:
    """inc/filter.h:1065-1071 (_CFL_Filter :481-583): BFS-tree generation and refinement on the
    device, then FilterParamConstruct (tables)."""
    return _filter(data, query, can, L.SGM_FILTER_CFL, build_tables)


def Filter_CECI(data: Graph, query: Query, can: CandidateParam, build_tables: bool = True) -> bool:
    """inc/filter.h:1073-1079 (_CECI_Filter :602-920).  Where the reference ends the process on
    an empty candidate set (:700-703), this returns False."""
    return _filter(data, query, can, L.SGM_FILTER_CECI, build_tables)


def Filter_DAF(data: Graph, query: Query, can: CandidateParam, build_tables: bool = True) -> bool:
    """inc/filter.h:1081-1087 (_DAF_Filter :942-1020): NLF, then three refinement passes."""
    return _filter(data, query, can, L.SGM_FILTER_DAF, build_tables)


def filter_batch(cans, kind=L.SGM_FILTER_GQL):
    """Filter many queries of one data graph in one set of launches; returns [bool]."""
    if not cans:
        return []
    ctx = cans[0].ctx
    arr = (C.c_void_p * len(cans))(*[c._q.value for c in cans])
    status = np.zeros(len(cans), dtype=np.int32)
    _check(L.lib().sgm_filter_batch(arr, len(cans), kind, L.ptr(status)), ctx)
    for c in cans:
        c._after_filter()
    return [int(s) == L.SGM_OK for s in status]


def Order_RI(query: Query) -> list:
    """inc/order.h:1138-1281."""
    n = query.getVertexCnt()
    o = np.zeros(n, dtype=np.uint32)
    qoff, qadj = L.u32(query.off), L.u32(query.adj)
    code = L.lib().sgm_order_ri(L.ptr(qoff), L.ptr(qadj), n, L.ptr(o))
    if code != L.SGM_OK:
        raise L.SgmError(code, "Order_RI")
    return [int(x) for x in o]


def Order_GQL(data: Graph, query: Query, can: CandidateParam) -> list:
    """inc/order.h:563-597 (needs filtered candidate counts)."""
    o = np.zeros(query.getVertexCnt(), dtype=np.uint32)
    _check(L.lib().sgm_order_gql(can._q, L.ptr(o)), can.ctx)
    return [int(x) for x in o]


def _order_of(fn, name, query, can):
    o = np.zeros(query.getVertexCnt(), dtype=np.uint32)
    code = fn(can._q, L.ptr(o))
    if code != L.SGM_OK:
        msg = L.lib().sgm_ctx_last_error(can.ctx.handle)
        raise L.SgmError(code, f"{name}: {msg.decode() if msg else ''}")
    return [int(x) for x in o]


def Order_QSI(data: Graph, query: Query, can: CandidateParam) -> list:
    """inc/order.h:610-668 (needs filtered candidates; builds the aux tables if missing)."""
    return _order_of(L.lib().sgm_order_qsi, "Order_QSI", query, can)


def Order_CFL(data: Graph, query: Query, can: CandidateParam) -> list:
    """inc/order.h:781-976 (path estimates over the aux tables run on the device)."""
    return _order_of(L.lib().sgm_order_cfl, "Order_CFL", query, can)


def Order_CECI(data: Graph, query: Query, can: CandidateParam | None = None) -> list:
    """inc/order.h:599-607: the BFS order from CECIStartVertex (NLF counts only: ``can`` may be
    unfiltered, or None for a temporary candidate space on the default context)."""
    if can is None:
        can = CandidateParam(data, query)
        try:
            return _order_of(L.lib().sgm_order_ceci, "Order_CECI", query, can)
        finally:
            can.close()
    return _order_of(L.lib().sgm_order_ceci, "Order_CECI", query, can)


def Order_GuP(data: Graph, query: Query, can: CandidateParam) -> list:
    """inc/order.h:537-561.  Raises SgmError(SGM_EINVAL) for a query whose vertices all have
    degree <= 2, where the reference asserts (:484-492)."""
    return _order_of(L.lib().sgm_order_gup, "Order_GuP", query, can)


def query_kcore(query: Query) -> list:
    """Query::KCore core numbers (inc/graph.h:387-458)."""
    n = query.getVertexCnt()
    core = np.zeros(n, dtype=np.int32)
    qoff, qadj = L.u32(query.off), L.u32(query.adj)
    code = L.lib().sgm_query_kcore(L.ptr(qoff), L.ptr(qadj), n, L.ptr(core))
    if code != L.SGM_OK:
        raise L.SgmError(code, "query_kcore")
    return [int(x) for x in core]


def plan_export(query: Query, order):
    """The intersection-cache plan (inc/code_gen.h:682-847) as built by the library; ops of
    a labelled query are keyed by the follow vertex's label."""
    n = query.getVertexCnt()
    qoff, qadj, o = L.u32(query.off), L.u32(query.adj), L.u32(order)
    qlab = None if query.labels is None else L.u32(query.labels)
    nops = C.c_uint32()
    arrs = [np.zeros(512, dtype=np.uint32) for _ in range(4)]
    it = np.zeros(n, dtype=np.uint32)
    code = L.lib().sgm_plan_export_labeled(L.ptr(qoff), L.ptr(qadj), None if qlab is None else L.ptr(qlab), n,
                                           L.ptr(o), C.byref(nops), *[L.ptr(a) for a in arrs], L.ptr(it))
    if code != L.SGM_OK:
        raise L.SgmError(code, "plan")
    k = nops.value
    ops = []
    for i in range(k):
        m = int(arrs[3][i])
        ops.append({"depth": int(arrs[0][i]), "parent": None if arrs[1][i] == 0xFFFFFFFF else int(arrs[1][i]),
                    "vdepth": int(arrs[2][i]), "edges": [b for b in range(32) if (m >> b) & 1]})
    return {"ops": ops, "iter_set": [None if x == 0xFFFFFFFF else int(x) for x in it]}


class ReuseParam:
    """inc/api.h:167-182: owns the intersection-cache plan for (query, order)."""

    def __init__(self, query: Query, order):
        self.order = list(order)
        self.plan = plan_export(query, order)


@dataclass
class Config:
    """inc/api.h:184-227 (non-owning)."""
    useFilter: bool = False
    filterType: str = "gql"
    orderType: str = "gql"
    useCache: bool = False
    usePrune: bool = False
    can: CandidateParam | None = None
    reuse: ReuseParam | None = None
    useDAFFailingPrune: bool = False

    def Check(self):
        if self.useFilter and self.can is None:
            raise ValueError("config error, use filter but canParam is nullptr")
        if self.useCache and self.reuse is None:
            raise ValueError("config error, use cache but reuseParam is nullptr")


class Profiler:
    """Mirror of the reference's Profiler counters (inc/common_type.h:150-226)."""
    _inst = None

    def __init__(self):
        self.reset()

    @classmethod
    def getInst(cls):
        if cls._inst is None:
            cls._inst = Profiler()
        return cls._inst

    def reset(self):
        self.total_intersection = 0
        self.total_iter_count = [0] * 300
        self.daf_prune_cnt = 0
        self.gup_cd_cnt = 0


@dataclass
class ExploreResult:
    time: float            # ms, enumeration kernel time
    ans: int               # min(#embeddings, output_limit)
    intersections: int = 0
    nodes: list = field(default_factory=list)
    fs_prunes: int = 0
    elems_read: int = 0
    elems_out: int = 0
    elems_scanned: int = 0
    capped: bool = False
    retries: int = 0
    stopped: bool = False  # ended early by a stop request: the count is partial
    busy_ns: int = 0       # wave time spent on the query, summed over waves
    timed_out: bool = False  # passed its time_limit_ms work budget: the count is partial
    probe_words: int = 0     # words its intersections read: keys plus bitmap / binary-search probes


def enumerate_query(can: CandidateParam, order, method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, blocks=0,
                    stack_words=0, flags=0, time_limit_ms=0, roots=None, split_chunks=0) -> ExploreResult:
    """``time_limit_ms``: per-query work budget in device wave time (0 = none); a query that
    passes it comes back with ``timed_out`` set and a partial count.  ``roots`` = (lo, hi):
    search only the depth-0 candidates with lo <= id < hi (disjoint ranges add up).  For the
    dynamic-order methods ``order`` is the start order and ``nodes`` counts the extensions
    per depth of the search, whose vertex at a depth varies (the reference's dynamic searches
    count no iterations, so Profiler.total_iter_count is left alone)."""
    opts = _opts(method, cap, blocks, stack_words, flags, time_limit_ms=time_limit_ms, roots=roots,
                 split_chunks=split_chunks)
    st = L.Stats()
    o = L.u32(order)
    _check(L.lib().sgm_enumerate(can._q, L.ptr(o), C.byref(opts), C.byref(st)), can.ctx)
    res = _result(st, can.query.getVertexCnt())
    prof = Profiler.getInst()
    prof.total_intersection += res.intersections
    if method not in (L.SGM_ENUM_DYN_DAF, L.SGM_ENUM_DYN_VEQ):
        for d, x in enumerate(res.nodes):
            prof.total_iter_count[d] += x
    if method == L.SGM_ENUM_AUX_TABLES_GUPCD:
        prof.gup_cd_cnt += res.fs_prunes
    else:
        prof.daf_prune_cnt += res.fs_prunes
    return res


def enumerate_batch(cans, orders=None, method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, blocks=0, stack_words=0,
                    flags=0, time_limit_ms=0, roots=None, split_chunks=0):
    """Enumerate many filtered queries of one data graph in one persistent launch.
    ``orders`` None (or a None entry): the RI order, computed by the library.  Returns one
    ExploreResult per query (``time`` is the whole batch's kernel time).  ``time_limit_ms``:
    per-query work budget (device wave time); queries past it come back ``timed_out``.
    ``roots`` = (lo, hi): every query searches only its depth-0 candidates in that id range."""
    if not cans:
        return []
    ctx = cans[0].ctx
    opts = _opts(method, cap, blocks, stack_words, flags, time_limit_ms=time_limit_ms, roots=roots,
                 split_chunks=split_chunks)
    nq = len(cans)
    ords = [None if o is None else L.u32(o) for o in orders] if orders is not None else None
    qarr = (C.c_void_p * nq)(*[c._q.value for c in cans])
    oarr = None if ords is None else (C.c_void_p * nq)(*[None if o is None else o.ctypes.data for o in ords])
    stats = (L.Stats * nq)()
    _check(L.lib().sgm_enumerate_batch(qarr, oarr, nq, C.byref(opts), stats), ctx)
    out = []
    prof = Profiler.getInst()
    for can, st in zip(cans, stats):
        r = _result(st, can.query.getVertexCnt())
        prof.total_intersection += r.intersections
        out.append(r)
    return out


def explore_Intersection_Cache_DAF_Filter(data: Graph, query: Query, order, cfg: Config,
                                          output_limit: int | None = None) -> ExploreResult:
    """inc/explore.h:1073-1129 (RI+cache+filter+fs row)."""
    cfg.Check()
    if cfg.can is None:
        raise ValueError("the cache+filter+DAF search needs filtered candidates (cfg.can)")
    return enumerate_query(cfg.can, order, L.SGM_ENUM_CACHE_FILTER_FS, output_limit)


def explore_Intersection_All(data: Graph, query: Query, order, cfg: Config, output_limit: int | None = None,
                             DAFFailingPrune: bool = False, GuPFailingPrune: bool = False,
                             GUPConflictDetection: bool = False, BICEConflictDetection: bool = False,
                             VEQAutomorphismPrune: bool = False) -> ExploreResult:
    """inc/explore.h:2757-2848, the template flags as keywords: plain (the "RI" / "GQL" rows),
    with DAF failing sets, with GuP's conflict detection, or both.  GuP failing sets, BICE
    and VEQ pruning are not on the accelerated path (NotImplementedError)."""
    cfg.Check()
    if cfg.can is None:
        raise ValueError("the candidate-space search needs filtered candidates (cfg.can)")
    if GuPFailingPrune or BICEConflictDetection or VEQAutomorphismPrune:
        raise NotImplementedError("explore_Intersection_All with GuP failing sets, BICE or VEQ pruning is not on "
                                  "the accelerated path")
    method = {(False, False): L.SGM_ENUM_AUX_TABLES, (True, False): L.SGM_ENUM_AUX_TABLES_FS,
              (False, True): L.SGM_ENUM_AUX_TABLES_GUPCD,
              (True, True): L.SGM_ENUM_AUX_TABLES_FS_GUPCD}[(bool(DAFFailingPrune), bool(GUPConflictDetection))]
    return enumerate_query(cfg.can, order, method, output_limit)


def gup_guards(can: CandidateParam, order):
    """GUPConflictDetection::Init's reservation guards for ``order`` (inc/prune.h:828-905),
    computed by the library: (goff[n+1], words) with 4 words per (depth, candidate position)
    in depth order -- the guard's size (0 = none), then its data vertices ascending."""
    n = can.query.getVertexCnt()
    o = L.u32(order)
    goff = np.zeros(n + 1, dtype=np.uint32)
    _check(L.lib().sgm_gup_guards(can._q, L.ptr(o), L.ptr(goff), None, 0), can.ctx)
    words = np.zeros(max(4 * int(goff[n]), 1), dtype=np.uint32)
    _check(L.lib().sgm_gup_guards(can._q, L.ptr(o), None, L.ptr(words), words.size), can.ctx)
    return goff, words[:4 * int(goff[n])]


def DAFStartVertex(data: Graph, query: Query, can: CandidateParam) -> int:
    """inc/filter.h:922-940: the start vertex of the DAF BFS order (NLF counts only)."""
    s = C.c_uint32()
    _check(L.lib().sgm_daf_start_vertex(can._q, C.byref(s)), can.ctx)
    return int(s.value)


def computeWeightArray(data: Graph, query: Query, can: CandidateParam, bfs_order: list) -> list:
    """inc/order.h:978-1031 on the device: ``bfs_order`` (a list, cleared first) receives the BFS
    order from DAFStartVertex; returns weight_array, one uint32 array of candidates_count[u]
    weights per query vertex (the reference fills a caller-allocated uint32_t**)."""
    n = query.getVertexCnt()
    cnt = can.candidates_count
    w = [np.zeros(max(int(c), 1), dtype=np.uint32) for c in cnt]
    o = np.zeros(n, dtype=np.uint32)
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in w])
    _check(L.lib().sgm_daf_weights(can._q, L.ptr(o), ptrs), can.ctx)
    bfs_order.clear()
    bfs_order.extend(int(x) for x in o)
    return [w[u][:int(cnt[u])] for u in range(n)]


def _dyn_check(query, bfs_order, cfg):
    cfg.Check()
    if cfg.can is None:
        raise ValueError("the dynamic-order search needs filtered candidates (cfg.can)")
    if len(bfs_order) != query.getVertexCnt():
        raise L.SgmError(L.SGM_EORDER, "bfs_order length differs from the query size")


def _dynamic(query, bfs_order, match_order, cfg, method, output_limit):
    # the reference writes match_order past its cleared size (inc/explore.h:1877-1879), so
    # the caller's vector comes back empty
    match_order.clear()
    return enumerate_query(cfg.can, bfs_order, method, output_limit)


def explore_Intersection_DynamicOrder_DAF(data: Graph, query: Query, bfs_order, match_order: list, cfg: Config,
                                          weight_array, output_limit: int | None = None) -> ExploreResult:
    """inc/explore.h:1849-1931: the next query vertex per search node by the DAF weights in
    ``weight_array`` (from computeWeightArray over ``bfs_order``)."""
    _dyn_check(query, bfs_order, cfg)
    n = query.getVertexCnt()
    if len(weight_array) != n:
        raise ValueError("weight_array needs one array per query vertex")
    cnt = cfg.can.candidates_count
    ws = []
    for u in range(n):
        a = np.ascontiguousarray(weight_array[u], dtype=np.uint32)
        if a.size < int(cnt[u]):
            raise ValueError(f"weight_array[{u}] holds {a.size} weights for {int(cnt[u])} candidates")
        ws.append(a if a.size else np.zeros(1, dtype=np.uint32))
    o = L.u32(bfs_order)
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in ws])
    _check(L.lib().sgm_set_daf_weights(cfg.can._q, L.ptr(o), ptrs), cfg.can.ctx)
    return _dynamic(query, bfs_order, match_order, cfg, L.SGM_ENUM_DYN_DAF, output_limit)


def explore_Intersection_DynamicOrder_VEQ(data: Graph, query: Query, bfs_order, match_order: list, cfg: Config,
                                          output_limit: int | None = None) -> ExploreResult:
    """inc/explore.h:1993-2062: the same search ranked by candidate counts."""
    _dyn_check(query, bfs_order, cfg)
    return _dynamic(query, bfs_order, match_order, cfg, L.SGM_ENUM_DYN_VEQ, output_limit)


def _opts(method, cap, blocks=0, stack_words=0, flags=0, hub_min_degree=0, time_limit_ms=0, roots=None,
          split_chunks=0):
    o = L.EnumOpts()
    o.cap = (1 << 64) - 1 if cap is None else int(cap)
    o.method = method
    o.blocks = blocks
    o.stack_words = stack_words
    o.flags = flags
    o.hub_min_degree = hub_min_degree
    o.time_limit_ms = int(time_limit_ms)
    o.split_chunks = int(split_chunks)
    if roots is not None:
        lo, hi = int(roots[0]), int(roots[1])
        if not 0 <= lo < hi <= 0xFFFFFFFF:
            raise ValueError(f"root range must satisfy 0 <= lo < hi < 2^32 (got {roots})")
        o.root_begin, o.root_end = lo, hi
    return o


def _result(st, n):
    return ExploreResult(time=float(st.kernel_ms), ans=int(st.count), intersections=int(st.intersections),
                         nodes=[int(x) for x in st.nodes[:n]], fs_prunes=int(st.fs_prunes),
                         elems_read=int(st.elems_read), elems_out=int(st.elems_out),
                         elems_scanned=int(st.elems_scanned), capped=bool(st.capped), retries=int(st.retries),
                         stopped=bool(st.stopped), busy_ns=int(st.busy_ns), timed_out=bool(st.timed_out),
                         probe_words=int(st.probe_words))


class Batch:
    """Batched runner over one resident data graph (sgm_batch_*): up to ``slots`` queries
    per batch, filtered and enumerated by replayed hipGraphs, with per-slot device state
    allocated once.  ``filter`` loads a batch; ``enumerate`` runs it with RI orders (None),
    or per-query orders chosen in between (``order_gql`` or any connected order)."""

    def __init__(self, data: Graph, slots: int = 64, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.data = data
        self.dg = self.ctx.upload(data)
        self.slots = int(slots)
        h = C.c_void_p()
        _check(L.lib().sgm_batch_create(self.dg.handle, self.slots, C.byref(h)), self.ctx)
        self.handle = h
        self.queries = []
        self.ctx.adopt(self)

    def close(self):
        if self.handle:
            L.lib().sgm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def filter(self, queries, kind=L.SGM_FILTER_GQL):
        """Returns (ok[nq] bools, cand_counts[nq, 32], cand_degsum[nq, 32]).  Labelled queries
        (or a labelled data graph) run the labelled NLF / GQL."""
        nq = len(queries)
        self._keep = [(L.u32(q.off), L.u32(q.adj)) for q in queries]
        offs = (C.c_void_p * nq)(*[a.ctypes.data for a, _ in self._keep])
        adjs = (C.c_void_p * nq)(*[b.ctypes.data for _, b in self._keep])
        ns = L.u32([q.getVertexCnt() for q in queries])
        status = np.zeros(nq, dtype=np.int32)
        cnt = np.zeros((nq, 32), dtype=np.uint32)
        dsum = np.zeros((nq, 32), dtype=np.uint64)
        if any(q.labels is not None for q in queries):
            self._keep_lab = [None if q.labels is None else L.u32(q.labels) for q in queries]
            labs = (C.c_void_p * nq)(*[None if a is None else a.ctypes.data for a in self._keep_lab])
            _check(L.lib().sgm_batch_filter_labeled(self.handle, nq, offs, adjs, labs, L.ptr(ns), kind, L.ptr(status),
                                                    L.ptr(cnt), L.ptr(dsum)), self.ctx)
        else:
            _check(L.lib().sgm_batch_filter(self.handle, nq, offs, adjs, L.ptr(ns), kind, L.ptr(status), L.ptr(cnt),
                                            L.ptr(dsum)), self.ctx)
        self.queries = list(queries)
        return [int(x) == L.SGM_OK for x in status], cnt, dsum

    def order_gql(self, i: int) -> list:
        o = np.zeros(self.queries[i].getVertexCnt(), dtype=np.uint32)
        _check(L.lib().sgm_batch_order_gql(self.handle, i, L.ptr(o)), self.ctx)
        return [int(x) for x in o]

    def _orders(self, orders):
        if orders is None:
            return None, None
        keep = [None if o is None else L.u32(o) for o in orders]
        arr = (C.c_void_p * len(keep))(*[None if o is None else o.ctypes.data for o in keep])
        return arr, keep

    def enumerate(self, orders=None, method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, blocks=0, stack_words=0,
                  flags=0, hub_min_degree=0, time_limit_s=None, time_limit_ms=0, roots=None, split_chunks=0):
        """With ``time_limit_s``, a stop request ends the launch after that many seconds;
        queries still running then come back with ``stopped`` set and partial counts.
        ``time_limit_ms``: per-query work budget (device wave time); queries past it come back
        ``timed_out``.  ``split_chunks``: chunks per heavy root's depth-1 loop (0 = 32)."""
        arr, _keep = self._orders(orders)
        opts = _opts(method, cap, blocks, stack_words, flags, hub_min_degree, time_limit_ms, roots, split_chunks)
        stats = (L.Stats * len(self.queries))()
        timer = None
        if time_limit_s:
            import threading
            timer = threading.Timer(time_limit_s, self.ctx.stop)
            timer.start()
        try:
            code = L.lib().sgm_batch_enumerate(self.handle, arr, C.byref(opts), stats)
        finally:
            if timer:
                timer.cancel()
        if code != L.SGM_ESTOPPED or not time_limit_s:
            _check(code, self.ctx)
        return [_result(st, q.getVertexCnt()) for st, q in zip(stats, self.queries)]

    def enumerate_start(self, orders=None, method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None, blocks=0, stack_words=0,
                        flags=0, hub_min_degree=0, time_limit_ms=0, roots=None, split_chunks=0):
        """Launch the loaded batch's enumeration and return at once
        (sgm_batch_enumerate_launch); ``enumerate_wait`` returns the results.  In between,
        another Batch of the same context can filter and launch the next batch, which then
        fills the GPU as this one runs out of work."""
        arr, _keep = self._orders(orders)
        opts = _opts(method, cap, blocks, stack_words, flags, hub_min_degree, time_limit_ms, roots, split_chunks)
        _check(L.lib().sgm_batch_enumerate_launch(self.handle, arr, C.byref(opts)), self.ctx)

    def enumerate_wait(self):
        stats = (L.Stats * len(self.queries))()
        _check(L.lib().sgm_batch_enumerate_wait(self.handle, stats), self.ctx)
        return [_result(st, q.getVertexCnt()) for st, q in zip(stats, self.queries)]

    def rank_orders(self, model, n_samples=32, seed=0, margin=0.5):
        """Learned orders of the loaded batch (sgm_batch_rank_orders): returns (orders,
        predicted log2 nodes of the chosen order, of the RI order, learned-order flags)."""
        nq = len(self.queries)
        o = np.zeros((nq, 32), dtype=np.uint32)
        pb = np.zeros(nq, dtype=np.float32)
        pr = np.zeros(nq, dtype=np.float32)
        fl = np.zeros(nq, dtype=np.int32)
        _check(L.lib().sgm_batch_rank_orders(self.handle, model.handle, n_samples, int(seed) & (2**64 - 1),
                                             C.c_float(margin), L.ptr(o), L.ptr(pb), L.ptr(pr), L.ptr(fl)),
               self.ctx)
        orders = [[int(x) for x in o[i, :q.getVertexCnt()]] for i, q in enumerate(self.queries)]
        return orders, pb, pr, fl.astype(bool)

    def autotune(self, orders=None, method=L.SGM_ENUM_CACHE_FILTER_FS, cap=None):
        """Returns ({blocks, flags, hub_min_degree, split_chunks, stack_words}, best_ms) of the
        fastest setting (sgm_batch_autotune)."""
        arr, _keep = self._orders(orders)
        base = _opts(method, cap)
        best = L.EnumOpts()
        ms = C.c_double()
        _check(L.lib().sgm_batch_autotune(self.handle, arr, C.byref(base), C.byref(best), C.byref(ms)), self.ctx)
        return {"blocks": int(best.blocks), "flags": int(best.flags), "hub_min_degree": int(best.hub_min_degree),
                "split_chunks": int(best.split_chunks), "stack_words": int(best.stack_words)}, float(ms.value)

    def last_ms(self):
        f, e = C.c_double(), C.c_double()
        _check(L.lib().sgm_batch_last_ms(self.handle, C.byref(f), C.byref(e)), self.ctx)
        return float(f.value), float(e.value)
