"""Learned matching orders: the cost model (sgm_costmodel_*) and the order ranking built on it.

The model predicts log2(1 + search-tree nodes) of a (query, order) pair from the query graph,
the order and the candidate statistics of the filtered query; its kernels run on the GPU
(BF16 MFMA, FP32 accumulation) for training and inference.  Ranking changes only the order
of a query, never its count: every order of a connected query enumerates the same
embeddings.  ``choose_orders`` falls back to the heuristic order unless the model predicts a
clear gain.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .api import Context, _check, default_context
from .graph import Query

__all__ = ["CostModel", "sample_order", "candidate_orders", "choose_orders", "node_label", "collect_training",
           "train_model", "LABELS", "LABEL_METHOD"]


def sample_order(q: Query, seed: int) -> list:
    """A seeded random connected order (sgm_order_sample)."""
    o = np.zeros(q.getVertexCnt(), dtype=np.uint32)
    qo, qa = L.u32(q.off), L.u32(q.adj)
    _check(L.lib().sgm_order_sample(L.ptr(qo), L.ptr(qa), q.getVertexCnt(), int(seed) & (2**64 - 1), L.ptr(o)))
    return [int(x) for x in o]


def node_label(nodes) -> float:
    """Training target of an enumeration: log2(1 + internal search-tree nodes), i.e. the
    nodes of every depth but the last.  Last-depth nodes are the embeddings themselves (at
    most the cap, and the same for every order), so they carry no ranking signal."""
    return math.log2(1.0 + float(sum(nodes[:-1])))


def busy_label(result) -> float:
    """Training target measuring what an order costs the enumerator: log2(1 + microseconds
    of wave time spent on the query), summed over the waves (``busy_ns``).  Node counts
    weigh every search-tree node alike, but a last-level step over a list cached at an
    earlier depth is evaluated once per chunk of extensions while an intersection is
    evaluated per extension, so two orders with the same node count can differ several-fold
    in time."""
    return math.log2(1.0 + result.busy_ns / 1000.0)


LABELS = {"busy": busy_label, "nodes": lambda r: node_label(r.nodes), "plain-nodes": lambda r: node_label(r.nodes)}
# the search each label's training runs are logged with: the cache + filter + failing-set
# search the orders are chosen for, or the same search without failing sets ("plain": the
# reference's node counts of explore_Intersection_Recursive, inc/explore.h:555-557,646-648)
LABEL_METHOD = {"busy": L.SGM_ENUM_CACHE_FILTER_FS, "nodes": L.SGM_ENUM_CACHE_FILTER_FS,
                "plain-nodes": L.SGM_ENUM_CACHE_FILTER}


class _Samples:
    """ctypes array of sgm_cm_sample with the numpy buffers it points into."""

    def __init__(self, items):
        self.keep = []
        self.arr = (L.CmSample * len(items))()
        for k, (q, order, counts, degsum) in enumerate(items):
            n = q.getVertexCnt()
            bufs = (L.u32(q.off), L.u32(q.adj), L.u32(order), L.u32(np.asarray(counts)[:n]),
                    np.ascontiguousarray(np.asarray(degsum)[:n], dtype=np.uint64))
            self.keep.append(bufs)
            s = self.arr[k]
            s.n = n
            s.qoff, s.qadj, s.order, s.cand_count, s.cand_degsum = [b.ctypes.data for b in bufs]


class CostModel:
    """3-layer message-passing cost model, hidden 64 (see include/sgm_c.h)."""

    def __init__(self, seed: int = 0, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        _check(L.lib().sgm_costmodel_create(self.ctx.handle, int(seed), C.byref(h)), self.ctx)
        self.handle = h
        self.ctx.adopt(self)
        self.n_params = int(L.lib().sgm_costmodel_param_count())

    def close(self):
        if self.handle:
            L.lib().sgm_costmodel_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        _check(L.lib().sgm_costmodel_params(self.handle, L.ptr(p), None), self.ctx)
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _check(L.lib().sgm_costmodel_params(self.handle, None, L.ptr(p)), self.ctx)

    @staticmethod
    def features(item):
        """(feats[32, 32], back[32], fwd[32]) of one (query, order, counts, degsum) item."""
        s = _Samples([item])
        f = np.zeros((32, 32), dtype=np.float32)
        b = np.zeros(32, dtype=np.uint32)
        w = np.zeros(32, dtype=np.uint32)
        _check(L.lib().sgm_costmodel_features(C.byref(s.arr[0]), L.ptr(f), L.ptr(b), L.ptr(w)))
        return f, b, w

    def predict(self, items) -> np.ndarray:
        out = np.zeros(len(items), dtype=np.float32)
        if items:
            s = _Samples(items)
            _check(L.lib().sgm_costmodel_predict(self.handle, s.arr, len(items), L.ptr(out)), self.ctx)
        return out

    def grad(self, items, targets):
        s = _Samples(items)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(L.lib().sgm_costmodel_grad(self.handle, s.arr, L.ptr(t), len(items), L.ptr(g), L.ptr(loss)), self.ctx)
        return g, float(loss[0])

    def train(self, items, targets, epochs=30, minibatch=256, lr=2e-3, seed=0) -> np.ndarray:
        s = _Samples(items)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(L.lib().sgm_costmodel_train(self.handle, s.arr, L.ptr(t), len(items), epochs, minibatch,
                                           C.c_float(lr), int(seed), L.ptr(losses)), self.ctx)
        return losses

    def timing(self) -> dict:
        v = [C.c_double() for _ in range(4)]
        _check(L.lib().sgm_costmodel_timing(self.handle, *[C.byref(x) for x in v]), self.ctx)
        return {"train_ms": v[0].value, "train_flops": v[1].value, "infer_ms": v[2].value,
                "infer_flops": v[3].value}


def candidate_orders(q: Query, ri, gql, n_samples: int, seed: int) -> list:
    """RI, GQL and n_samples seeded sampled orders (duplicates removed, first kept)."""
    out, seen = [], set()
    for o in [ri, gql] + [sample_order(q, (seed << 20) + k) for k in range(n_samples)]:
        if o is None:
            continue
        t = tuple(o)
        if t not in seen:
            seen.add(t)
            out.append(list(o))
    return out


def choose_orders(model: CostModel, queries, counts, degsums, cands, fallback, margin: float = 0.5):
    """Per query, the candidate order with the lowest predicted log2 node count, unless it
    beats the fallback order's prediction by less than ``margin`` (then the fallback).
    Returns (orders, predicted, predicted_fallback)."""
    items, owner = [], []
    for i, q in enumerate(queries):
        for o in cands[i]:
            items.append((q, o, counts[i], degsums[i]))
            owner.append(i)
    pred = model.predict(items)
    best = [None] * len(queries)
    best_p = [math.inf] * len(queries)
    fb_p = [math.inf] * len(queries)
    for (q, o, _, _), i, p in zip(items, owner, pred):
        if list(o) == list(fallback[i]):
            fb_p[i] = float(p)
        if p < best_p[i]:
            best_p[i], best[i] = float(p), list(o)
    orders = []
    for i in range(len(queries)):
        use = best[i] if best[i] is not None and best_p[i] < fb_p[i] - margin else list(fallback[i])
        orders.append(use)
    return orders, best_p, fb_p


def collect_training(runner, queries, n_samples=8, cap=10_000_000, seed=0, time_limit_s=5.0, log=None,
                     label="busy"):
    """Training triples logged by the enumerator: for every query, its RI order, GQL order
    and ``n_samples`` sampled orders are enumerated through the batched runner and labelled
    with ``LABELS[label]``: log2(1 + internal search-tree nodes) of the cache + filter +
    failing-set search the orders are chosen for ("nodes") or of the same search without
    failing sets ("plain-nodes"), or log2(1 + wave microseconds) of the former ("busy").
    Launches cut off by the time limit leave their unfinished queries out.  Returns (items,
    labels, info)."""
    label_of = LABELS[label]
    method = LABEL_METHOD[label]
    from .api import Order_RI
    items, labels = [], []
    dropped = enumerated = 0
    B = runner.slots
    for lo in range(0, len(queries), B):
        part = queries[lo:lo + B]
        ok, cnt, dsum = runner.filter(part)
        sets = []
        for i, q in enumerate(part):
            cands = candidate_orders(q, Order_RI(q), runner.order_gql(i), n_samples, seed * 1000003 + lo + i)
            while len(cands) < 2 + n_samples:  # tiny queries have few distinct orders
                cands.append(cands[len(cands) % max(1, len(cands))])
            sets.append(cands)
        for k in range(2 + n_samples):
            orders = [sets[i][k] for i in range(len(part))]
            res = runner.enumerate(orders=orders, method=method, cap=cap, time_limit_s=time_limit_s)
            for i, r in enumerate(res):
                if not ok[i]:
                    continue
                enumerated += 1
                if r.stopped:
                    dropped += 1
                    continue
                n = part[i].getVertexCnt()
                items.append((part[i], orders[i], cnt[i, :n].copy(), dsum[i, :n].copy()))
                labels.append(label_of(r))
        if log:
            log(f"collected {len(items)} samples from {lo + len(part)} queries ({dropped} cut off)")
    return items, np.asarray(labels, dtype=np.float32), {"enumerated": enumerated, "cut_off": dropped}


def train_model(items, labels, seed=0, epochs=30, minibatch=256, lr=2e-3, ctx=None):
    """A cost model trained on the triples; returns (model, per-epoch losses)."""
    m = CostModel(seed=seed, ctx=ctx)
    losses = m.train(items, labels, epochs=epochs, minibatch=minibatch, lr=lr, seed=seed + 1)
    return m, losses
