#!/usr/bin/env python3
"""Headline benchmark: exact embeddings/s and queries/s on one MI355X for the
filter -> order -> enumerate path of unlabeled subgraph matching.

Workloads (one step = one pass of the path over the whole query set):

  chunglu100k (default)
    data graph  Chung-Lu, 100,000 vertices, degree exponent 2.3, average degree 16, seed 2
                (unlabeled, simple, sorted u32 CSR; generated once, resident in HBM)
    queries     200 dense 16-vertex random-walk queries (average degree > 3), seed 22
    path        NLF + GQL filtering of all 200 queries in one set of launches, RI order
                (host), cache + filter + failing-set enumeration of all 200 queries in one
                persistent launch (per-query caps and statistics)
  chunglu10k
    data graph  Chung-Lu, 10,000 vertices, degree exponent 2.5, average degree 7.4, seed 1
    query       one 8-vertex sparse random-walk query (average degree <= 3), seed 11
    path        GQL filter, RI order, cache + filter + failing-set enumeration of the one
                query (its embeddings pass 1.5e13, so the step is capped like the CPU side;
                an uncapped pass under a work budget reports the count it reached)
  rmat20-learned
    the rmat20 graph and query mix with learned orders.  Two cost models are trained (untimed)
    on 5,000 training queries (seed 44) x 10 orders logged by the enumerator: the ranking model
    on the wave time each order costs the search (--learn-label busy) and the node-count model
    on the search-tree node counts of the plain search (failing sets off).  Each step filters,
    ranks RI + GQL + 32 sampled orders per query on the GPU with the ranking model, keeps RI
    unless it predicts a clear gain, and enumerates; the same pass with RI orders and one with
    the node-count model's orders are timed after it (learned-vs-RI speedups, and predicted vs
    measured node counts of the node-count model's orders)
  rmat22-serve
    data graph  R-MAT scale 22, average degree 32 (2^26 undirected edges), seed 5
    queries     1,200 random-walk queries, sizes uniform in 8..32, mixed sparse / dense,
                seed 55, cap 10^6, arriving as a Poisson stream at --serve-load (0.7) of the
                capacity measured by a back-to-back pass (seed 77)
    path        serving loop: dynamic batches of every arrived query, up to the slot count
                admitted by a device-memory budget (--serve-mem-gb), learned orders with RI
                fallback (a cost model trained untimed on 600 queries, seed 66); one step = one
                serving run over the stream; the same stream served with RI orders is
                reported beside it (queries/s, p50 / p99 latency)
  rmat20
    data graph  R-MAT scale 20 (a,b,c,d = .57,.19,.19,.05), 2*10^7 undirected edges, seed 3
    queries     1,000 random-walk queries, sizes uniform in 4..32, even ids dense / odd
                sparse, seed 33
    path        batches of 64 queries of similar size (grouped by vertex count) through the
                batched runner: hipGraph-replayed NLF + GQL
                filter, RI orders, hipGraph-replayed persistent enumeration with the
                autotuned grid size and intersection strategy (tuned once on the first
                batch, untimed); a single-launch pass over all 1,000 queries is also timed

Counts are exact, capped at --cap embeddings per query (default 10^7, the same cap for the
GPU run and the CPU baseline).  The CPU baseline is the project's C restatement of the
reference path (oracle/, the checker; AVX2 set intersections as in the reference's build),
run single-threaded per query on this host over a bounded sample of the same queries; the
GPU counts and candidate-set digests of the sampled queries are compared with it.

The default run also reports the rmat20-learned comparison (one timed pass each) under
"learned_vs_ri" (--no-learned skips it).

The dense 16-vertex Chung-Lu 10^5 workload is specified without a cap; its queries hold more than 10^9 embeddings each (no query
finishes on either side within any budget this run can afford), so the default workload
also runs the 200 queries uncapped once with a per-query work budget (--uncapped-budget-ms
of wave time each) and reports how many finished, timed-out queries as timed out with the
count they had reached (a lower bound, not a result), beside the capped headline.
"""
from __future__ import annotations

import argparse
import collections
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "exact embeddings/sec & queries/sec on 1 MI355X; learned vs RI order speedup"


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["chunglu100k", "chunglu100k-labelled", "chunglu100k-aux", "chunglu100k-dyn",
                                           "chunglu10k", "rmat20", "rmat20-learned", "rmat22-serve"],
                    default="chunglu100k")
    ap.add_argument("--serve-mem-gb", type=float, default=32.0,
                    help="rmat22-serve: device memory budget for query slots and DFS stacks")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="rmat22-serve: arrival rate as a fraction of the measured capacity")
    ap.add_argument("--train-queries", type=int, default=5000,
                    help="rmat20-learned: training queries (10 logged orders each: RI, GQL, 8 sampled)")
    ap.add_argument("--learn-label", choices=["busy", "nodes", "plain-nodes"], default="busy",
                    help="cost-model target: wave time per query, internal search-tree nodes of the cache + filter + "
                         "failing-set search, or of the same search without failing sets (plain)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="chunglu100k: skip the learned-vs-RI section (R-MAT 2^20 mix, about a minute)")
    ap.add_argument("--rank-margin", type=float, default=0.5,
                    help="rmat20-learned: keep RI unless the model predicts a gain of more than this (log2)")
    ap.add_argument("--rank-samples", type=int, default=32,
                    help="rmat20-learned: sampled orders ranked per query besides RI and GQL")
    ap.add_argument("--queries", type=int, default=0, help="number of queries (0 = the workload's own)")
    ap.add_argument("--labels", type=int, default=8,
                    help="chunglu100k-labelled: vertex labels 0..N-1, label l drawn with weight 1/(l+1)")
    ap.add_argument("--cap", type=int, default=None,
                    help="embeddings per query (default 10^7; chunglu100k-dyn: 10^5, the output limit of the "
                         "reference's order comparison)")
    ap.add_argument("--uncapped-budget-ms", type=int, default=-1,
                    help="chunglu100k / chunglu10k: per-query wave-time budget of the uncapped pass (0 skips it; "
                         "default 10 s per query for the 200 dense 16-vertex Chung-Lu 10^5 queries, 2000 s for the one 8-vertex Chung-Lu 10^4 query: "
                         "about half a second of the whole GPU either way)")
    ap.add_argument("--per-query", action="store_true",
                    help="chunglu100k: one enumeration launch per query instead of one batched launch")
    ap.add_argument("--blocks", type=int, default=0,
                    help="persistent enumeration blocks of 4 waves (0 = library default / autotuned)")
    ap.add_argument("--batch", type=int, default=64, help="rmat20: queries per graph-replayed batch")
    ap.add_argument("--runners", type=int, default=2,
                    help="rmat20: overlapped runners (batch k+1.. filtered and launched while batch k enumerates)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="rmat20: one runner, each batch waited for before the next is filtered")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="wall-clock budget of the CPU baseline sample (0 disables it)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per launch (from rocprofv3 --pmc), if present")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES bytes (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    args = ap.parse_args()
    if args.cap is None:
        args.cap = 100_000 if args.workload == "chunglu100k-dyn" else 10_000_000
    if args.uncapped_budget_ms < 0:
        args.uncapped_budget_ms = 2_000_000 if args.workload == "chunglu10k" else 10_000
    return args


def copy_calibration(nbytes):
    import torch
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "ok": bool(y[-1].item() == 1)}), flush=True)


def new_rec():
    return {"emb": 0, "inter": 0, "elems_read": 0, "elems_out": 0, "elems_scanned": 0, "probes": 0, "enum_ms": 0.0,
            "enum_launches": 0, "capped": 0, "counts": [], "gql_ms": 0.0, "nlf_ms": 0.0, "cand_ms": 0.0,
            "filter_ms": 0.0}


def add_result(rec, r):
    if r is None:
        rec["counts"].append(0)
        return
    rec["emb"] += r.ans
    rec["inter"] += r.intersections
    rec["elems_read"] += r.elems_read
    rec["elems_out"] += r.elems_out
    rec["elems_scanned"] += r.elems_scanned
    rec["probes"] += r.probe_words
    rec["capped"] += int(r.capped)
    rec["counts"].append(r.ans)


# ------------------------------------------------------------------------- chunglu100k
def setup_chunglu(args):
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd import _lib as L
    from subgraphmatching_amd.generators import chung_lu, query_set, with_labels

    c10k = args.workload == "chunglu10k"
    g = chung_lu(10_000, 2.5, 7.4, seed=1) if c10k else chung_lu(100_000, 2.3, 16.0, seed=2)
    labelled = args.workload == "chunglu100k-labelled"
    aux = args.workload == "chunglu100k-aux"  # tables + explore_Intersection_All<true,...>, per query
    dyn = args.workload == "chunglu100k-dyn"  # computeWeightArray + explore_Intersection_DynamicOrder_DAF
    if labelled:  # the same CSR with skewed labels; the queries' walks carry them
        g = with_labels(g, args.labels, seed=23, skew=1.0)
    if c10k:
        queries = query_set(g, 1, 8, False, seed=11)
    else:
        queries = query_set(g, args.queries or 200, 16, True, seed=22)
    ctx = sgm.default_context()
    cans = [sgm.CandidateParam(g, q, ctx) for q in queries]

    def step():
        rec = new_rec()
        t0 = time.perf_counter()
        ok = sgm.filter_batch(cans, L.SGM_FILTER_GQL)
        rec["nlf_ms"], rec["gql_ms"], rec["cand_ms"] = ctx.filter_ms()
        rec["filter_ms"] = rec["nlf_ms"] + rec["gql_ms"] + rec["cand_ms"]
        t1 = time.perf_counter()
        if dyn:
            live = [i for i, good in enumerate(ok) if good]
            # start orders (BFS from DAFStartVertex), DAF weights and missing tables are made
            # by the library inside the call; one launch over every query
            batch = sgm.enumerate_batch([cans[i] for i in live], None, L.SGM_ENUM_DYN_DAF, cap=args.cap,
                                        blocks=args.blocks)
            results = [None] * len(queries)
            for i, r in zip(live, batch):
                results[i] = r
            if batch:
                rec["enum_ms"] += batch[0].time
                rec["enum_launches"] += 1
        elif aux:
            live = [i for i, good in enumerate(ok) if good]
            # FilterParamConstruct's tables (inc/table.h:7-137) of every query are built in one
            # batch inside the call, RI orders too; then one launch over every query
            batch = sgm.enumerate_batch([cans[i] for i in live], None, L.SGM_ENUM_AUX_TABLES_FS, cap=args.cap,
                                        blocks=args.blocks)
            results = [None] * len(queries)
            for i, r in zip(live, batch):
                results[i] = r
            if batch:
                rec["enum_ms"] += batch[0].time
                rec["enum_launches"] += 1
        elif args.per_query:
            orders = [sgm.Order_RI(q) for q in queries]
            results = []
            for can, order, good in zip(cans, orders, ok):
                r = sgm.enumerate_query(can, order, L.SGM_ENUM_CACHE_FILTER_FS, cap=args.cap,
                                        blocks=args.blocks) if good else None
                if r is not None:
                    rec["enum_ms"] += r.time
                    rec["enum_launches"] += 1
                results.append(r)
        else:
            live = [i for i, good in enumerate(ok) if good]
            # RI orders computed by the library inside the call (orders=None)
            batch = sgm.enumerate_batch([cans[i] for i in live], None,
                                        L.SGM_ENUM_CACHE_FILTER_FS, cap=args.cap, blocks=args.blocks)
            results = [None] * len(queries)
            for i, r in zip(live, batch):
                results[i] = r
            if batch:
                rec["enum_ms"] += batch[0].time
                rec["enum_launches"] += 1
        for r in results:
            add_result(rec, r)
        rec["wall_s"] = time.perf_counter() - t0
        rec["filter_wall_s"] = t1 - t0
        return rec

    name = "chunglu_V100000_g2.3_d16_seed2__200x16v_dense_rw_seed22__gql_ri_cache_filter_fs"
    if c10k:
        name = "chunglu_V10000_g2.5_d7.4_seed1__1x8v_sparse_rw_seed11__gql_ri_cache_filter_fs"
    if labelled:
        name = name.replace("seed2__", f"seed2_labels{args.labels}_skew1_seed23__")
    if aux:
        name = name.replace("gql_ri_cache_filter_fs", "gql_tables_ri_aux_fs")
    if dyn:
        name = name.replace("gql_ri_cache_filter_fs", "gql_daf_weights_dynamic_order")
    config = {"workload": name,
              "queries_per_step": len(queries), "cap_per_query": args.cap,
              "order": "DAF dynamic order (BFS from DAFStartVertex, computeWeightArray weights)" if dyn else "RI",
              "filter": "labelled NLF+GQL" if labelled else "NLF+GQL",
              "enumeration": ("aux tables + failing sets" if aux else
                              "dynamic-order search (explore_Intersection_DynamicOrder_DAF)" if dyn else
                              "cache+filter+failing-sets"),
              "launch": "per-query" if args.per_query else "one filter batch + one persistent enumeration launch"}

    def uncapped(budget_ms):
        """The same queries uncapped, one launch, each with a work budget of budget_ms of wave
        time: finished queries are exact results; timed-out ones report the count they had
        reached, a lower bound of their embedding count and never a result."""
        ok = sgm.filter_batch(cans, L.SGM_FILTER_GQL)
        live = [c for c, good in zip(cans, ok) if good]
        t0 = time.perf_counter()
        res = sgm.enumerate_batch(live, None, L.SGM_ENUM_CACHE_FILTER_FS, time_limit_ms=budget_ms)
        wall = time.perf_counter() - t0
        done = [r for r in res if not r.timed_out]
        lb = sorted(r.ans for r in res if r.timed_out)
        emb = sum(r.ans for r in done)
        return {"time_limit_ms_per_query": budget_ms, "budget_kind": "wave time per query, summed over its waves",
                "queries": len(res), "completed": len(done), "timed_out": len(lb),
                "completed_embeddings": emb, "launch_wall_ms": 1e3 * wall,
                "kernel_ms": res[0].time if res else 0.0,
                "embeddings_per_s_completed": emb / wall if done and wall > 0 else None,
                "timed_out_count_lower_bound": ({"min": lb[0], "median": lb[len(lb) // 2], "max": lb[-1]}
                                                if lb else None)}

    def gpu_digests():
        """Candidate-set digests of every query (the oracle module's digest form)."""
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sgm.filter_batch(cans, L.SGM_FILTER_GQL)
        return [O.cand_digest(c.candidates) for c in cans]

    extra = ({} if (labelled or aux) else {"digests": gpu_digests} if dyn else
             {"uncapped": uncapped, "digests": gpu_digests})
    return g, queries, step, config, extra


# ------------------------------------------------------------------------------ rmat20
def setup_rmat(args):
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd import _lib as L
    from subgraphmatching_amd.generators import query_set, rmat

    g = rmat(20, 20_000_000, seed=3)
    nq = args.queries or 1000
    queries = query_set(g, nq, lambda i, r: int(r.integers(4, 33)), lambda i, r: i % 2 == 0, seed=33)
    B = args.batch
    runner = sgm.Batch(g, slots=B)
    # batches group queries of similar size (a batch runs 2 x its largest query's GQL phases)
    by_size = sorted(range(nq), key=lambda i: queries[i].getVertexCnt())
    chunks = [[queries[i] for i in by_size[lo:lo + B]] for lo in range(0, nq, B)]
    # autotune once on the first batch (untimed)
    runner.filter(chunks[0])
    tuned, tuned_ms = runner.autotune(cap=args.cap)
    if args.blocks:
        tuned["blocks"] = args.blocks
    print(f"[bench] autotuned on the first batch: {tuned} ({tuned_ms:.2f} ms)", flush=True)

    # runners on their own streams: batch k+1 is filtered and launched while batch k
    # enumerates, and its grid takes the CUs batch k's waves leave at the end
    runners = [runner] + ([sgm.Batch(g, slots=B, ctx=runner.ctx) for _ in range(args.runners - 1)]
                          if not args.no_overlap else [])

    def step():
        rec = new_rec()
        t0 = time.perf_counter()
        counts = [0] * nq
        pending = collections.deque()  # launched batches: (runner, ok flags, position of its first query)

        def collect(p):
            r, ok, start = p
            res = r.enumerate_wait()
            _, e_ms = r.last_ms()
            rec["enum_ms"] += e_ms
            rec["enum_launches"] += 1
            for j, (good, x) in enumerate(zip(ok, res)):
                add_result(rec, x if good else None)
                counts[by_size[start + j]] = rec["counts"][-1]

        pos = 0
        for k, part in enumerate(chunks):
            r = runners[k % len(runners)]
            while any(p[0] is r for p in pending):  # the runner's previous batch, and those before it
                collect(pending.popleft())
            ok, _, _ = r.filter(part)
            f_ms, _ = r.last_ms()
            rec["filter_ms"] += f_ms
            rec["gql_ms"] += f_ms
            r.enumerate_start(cap=args.cap, **tuned)
            pending.append((r, ok, pos))
            pos += len(part)
        while pending:
            collect(pending.popleft())
        rec["counts"] = counts  # in query order
        rec["wall_s"] = time.perf_counter() - t0
        rec["filter_wall_s"] = rec["filter_ms"] / 1e3
        return rec

    def single_launch():
        """The same work as one filter batch and one persistent launch over all queries."""
        ctx = sgm.default_context()
        cans = [sgm.CandidateParam(g, q, ctx) for q in queries]
        t0 = time.perf_counter()
        ok = sgm.filter_batch(cans, L.SGM_FILTER_GQL)
        nlf_ms, gql_ms, cand_ms = ctx.filter_ms()
        live = [i for i, good in enumerate(ok) if good]
        res = sgm.enumerate_batch([cans[i] for i in live], [sgm.Order_RI(queries[i]) for i in live],
                                  L.SGM_ENUM_CACHE_FILTER_FS, cap=args.cap, blocks=tuned["blocks"])
        wall = time.perf_counter() - t0
        counts = [0] * len(queries)
        for i, r in zip(live, res):
            counts[i] = r.ans
        for c in cans:
            c.close()
        return {"wall_ms": 1e3 * wall, "filter_device_ms": nlf_ms + gql_ms + cand_ms,
                "enum_device_ms": res[0].time if res else 0.0}, counts

    config = {"workload": "rmat_s20_e2e7_abcd.57.19.19.05_seed3__1000x4-32v_mixed_rw_seed33__gql_ri_cache_filter_fs",
              "queries_per_step": nq, "cap_per_query": args.cap, "order": "RI", "filter": "NLF+GQL",
              "enumeration": "cache+filter+failing-sets",
              "launch": f"hipGraph-replayed batches of {B} (filter graph + persistent enumeration graph)"
                        + ("" if args.no_overlap else "; two runners on two streams, batch k+1 filtered and "
                                                      "launched while batch k enumerates"),
              "autotuned": tuned}
    return g, queries, step, config, {"single_launch": single_launch}


# ----------------------------------------------------------------------- rmat20-learned
def setup_rmat_learned(args):
    import numpy as np
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd import _lib as L
    from subgraphmatching_amd.generators import query_set, rmat
    from subgraphmatching_amd.learned import LABELS, collect_training, train_model

    g = rmat(20, 20_000_000, seed=3)
    nq = args.queries or 1000
    mix = lambda count, seed: query_set(g, count, lambda i, r: int(r.integers(4, 33)),  # noqa: E731
                                        lambda i, r: i % 2 == 0, seed=seed)
    queries = mix(nq, 33)
    train_q = mix(args.train_queries, 44)
    B = args.batch
    runner = sgm.Batch(g, slots=B)
    by_size = sorted(range(nq), key=lambda i: queries[i].getVertexCnt())  # batches of similar size
    chunks = [[queries[i] for i in by_size[lo:lo + B]] for lo in range(0, nq, B)]
    runner.filter(chunks[0])
    tuned, _ = runner.autotune(cap=args.cap)
    log = lambda m: print(f"[bench] {m}", flush=True)  # noqa: E731

    def fit(label):
        t0 = time.perf_counter()
        items, labels, info = collect_training(runner, train_q, n_samples=8, cap=args.cap, seed=1,
                                               time_limit_s=5.0, label=label, log=log)
        collect_s = time.perf_counter() - t0
        model, losses = train_model(items, labels, seed=7, epochs=30, minibatch=256, lr=2e-3)
        tm = model.timing()
        log(f"cost model ({label}): {len(items)} triples ({info['cut_off']} cut off) logged in {collect_s:.1f}s, "
            f"loss {losses[0]:.3f} -> {losses[-1]:.3f}, training {tm['train_ms']:.0f} ms")
        return {"model": model, "label": label, "items": len(items), "info": info, "losses": losses, "timing": tm,
                "collect_s": collect_s}

    # the ranking model (label --learn-label), and the search-tree node-count model of the
    # plain search (failing sets off), whose predictions are checked against measured node
    # counts and whose ranking is timed beside it
    rank = fit(args.learn_label)
    nodes = fit("plain-nodes") if args.learn_label != "plain-nodes" else rank
    stats = {}

    # two runners on their own streams, as in the rmat20 workload: batch k+1 is filtered, ranked
    # and launched while batch k enumerates (the same for the RI and the learned passes)
    runners = [runner, sgm.Batch(g, slots=B, ctx=runner.ctx)]

    def run(fitted):
        """One pass over the 1,000 queries: filter, order (RI when fitted is None), enumerate."""
        rec = new_rec()
        t_start = time.perf_counter()
        st = {"learned_used": 0, "pred": [], "meas": [], "nodes": 0, "busy_ns": 0, "rank_s": 0.0, "orders": []}
        pending = collections.deque()  # launched batches: (runner, ok flags, predictions)

        def collect(p):
            r, ok, pb = p
            res = r.enumerate_wait()
            rec["enum_ms"] += r.last_ms()[1]
            rec["enum_launches"] += 1
            for i, (good, x) in enumerate(zip(ok, res)):
                add_result(rec, x if good else None)
                if good:
                    st["nodes"] += sum(x.nodes[:-1])
                    st["busy_ns"] += x.busy_ns
                    if fitted is not None and fitted["label"] != "plain-nodes":
                        st["pred"].append(float(pb[i]))
                        st["meas"].append(float(LABELS[fitted["label"]](x)))

        for k, part in enumerate(chunks):
            r = runners[k % len(runners)]
            while any(p[0] is r for p in pending):  # the runner's previous batch, and those before it
                collect(pending.popleft())
            ok, _, _ = r.filter(part)
            f_ms, _ = r.last_ms()
            rec["filter_ms"] += f_ms
            rec["gql_ms"] += f_ms
            orders = pb = None
            if fitted is not None:
                t1 = time.perf_counter()
                orders, pb, _, fl = r.rank_orders(fitted["model"], n_samples=args.rank_samples, seed=5,
                                                  margin=args.rank_margin)
                st["rank_s"] += time.perf_counter() - t1
                st["learned_used"] += int(fl.sum())
                st["orders"].append((orders, ok, [float(x) for x in pb]))
            r.enumerate_start(orders=orders, cap=args.cap, **tuned)
            pending.append((r, ok, pb))
        while pending:
            collect(pending.popleft())
        rec["wall_s"] = time.perf_counter() - t_start
        rec["filter_wall_s"] = rec["filter_ms"] / 1e3
        return rec, st

    def step():
        rec, st = run(rank)
        stats["rank"] = st
        return rec

    def measured_plain(st):
        """Predicted vs measured plain node counts of the orders a pass chose: the same batches
        without failing sets (untimed; launches past 5 s leave their unfinished queries out)."""
        pred, meas, cut = [], [], 0
        for part, (orders, ok, pb) in zip(chunks, st["orders"]):
            runner.filter(part)
            res = runner.enumerate(orders=orders, method=L.SGM_ENUM_CACHE_FILTER, cap=args.cap, time_limit_s=5.0)
            for i, (good, r) in enumerate(zip(ok, res)):
                if good and not r.stopped:
                    pred.append(pb[i])
                    meas.append(LABELS["plain-nodes"](r))
                cut += bool(good and r.stopped)
        return pred, meas, cut

    def pvm(pred, meas, label, cut=0):
        pred, meas = np.array(pred, dtype=float), np.array(meas, dtype=float)
        err = np.abs(pred - meas)
        return {"label": label, "queries": int(len(pred)), "cut_off": cut,
                "abs_error_log2": {"mean": float(err.mean()), "median": float(np.median(err))} if len(err) else None,
                "within_1_log2": float((err <= 1.0).mean()) if len(err) else None,
                "pearson_log2": float(np.corrcoef(pred, meas)[0, 1]) if len(pred) > 2 else None,
                "mean_predicted_log2": float(pred.mean()) if len(pred) else None,
                "mean_measured_log2": float(meas.mean()) if len(meas) else None}

    def cm(f):
        tm, ms = f["timing"], f["timing"]["train_ms"]
        t2 = f["model"].timing()
        return {"label": f["label"], "training_triples": f["items"], "cut_off": f["info"]["cut_off"], "epochs": 30,
                "logged_s": f["collect_s"], "loss_first_epoch": float(f["losses"][0]),
                "loss_last_epoch": float(f["losses"][-1]), "train_ms": ms,
                "train_tflops": tm["train_flops"] / ms / 1e9 if ms else 0.0,
                "train_mfma_util_vs_bf16_dense_2.5pf": tm["train_flops"] / ms / 1e9 / 2500.0 if ms else 0.0,
                "infer_ms": t2["infer_ms"], "infer_flops": t2["infer_flops"], "layers": 3, "hidden": 64,
                "dtype": "bf16 operands, fp32 accumulation"}

    def compare(K, learned_counts, learned_s):
        ri = [run(None) for _ in range(K)]
        ri_s = sum(r["wall_s"] for r, _ in ri)
        st = stats["rank"]
        if rank["label"] == "plain-nodes":
            p, m, cut = measured_plain(st)
            rank_pvm = pvm(p, m, "plain-nodes", cut)
        else:
            rank_pvm = pvm(st["pred"], st["meas"], rank["label"])
        out = {"ri_queries_per_s": nq * K / ri_s, "learned_queries_per_s": nq * K / learned_s,
               "learned_vs_ri_speedup": ri_s / learned_s,
               "counts_identical": all(r["counts"] == learned_counts for r, _ in ri),
               "internal_nodes_per_step": {"ri": ri[0][1]["nodes"], "learned": st["nodes"]},
               "wave_busy_ms_per_step": {"ri": ri[0][1]["busy_ns"] / 1e6, "learned": st["busy_ns"] / 1e6},
               "label": rank["label"],
               "learned_orders_used_per_step": st["learned_used"],
               "ranking_host_s_per_step": st["rank_s"],
               "predicted_vs_measured": rank_pvm,
               "costmodel": cm(rank)}
        if nodes is not rank:
            # the node-count model: its predictions against the plain node counts of the
            # orders it picks, and the same 1,000 queries timed with those orders
            nrec, nst = run(nodes)
            p, m, cut = measured_plain(nst)
            out["node_count_model"] = {
                "predicted_vs_measured_nodes": pvm(p, m, "plain-nodes", cut),
                "queries_per_s_ranked_by_it": nq / nrec["wall_s"],
                "speedup_vs_ri_ranked_by_it": (ri_s / K) / nrec["wall_s"],
                "counts_identical": nrec["counts"] == learned_counts,
                "internal_nodes_per_step": nst["nodes"], "wave_busy_ms_per_step": nst["busy_ns"] / 1e6,
                "learned_orders_used": nst["learned_used"], "costmodel": cm(nodes)}
        return out

    config = {"workload": "rmat_s20_e2e7_abcd.57.19.19.05_seed3__1000x4-32v_mixed_rw_seed33__gql_learned_order_cache_filter_fs",
              "queries_per_step": nq, "cap_per_query": args.cap, "order": f"learned (RI fallback, margin {args.rank_margin} log2)",
              "filter": "NLF+GQL", "enumeration": "cache+filter+failing-sets",
              "launch": f"hipGraph-replayed batches of {B}", "autotuned": tuned,
              "training": f"{args.train_queries} queries (seed 44) x 10 orders, labels "
                          + {"busy": "log2(1 + wave microseconds)",
                             "nodes": "log2(1 + internal search-tree nodes, failing sets on)",
                             "plain-nodes": "log2(1 + internal search-tree nodes, failing sets off)"}[args.learn_label]
                          + ("; a second model on log2(1 + internal nodes, failing sets off)"
                             if args.learn_label != "plain-nodes" else ""),
              "ranked_orders_per_query": f"RI + GQL + {args.rank_samples} sampled"}
    return g, queries, step, config, {"compare": compare}


def learned_section(args):
    """The learned-vs-RI comparison of the R-MAT 2^20 1,000-query mix (cost models trained on
    5,000 training queries x 10 orders), run after the headline workload: one untimed and one
    timed pass with learned orders, one with RI orders, and the node-count model's pass."""
    import copy
    a = copy.copy(args)
    a.queries = 0
    g, queries, step, config, extra = setup_rmat_learned(a)
    step()
    t0 = time.perf_counter()
    rec = step()
    out = extra["compare"](1, rec["counts"], time.perf_counter() - t0)
    out["workload"] = config["workload"]
    out["training"] = config["training"]
    return out


# ------------------------------------------------------------------------- rmat22-serve
def setup_rmat_serve(args):
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd.generators import query_set, rmat
    from subgraphmatching_amd.learned import collect_training, train_model
    from subgraphmatching_amd.serving import admit_slots, poisson_arrivals, serve

    cap = args.cap if args.cap != 10_000_000 else 1_000_000
    g = rmat(22, 1 << 26, seed=5)
    nq = args.queries or 1200
    mix = lambda count, seed: query_set(g, count, lambda i, r: int(r.integers(8, 33)),  # noqa: E731
                                        lambda i, r: i % 2 == 0, seed=seed)
    queries = mix(nq, 55)
    train_q = mix(600, 66)
    ctx = sgm.default_context()
    num_cus = 256
    slots, stack, per_slot = admit_slots(g.getVertexCnt(), g.getMaxDgree(), int(args.serve_mem_gb * (1 << 30)),
                                         waves=num_cus * 4 * 4)
    runner = sgm.Batch(g, slots=slots, ctx=ctx)
    opts = {"blocks": num_cus * 4}
    items, labels, info = collect_training(runner, train_q, n_samples=8, cap=cap, seed=2, time_limit_s=5.0)
    model, losses = train_model(items, labels, seed=8, epochs=30, minibatch=256, lr=2e-3)
    rank = lambda r: (lambda o: (o[0], o[3]))(r.rank_orders(model, n_samples=32, seed=5, margin=0.5))  # noqa: E731
    # capacity: the whole stream at once, RI orders
    zero = [0.0] * nq
    serve(runner, queries[:2 * slots], zero[:2 * slots], cap, enum_opts=opts)  # warm the captured graphs
    capacity = serve(runner, queries, zero, cap, enum_opts=opts).queries_per_s
    rate = args.serve_load * capacity
    arrivals = poisson_arrivals(nq, rate, seed=77)
    print(f"[bench] serving: {slots} slots admitted ({per_slot / 2**20:.0f} MiB each, stacks {stack / 2**30:.1f} GiB), "
          f"capacity {capacity:.0f} q/s, Poisson rate {rate:.0f} q/s; cost model on {len(items)} triples, "
          f"loss {losses[0]:.2f} -> {losses[-1]:.2f}", flush=True)
    last = {}

    def step():
        res = serve(runner, queries, arrivals, cap, order_fn=rank, enum_opts=opts)
        last["learned"] = res
        rec = new_rec()
        rec["emb"] = sum(res.counts)
        rec["counts"] = res.counts
        rec["filter_ms"] = rec["gql_ms"] = res.device_ms["filter"]
        rec["enum_ms"] = res.device_ms["enumerate"]
        rec["enum_launches"] = res.batches
        rec["capped"] = sum(int(c >= cap) for c in res.counts)
        rec["wall_s"] = res.elapsed_s
        return rec

    def compare(K, learned_counts, learned_s):
        ri = serve(runner, queries, arrivals, cap, enum_opts=opts)
        le = last["learned"]
        summary = lambda r: {"queries_per_s": r.queries_per_s, "latency_ms": r.latency_ms,  # noqa: E731
                             "batches": r.batches, "mean_batch": r.mean_batch, "device_ms": r.device_ms}
        return {"arrival_rate_q_per_s": rate, "capacity_q_per_s_ri_back_to_back": capacity, "slots": slots,
                "learned": dict(summary(le), learned_orders_used=le.learned_used),
                "ri": summary(ri), "counts_identical": ri.counts == learned_counts,
                "p99_ri_over_learned": ri.latency_ms["p99"] / le.latency_ms["p99"]}

    config = {"workload": "rmat_s22_e2^26_seed5__1200x8-32v_mixed_rw_seed55__poisson_serving_learned_order",
              "queries_per_step": nq, "cap_per_query": cap, "order": "learned (RI fallback)",
              "filter": "NLF+GQL", "enumeration": "cache+filter+failing-sets",
              "launch": f"dynamic batches of up to {slots} (hipGraph-replayed)",
              "memory_budget_gb": args.serve_mem_gb, "arrival_load": args.serve_load}
    return g, queries, step, config, {"compare": compare, "serving": True}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path shards by query "
              "and nothing in this build uses a second GPU", file=sys.stderr)
        sys.exit(2)

    if (args.workload == "rmat20" and not args.no_overlap) or args.workload == "rmat20-learned" or \
            (args.workload == "chunglu100k" and args.learned):
        # the two overlapped runners' streams need hardware queues of their own (HIP shares
        # its default 4 among the context, stop, null and runner streams); set before the
        # first HIP call of the process
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
    t_setup = time.perf_counter()
    setup = {"rmat20": setup_rmat, "rmat20-learned": setup_rmat_learned,
             "rmat22-serve": setup_rmat_serve}.get(args.workload, setup_chunglu)
    g, queries, step, config, extra = setup(args)
    print(f"[bench] {args.workload}: graph V={g.getVertexCnt()} E={g.getEdgeCnt()} maxdeg={g.getMaxDgree()}; "
          f"{len(queries)} queries; setup {time.perf_counter() - t_setup:.1f}s", flush=True)

    for i in range(args.warmup):
        r = step()
        print(f"[bench] warmup {i}: {r['wall_s']:.3f}s", flush=True)
    recs = []
    t_all = time.perf_counter()
    for i in range(args.steps):
        r = step()
        recs.append(r)
        print(f"[bench] step {i}: {r['wall_s']:.3f}s  filter {r['filter_ms']:.1f} ms (gql {r['gql_ms']:.1f} ms)  "
              f"enum {r['enum_ms']:.1f} ms  emb {r['emb']}", flush=True)
    total_s = time.perf_counter() - t_all
    counts0 = recs[0]["counts"]
    assert all(r["counts"] == counts0 for r in recs), "counts differ between steps"

    K = len(recs)
    emb = sum(r["emb"] for r in recs)
    nq = len(queries) * K
    inter = sum(r["inter"] for r in recs)
    enum_ms = sum(r["enum_ms"] for r in recs)
    filter_ms = sum(r["filter_ms"] for r in recs)
    launches = sum(r["enum_launches"] for r in recs)
    # Algorithmic bytes of the enumeration kernel, per the strategy each intersection used (a
    # lower bound of what the search must move): 4 B per key of an intersection's shorter list
    # and per word it probes in the longer one (one bitmap word, or one word per binary-search
    # step and the final load), 4 B per classified candidate (its filter word).
    algo_bytes = sum(4 * (r["probes"] + r["elems_scanned"]) for r in recs)
    # the merge model (every intersection reads |A|+|B| and writes |A n B|, every candidate
    # 4 B + 1 bit), kept for comparison: not a bound for bitmap or binary-search intersections
    merge_bytes = sum(4 * (r["elems_read"] + r["elems_out"]) + 4.125 * r["elems_scanned"] for r in recs)
    enum_kernel = {"chunglu100k-aux": "k_enum_aux", "chunglu100k-dyn": "k_enum_dyn"}.get(args.workload, "k_enum_cache")
    achieved = algo_bytes / (enum_ms / 1e3) / 1e9 if enum_ms > 0 else 0.0
    traffic = measured = None
    traffic_note = "no measured traffic for this workload"
    # the workloads whose launches' HBM traffic profiles/traffic.json records
    traffic_key = {"chunglu100k": "k_enum_cache_bytes_per_launch", "chunglu100k-aux": "k_enum_aux_bytes_per_launch",
                   "chunglu100k-dyn": "k_enum_dyn_bytes_per_launch"}.get(args.workload)
    if traffic_key and os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tj = json.load(f)
            if tj.get("source_stamp") == source_stamp() and tj.get(traffic_key):
                traffic = tj.get(traffic_key)
                traffic_note = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel build (profiles/traffic.json)"
            elif tj.get("source_stamp") != source_stamp():
                traffic_note = "profiles/traffic.json was measured on other kernel sources: not used"
        except Exception as e:  # noqa: BLE001
            traffic_note = f"profiles/traffic.json unreadable: {e}"
    if traffic and launches and not args.per_query:
        measured = traffic / (enum_ms / launches / 1e3) / 1e9

    extras = {}
    if "compare" in extra:
        extras["learned_vs_ri"] = extra["compare"](K, counts0, total_s)
        key = "serving" if extra.get("serving") else "learned_vs_ri"
        extras[key] = extras.pop("learned_vs_ri")
        print(f"[bench] {key}: {extras[key]}", flush=True)
    if "single_launch" in extra:
        sl, sl_counts = extra["single_launch"]()
        extras["single_launch"] = dict(sl, counts_match=sl_counts == counts0)
        print(f"[bench] single launch over all queries: {sl}", flush=True)
    if "uncapped" in extra and args.uncapped_budget_ms > 0:
        extras["uncapped"] = extra["uncapped"](args.uncapped_budget_ms)
        print(f"[bench] uncapped pass: {extras['uncapped']}", flush=True)
    if args.workload == "chunglu100k" and args.learned:
        extras["learned_vs_ri"] = learned_section(args)
        print(f"[bench] learned_vs_ri: {extras['learned_vs_ri']}", flush=True)

    cpu = None
    gpu_digests = extra["digests"]() if "digests" in extra and args.cpu_seconds > 0 else None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(g, queries, counts0, config["cap_per_query"], args.cpu_seconds,
                           {"chunglu100k-aux": "aux_fs", "chunglu100k-dyn": "dyn_daf"}.get(args.workload, "cache"),
                           gpu_digests)

    out = {
        "metric": METRIC,
        "value": emb / total_s,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_s / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": config,
        "queries_per_s": nq / total_s,
        "intersections_per_s": inter / total_s,
        "embeddings_per_step": emb // K,
        "capped_queries_per_step": recs[0]["capped"],
        "device_ms_per_step": {"filter": filter_ms / K, "gql_filter": sum(r["gql_ms"] for r in recs) / K,
                               "enumerate": enum_ms / K},
        "roofline": {"kernel": enum_kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "achieved_basis": "strategy bytes per launch (4 B per intersection key and per bitmap word or "
                                       "binary-search word it probes, 4 B per classified candidate's filter "
                                       "word; a lower bound of what the search moves) / average launch time",
                     "traffic_basis": traffic_note,
                     "measured_gbs": measured,
                     "measured_frac": measured / HBM_PEAK_GBS if measured is not None else None,
                     "algorithmic_bytes_per_launch": algo_bytes / max(launches, 1),
                     "avg_launch_ms": enum_ms / max(launches, 1),
                     "merge_model_gbs": merge_bytes / (enum_ms / 1e3) / 1e9 if enum_ms > 0 else 0.0,
                     "dominant_kernel": enum_kernel if enum_ms >= filter_ms else "k_gql_phase"},
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


def source_stamp():
    """First 16 hex digits of the SHA-256 of the enumeration kernel's sources: a traffic
    record applies only to the build it was measured on."""
    h = hashlib.sha256()
    for f in ("sgm_enum.hip", "sgm_internal.h", "sgm_wave.h"):
        with open(os.path.join(ROOT, "subgraphmatching_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def _cpu_worker(job):
    """One CPU worker: GQL + RI + cache/filter/fs (or aux-table/fs, or the DAF dynamic-order
    search) enumeration of its queries, one after the other on one core, until its wall-clock
    budget ends (the query running then is cut off by the oracle's time limit and left out).
    Returns (query, count, enumeration thread-ms, candidate digest) per finished query."""
    g, items, cap, budget_s, method = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.perf_counter()
    out = []
    for i, q in items:
        left = budget_s - (time.perf_counter() - t0)
        if left <= 0:
            break
        ok, c = O.gql(g, q)
        dig = O.cand_digest(c)
        order = O.order_ri(q)
        cnt, ms = 0, 0.0
        if ok:
            if method == "aux_fs":
                st = O.enum_all(g, q, c, order, cap=cap, time_limit_ms=max(1.0, 1e3 * left), fs=True)
            elif method == "dyn_daf":  # start order and weights are computed before the clock starts
                st = O.enum_dynamic(g, q, c, "DAF", cap=cap, time_limit_ms=max(1.0, 1e3 * left))
            else:
                st = O.enum_cache(g, q, c, order, cap=cap, time_limit_ms=max(1.0, 1e3 * left))
            if st["timed_out"]:
                break
            cnt, ms = st["count"], st["enum_ms"]
        out.append((i, cnt, ms, dig))
    return out, time.perf_counter() - t0


def cpu_baseline(g, queries, gpu_counts, cap, budget_s, method="cache", gpu_digests=None):
    """The project's C restatement of the same path (the oracle: single-threaded per query as
    the reference runs, AVX2 intersections as the reference's build) on a bounded sample of
    the same queries: P worker processes, one per core, each taking every P-th query for
    budget_s seconds of wall time.  P = min(16, host cores): 16 cores are one GPU's share of
    the GPU box's host.  Reports the P-core rate and worker 0's single-core rate; the GPU
    counts and candidate-set digests of every sampled query are compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    P = max(1, min(16, os.cpu_count() or 1, len(queries)))
    jobs = [(g, [(i, queries[i]) for i in range(w, len(queries), P)], cap, budget_s, method) for w in range(P)]
    if P == 1:
        results = [_cpu_worker(jobs[0])]
    else:
        import multiprocessing as mp
        pool = mp.get_context("spawn").Pool(P)  # fresh interpreters: no GPU state in the workers
        try:
            results = pool.map(_cpu_worker, jobs)
        finally:
            pool.close()  # workers exit on their own (no SIGTERM, as leaving a `with` block would send)
            pool.join()
    done = sum(len(r) for r, _ in results)
    emb = sum(c for r, _ in results for _, c, _, _ in r)
    wall = max(w for _, w in results)
    parity = all(c == gpu_counts[i] for r, _ in results for i, c, _, _ in r)
    dig = None
    if gpu_digests is not None:
        pairs = [(d, gpu_digests[i]) for r, _ in results for i, _, _, d in r]
        dig = {"checked": len(pairs), "match": all(a == b for a, b in pairs)}
    r0, w0 = results[0]
    single = sum(c for _, c, _, _ in r0) / w0 if w0 > 0 else 0.0
    print(f"[bench] cpu baseline: {done} queries on {P} cores in {wall:.1f}s, parity={parity}, digests={dig}",
          flush=True)
    return {"value": emb / wall if wall > 0 else 0.0, "unit": "embeddings/s", "cores": P, "kind": "port",
            "sample": f"{done} of the {len(queries)} queries (GQL + "
                      + {"aux_fs": "RI + aux-table/fs", "dyn_daf": "DAF weights + dynamic-order"}.get(
                          method, "RI + cache/filter/fs")
                      + f" enumeration, cap {cap}) "
                      f"on {P} worker processes, one query at a time each, {wall:.1f} s wall",
            "intersection": "AVX2" if O.isect_is_avx2() else "scalar merge",
            "queries_per_s": done / wall if wall > 0 else 0.0,
            "single_core_value": single, "single_core_queries_per_s": len(r0) / w0 if w0 > 0 else 0.0,
            "enum_thread_ms": sum(ms for r, _ in results for _, _, ms, _ in r),
            "counts_match_gpu": bool(parity), "candidate_sets_match_gpu": dig}


if __name__ == "__main__":
    main()
