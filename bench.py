#!/usr/bin/env python3
"""Headline benchmark: exact embeddings/s and queries/s on one MI355X for the
filter -> order -> enumerate path of unlabeled subgraph matching.

Workload (one step = one pass of the path over one batch):
  data graph  Chung-Lu, 100,000 vertices, degree exponent 2.3, average degree 16, seed 2
              (unlabeled, simple, sorted u32 CSR; generated once, resident in HBM)
  queries     200 dense 16-vertex random-walk queries (average degree > 3), seed 22
  path        NLF + GQL candidate filtering for all 200 queries in one batch of launches,
              RI order (host), cache + filter + failing-set enumeration of all 200 queries
              in one persistent launch (per-query caps and statistics)
  output      exact counts, capped at --cap embeddings per query (default 10^7, the same
              cap for the GPU run and the CPU baseline; uncapped counts of these queries
              run past 10^9 and cannot be enumerated in bounded time)

The CPU baseline is the project's C restatement of the reference path (oracle/, the
checker), run single-threaded on this host over a bounded sample of the same queries; the
GPU counts of the sampled queries are compared with it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=200)
    ap.add_argument("--cap", type=int, default=10_000_000)
    ap.add_argument("--per-query", action="store_true",
                    help="one enumeration launch per query instead of one batched persistent launch")
    ap.add_argument("--blocks", type=int, default=0,
                    help="persistent enumeration blocks of 4 waves (0 = library default)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="wall-clock budget of the CPU baseline sample (0 disables it)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per launch (from rocprofv3 --pmc), if present")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES bytes (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    return ap.parse_args()


def copy_calibration(nbytes):
    import torch
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "ok": bool(y[-1].item() == 1)}), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path shards by query "
              "and nothing in this build uses a second GPU", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd import _lib as L
    from subgraphmatching_amd.generators import chung_lu, query_set

    t_setup = time.perf_counter()
    g = chung_lu(100_000, 2.3, 16.0, seed=2)
    queries = query_set(g, args.queries, 16, True, seed=22)
    ctx = sgm.default_context()
    cans = [sgm.CandidateParam(g, q, ctx) for q in queries]
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] graph V={g.getVertexCnt()} E={g.getEdgeCnt()} maxdeg={g.getMaxDgree()}; "
          f"{len(queries)} queries; setup {setup_s:.1f}s", flush=True)

    def step():
        rec = {"emb": 0, "inter": 0, "elems_read": 0, "elems_out": 0, "elems_scanned": 0,
               "enum_ms": 0.0, "enum_launches": 0, "capped": 0, "counts": []}
        t0 = time.perf_counter()
        ok = sgm.filter_batch(cans, L.SGM_FILTER_GQL)
        nlf_ms, gql_ms, cand_ms = ctx.filter_ms()
        t1 = time.perf_counter()
        orders = [sgm.Order_RI(q) for q in queries]
        if args.per_query:
            results = [sgm.enumerate_query(can, order, L.SGM_ENUM_CACHE_FILTER_FS, cap=args.cap, blocks=args.blocks)
                       if good else None for can, order, good in zip(cans, orders, ok)]
        else:
            live = [i for i, good in enumerate(ok) if good]
            batch = sgm.enumerate_batch([cans[i] for i in live], [orders[i] for i in live],
                                        L.SGM_ENUM_CACHE_FILTER_FS, cap=args.cap, blocks=args.blocks)
            results = [None] * len(queries)
            for i, r in zip(live, batch):
                results[i] = r
            if batch:
                rec["enum_ms"] += batch[0].time
                rec["enum_launches"] += 1
        for r in results:
            if r is None:
                rec["counts"].append(0)
                continue
            rec["emb"] += r.ans
            rec["inter"] += r.intersections
            rec["elems_read"] += r.elems_read
            rec["elems_out"] += r.elems_out
            rec["elems_scanned"] += r.elems_scanned
            if args.per_query:
                rec["enum_ms"] += r.time
                rec["enum_launches"] += 1
            rec["capped"] += int(r.capped)
            rec["counts"].append(r.ans)
        t2 = time.perf_counter()
        rec.update(wall_s=t2 - t0, filter_wall_s=t1 - t0, nlf_ms=nlf_ms, gql_ms=gql_ms, cand_ms=cand_ms)
        return rec

    for i in range(args.warmup):
        r = step()
        print(f"[bench] warmup {i}: {r['wall_s']:.3f}s", flush=True)
    recs = []
    t_all = time.perf_counter()
    for i in range(args.steps):
        r = step()
        recs.append(r)
        print(f"[bench] step {i}: {r['wall_s']:.3f}s  filter {r['filter_wall_s']:.3f}s "
              f"(gql {r['gql_ms']:.1f} ms)  enum {r['enum_ms']:.1f} ms  emb {r['emb']}", flush=True)
    total_s = time.perf_counter() - t_all
    counts0 = recs[0]["counts"]
    assert all(r["counts"] == counts0 for r in recs), "counts differ between steps"

    K = len(recs)
    emb = sum(r["emb"] for r in recs)
    nq = len(queries) * K
    inter = sum(r["inter"] for r in recs)
    enum_ms = sum(r["enum_ms"] for r in recs)
    gql_ms = sum(r["gql_ms"] for r in recs)
    launches = sum(r["enum_launches"] for r in recs)
    # algorithmic bytes of the enumeration kernel (merge-intersection model): every
    # intersection reads |A|+|B| and writes |A n B| u32; every iterated candidate costs a
    # 4-byte id plus one filter bit
    algo_bytes = sum(4 * (r["elems_read"] + r["elems_out"]) + 4.125 * r["elems_scanned"] for r in recs)
    achieved = algo_bytes / (enum_ms / 1e3) / 1e9 if enum_ms > 0 else 0.0
    traffic = None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                traffic = json.load(f).get("k_enum_cache_bytes_per_launch")
        except Exception:
            traffic = None
    dominant = "k_enum_cache" if enum_ms >= gql_ms else "k_gql_phase"

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(g, queries, counts0, args.cap, args.cpu_seconds)

    out = {
        "metric": "exact embeddings/sec & queries/sec on 1 MI355X; learned vs RI order speedup",
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
        "config": {"workload": "chunglu_V100000_g2.3_d16_seed2__200x16v_dense_rw_seed22__gql_ri_cache_filter_fs",
                   "queries_per_step": len(queries), "cap_per_query": args.cap, "order": "RI",
                   "filter": "NLF+GQL", "enumeration": "cache+filter+failing-sets"},
        "queries_per_s": nq / total_s,
        "intersections_per_s": inter / total_s,
        "embeddings_per_step": emb // K,
        "capped_queries_per_step": recs[0]["capped"],
        "device_ms_per_step": {"gql_filter": gql_ms / K, "enumerate": enum_ms / K,
                               "nlf": sum(r["nlf_ms"] for r in recs) / K,
                               "candidate_counts": sum(r["cand_ms"] for r in recs) / K},
        "roofline": {"kernel": "k_enum_cache", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "dominant_kernel": dominant,
                     "algorithmic_bytes_per_launch": algo_bytes / max(launches, 1),
                     "avg_launch_ms": enum_ms / max(launches, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(g, queries, gpu_counts, cap, budget_s):
    """Single-threaded CPU restatement of the same path on a bounded sample of the same
    queries (GQL filter + RI order + cache/filter/failing-set enumeration, same cap)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    t0 = time.perf_counter()
    done = emb = 0
    enum_ms = 0.0
    parity = True
    for i, q in enumerate(queries):
        if time.perf_counter() - t0 > budget_s:
            break
        ok, c = O.gql(g, q)
        order = O.order_ri(q)
        cnt = 0
        if ok:
            st = O.enum_cache(g, q, c, order, cap=cap)
            cnt = st["count"]
            enum_ms += st["enum_ms"]
        parity &= cnt == gpu_counts[i]
        emb += cnt
        done += 1
    wall = time.perf_counter() - t0
    print(f"[bench] cpu baseline: {done} queries in {wall:.1f}s, parity={parity}", flush=True)
    return {"value": emb / wall, "unit": "embeddings/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of the {len(queries)} queries (GQL + RI + cache/filter/fs enumeration, "
                      f"cap {cap}), single thread, {wall:.1f} s wall",
            "queries_per_s": done / wall, "enum_thread_ms": enum_ms, "counts_match_gpu": bool(parity)}


if __name__ == "__main__":
    main()
