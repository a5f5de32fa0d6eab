"""Writes tests/golden/rmat_fixtures.json: oracle results for sampled queries of the R-MAT
benchmark configurations, at their own sizes.

  rmat20    R-MAT scale 20 (a,b,c,d = .57,.19,.19,.05), 2*10^7 undirected edges, seed 3; the
            1,000-query mix of seed 33 (sizes uniform in 4..32, even ids dense, odd sparse);
            64 queries spread over the sizes (every 16th in size order: one batch of 64);
            GQL filter, RI order, cache + filter + failing-set search capped at 10^7
  rmat22    R-MAT scale 22, average degree 32 (2^26 undirected edges), seed 5; the serving
            stream of seed 55 (1,200 queries of 8..32 vertices, even ids dense); 16 queries
            (every 75th); the same path capped at 10^6

Per query: GQL success, candidate counts and the SHA-256 digest of the candidate sets, the RI
order and the capped count (min(#embeddings, cap), order-independent, so it also checks
learned orders).  Every sampled query reaches its cap, so for six R-MAT 2^20 queries and three
R-MAT 2^22 queries the file also holds root-range slices (tests/slices.py): exact, uncapped
sub-searches whose count, per-depth nodes, intersections and failing-set cut-offs the
oracle computed for all four enumeration variants.  The oracle's GQL takes seconds to
minutes per query on these graphs, so the GPU tests read these results instead of
recomputing them.  Usage:

    python tests/golden/make_rmat_fixtures.py            (about 15 minutes on 8 cores)
    python tests/golden/make_rmat_fixtures.py --slices   (adds the slices to the file)
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
from subgraphmatching_amd.generators import query_set, rmat  # noqa: E402

sys.path.insert(0, HERE)
from make_config_slices import find_ranges  # noqa: E402

OUT = os.path.join(HERE, "rmat_fixtures.json")
ENUM_LIMIT_MS = 120_000.0

_G = None
_QS = None


def rmat20():
    g = rmat(20, 20_000_000, seed=3)
    qs = query_set(g, 1000, lambda i, r: int(r.integers(4, 33)), lambda i, r: i % 2 == 0, seed=33)
    by_size = sorted(range(len(qs)), key=lambda i: qs[i].getVertexCnt())
    return g, qs, [by_size[k] for k in range(0, len(qs), 16)][:64]


def rmat22():
    g = rmat(22, 1 << 26, seed=5)
    qs = query_set(g, 1200, lambda i, r: int(r.integers(8, 33)), lambda i, r: i % 2 == 0, seed=55)
    return g, qs, list(range(0, len(qs), 75))


def _one(args):
    i, cap = args
    q = _QS[i]
    ok, c = O.gql(_G, q)
    rec = {"id": i, "n": q.getVertexCnt(), "ok": ok, "cand_counts": [len(x) for x in c],
           "cand_digest": O.cand_digest(c), "order": O.order_ri(q)}
    if ok:
        st = O.enum_cache(_G, q, c, rec["order"], cap=cap, time_limit_ms=ENUM_LIMIT_MS)
        rec["count"] = None if st["timed_out"] else st["count"]
    else:
        rec["count"] = 0
    print(f"  q{i} n={rec['n']} ok={ok} count={rec['count']}", flush=True)
    return rec


def run(name, build, cap, procs):
    global _G, _QS
    _G, _QS, ids = build()
    print(f"{name}: V={_G.getVertexCnt()} adj={len(_G.adj)} sample={len(ids)}", flush=True)
    pool = mp.get_context("fork").Pool(procs)
    try:
        recs = pool.map(_one, [(i, cap) for i in ids], chunksize=1)
    finally:
        pool.close()
        pool.join()
    return {"cap": cap, "queries": recs}


def _slices(i):
    q = _QS[i]
    ok, c = O.gql(_G, q)
    order = O.order_ri(q)
    ranges = find_ranges(_G, q, c, order, 300 + i, 2, 0, 0, 1000.0, min_nodes=150) if ok else []
    print(f"  q{i}: {len(ranges)} slices", flush=True)
    return i, {"order": order, "ranges": ranges}


def add_slices(procs):
    global _G, _QS
    with open(OUT) as f:
        out = json.load(f)
    for name, build, count in (("rmat20", rmat20, 6), ("rmat22", rmat22, 3)):
        _G, _QS, _ = build()
        ids = [r["id"] for r in out[name]["queries"]]
        pick = [ids[k] for k in range(0, len(ids), max(1, len(ids) // count))][:count]
        pool = mp.get_context("fork").Pool(procs)
        try:
            res = pool.map(_slices, pick, chunksize=1)
        finally:
            pool.close()
            pool.join()
        out[name]["slices"] = {str(i): rec for i, rec in res if rec["ranges"]}
        print(name, "slices of", len(out[name]["slices"]), "queries", flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


def main():
    O.build()
    if "--slices" in sys.argv:
        add_slices(min(8, os.cpu_count() or 1))
        return
    procs = min(8, os.cpu_count() or 1)
    out = {"rmat20": run("rmat20", rmat20, 10_000_000, procs),
           "rmat22": run("rmat22", rmat22, 1_000_000, procs)}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
