"""Writes tests/golden/config_slices.json: full-size parity fixtures from the CPU oracle.

Inputs are this project's seeded benchmark inputs:
  chunglu10k   Chung-Lu 10^4 vertices, exponent 2.5, average degree 7.4 (seed 1); the 8-vertex
            sparse random-walk query of seed 11; GQL filter, RI order
  chunglu100k  Chung-Lu 10^5 vertices, exponent 2.3, average degree 16 (seed 2); the 200 dense
            16-vertex random-walk queries of seed 22; GQL filter, RI order

Whole searches on these inputs run for hours (the 8-vertex Chung-Lu 10^4 query has more than 1.5e13
embeddings), so the fixture holds, per query, root ranges whose searches the oracle
finishes within a second or two (found by a seeded search over candidate roots) and the
oracle's statistics for each range and each enumeration variant; plus the candidate-set
digests of every dense 16-vertex Chung-Lu 10^5 query.  Usage:

    python tests/golden/make_config_slices.py        (about 5 minutes on 8 cores)
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from slices import VARIANTS, cand_digest, oracle_stats  # noqa: E402
from subgraphmatching_amd.generators import chung_lu, query_set  # noqa: E402

OUT = os.path.join(HERE, "config_slices.json")
C2_QUERIES = list(range(0, 200, 25))  # the Chung-Lu 10^5 queries whose slices are recorded


def cl10k():
    g = chung_lu(10000, 2.5, 7.4, seed=1)
    return g, query_set(g, 1, 8, False, seed=11)


def cl100k():
    g = chung_lu(100_000, 2.3, 16.0, seed=2)
    return g, query_set(g, 200, 16, True, seed=22)


def find_ranges(g, q, cands, order, seed, singles, windows, width, probe_ms, min_nodes=0):
    """Seeded search for root ranges whose every variant finishes: `singles` one-root ranges
    and `windows` id windows of `width` ids.  With min_nodes == 0 a range must hold
    embeddings; otherwise its search tree must have at least min_nodes nodes (on the
    Chung-Lu 10^5 inputs every root's subtree either dies within about a thousand nodes or holds
    more than 10^8 embeddings, so exact slices there have no embeddings)."""
    rng = np.random.default_rng(seed)
    roots = np.array(cands[order[0]])
    found = []

    def accept(r):
        st = oracle_stats(O, g, q, cands, order, "cache_fs", r, probe_ms)
        if st is None or (st["count"] == 0 if not min_nodes else sum(st["nodes"]) < min_nodes):
            return False
        rec = {"roots": list(r)}
        for name, _, _ in VARIANTS:
            s = oracle_stats(O, g, q, cands, order, name, r, 20 * probe_ms)
            if s is None:
                return False
            rec[name] = s
        found.append(rec)
        return True

    n1 = n2 = 0
    for v in rng.permutation(roots)[:400].tolist():
        if n1 >= singles:
            break
        n1 += accept((v, v + 1))
    V = g.getVertexCnt()
    for lo in rng.integers(0, V - width, 200).tolist():
        if n2 >= windows:
            break
        n2 += accept((lo, lo + width))
    return found


def c2_digest(i):
    g, qs = cl100k()
    ok, c = O.gql(g, qs[i])
    return i, bool(ok), cand_digest(c), [len(x) for x in c]


def c2_slices(i):
    g, qs = cl100k()
    q = qs[i]
    ok, c = O.gql(g, q)
    order = O.order_ri(q)
    return i, {"order": order, "ranges": find_ranges(g, q, c, order, 100 + i, 4, 0, 0, 1000.0, min_nodes=150)}


def main():
    out = {"about": "oracle statistics of root-range slices (see tests/slices.py); written by "
                    "tests/golden/make_config_slices.py"}
    g1, (q1,) = cl10k()
    ok, c = O.gql(g1, q1)
    order = O.order_ri(q1)
    out["cl10k"] = {"graph": "chung_lu(10000, 2.5, 7.4, seed=1)", "query": "query_set(g, 1, 8, False, seed=11)[0]",
                      "order": order, "cand_digest": cand_digest(c),
                      "ranges": find_ranges(g1, q1, c, order, 11, 3, 0, 0, 2000.0)}
    print("cl10k", len(out["cl10k"]["ranges"]), "ranges", flush=True)
    pool = mp.get_context("spawn").Pool(min(8, os.cpu_count() or 1))
    dig = sorted(pool.map(c2_digest, range(200)))
    out["cl100k"] = {"graph": "chung_lu(100000, 2.3, 16.0, seed=2)", "queries": "query_set(g, 200, 16, True, seed=22)",
                      "ok": [d[1] for d in dig], "cand_digest": [d[2] for d in dig], "cand_counts": [d[3] for d in dig],
                      "slices": {}}
    print("cl100k digests", flush=True)
    for i, rec in sorted(pool.map(c2_slices, C2_QUERIES)):
        out["cl100k"]["slices"][str(i)] = rec
        print("cl100k query", i, len(rec["ranges"]), "ranges", flush=True)
    pool.close()
    pool.join()
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
