"""Helpers shared by the full-size parity tests and the script that writes their fixtures
(tests/golden/make_config_slices.py).

A "slice" is a root range [lo, hi): only the depth-0 candidates with lo <= id < hi are
searched.  The reference's depth-0 loop never prunes and resets every deeper failing-set
slot before reading it (inc/explore.h:1030-1066, inc/prune.h:119-137), so a slice's count,
per-depth nodes, intersections and failing-set cut-offs are exact parts of the whole
search's (tests/test_oracle_pins.py checks that disjoint slices add up).  On the benchmark
graphs whole searches run for hours, while slices of cheap roots finish in a second on the
CPU oracle: they carry exact, uncapped parity to full-size inputs.
"""
from __future__ import annotations

from oracle import cand_digest  # noqa: F401  (one definition: the oracle's)

# the four enumeration variants compared on every slice: (name, cache search?, failing sets?)
VARIANTS = [("cache_fs", True, True), ("cache", True, False), ("aux", False, False), ("aux_fs", False, True)]


def oracle_stats(O, g, q, cands, order, variant, roots, time_limit_ms=0.0):
    """The oracle's statistics of one slice (count, intersections, nodes, fs_prunes) or
    None when it ran out of time."""
    name, cache, fs = next(v for v in VARIANTS if v[0] == variant)
    fn = O.enum_cache if cache else O.enum_all
    st = fn(g, q, cands, order, fs=fs, time_limit_ms=time_limit_ms, roots=roots)
    if st["timed_out"]:
        return None
    return {"count": st["count"], "intersections": st["intersections"], "nodes": st["nodes"],
            "fs_prunes": st["fs_prunes"]}


def gpu_stats(res) -> dict:
    return {"count": res.ans, "intersections": res.intersections, "nodes": res.nodes, "fs_prunes": res.fs_prunes}
