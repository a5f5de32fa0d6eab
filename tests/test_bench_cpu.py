"""bench.py's CPU-baseline leg on the CPU: the worker pool finishes its sample and exits
cleanly (closed and joined, never terminated: no "Aborted ... SIGTERM" traces), counts and
candidate-set digests are compared with the given device results, and a mismatch is
reported as one."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path[:0] = [ROOT, ROOT + "/oracle"]
import bench, oracle as O
from subgraphmatching_amd.generators import chung_lu, query_set
g = chung_lu(1500, 2.3, 6.0, seed=5)
qs = query_set(g, 6, 5, True, seed=6)
counts, digs = [], []
for q in qs:
    ok, c = O.gql(g, q)
    counts.append(O.enum_cache(g, q, c, O.order_ri(q), cap=20000)["count"] if ok else 0)
    digs.append(O.cand_digest(c))
r = bench.cpu_baseline(g, qs, counts, 20000, 5.0, "cache", digs)
assert r["counts_match_gpu"] and r["candidate_sets_match_gpu"] == {"checked": 6, "match": True}, r
assert r["cores"] == min(16, __import__("os").cpu_count(), 6) and r["intersection"] == "AVX2", r
bad = list(digs); bad[2] = "0" * 64
r = bench.cpu_baseline(g, qs, counts, 20000, 5.0, "cache", bad)
assert r["candidate_sets_match_gpu"]["match"] is False, r
print("OK")
"""


def test_cpu_baseline_pool_exits_cleanly():
    p = subprocess.run([sys.executable, "-c", SCRIPT.replace("ROOT", repr(ROOT))], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr[-3000:]
    assert "OK" in p.stdout
    assert "SIGTERM" not in p.stderr and "Aborted" not in p.stderr, p.stderr[-3000:]
