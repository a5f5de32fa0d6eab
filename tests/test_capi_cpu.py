"""CPU-side checks of the C ABI library: it loads without a GPU, exports every symbol
include/sgm_c.h declares, and its host-side planning (RI / GQL orders, intersection-cache
plan) equals the oracle.  No device compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels
from brute import query_from_edges

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "sgm_c.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sgm_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = header_symbols()
    assert syms, "no declarations parsed"
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(L.EXPORTS) == syms


def test_version_and_no_device_errors():
    assert L.lib().sgm_version().startswith(b"sgm-mi355x")
    n = C.c_int(-1)
    assert L.lib().sgm_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        assert L.lib().sgm_ctx_create(0, C.byref(h)) == L.SGM_ENODEV


def test_orders_and_plan_match_oracle():
    g = chung_lu(3000, 2.5, 7.4, seed=31)
    qs = query_set(g, 40, lambda i, r: int(r.integers(3, 33)), lambda i, r: bool(i % 2), seed=32)
    for q in qs:
        o = sgm.Order_RI(q)
        assert o == O.order_ri(q)
        assert sgm.plan_export(q, o) == O.plan(q, o)
        # the GQL order's rule, evaluated through the oracle on fake counts, also fixes plans
        counts = np.random.default_rng(len(o)).integers(1, 50, q.getVertexCnt())
        og = O.order_gql(g, q, counts)
        assert sgm.plan_export(q, og) == O.plan(q, og)


def test_labelled_plans_match_oracle():
    """Ops keyed by (label, edge list): the library's plans of labelled queries equal the
    oracle's, and labels change some plans (fewer shared ops)."""
    g = with_labels(chung_lu(3000, 2.5, 7.4, seed=31), 3, seed=33)
    qs = query_set(g, 40, lambda i, r: int(r.integers(3, 33)), lambda i, r: bool(i % 2), seed=34)
    changed = 0
    for q in qs:
        assert q.is_labelled()
        for o in (sgm.Order_RI(q), O.order_gql(g, q, np.random.default_rng(len(q.adj)).integers(1, 50, q.getVertexCnt()))):
            p = sgm.plan_export(q, o)
            assert p == O.plan(q, o)
            changed += p != sgm.plan_export(sgm.Query(q.off, q.adj), o)
    assert changed > 0


def test_plans_over_sampled_orders_match_oracle():
    """Many connected orders per query (seeded samples), labelled and not: the library's plan
    (op table keyed by label and edge set) equals the oracle's restatement of PlanIR."""
    from subgraphmatching_amd.learned import sample_order
    g = chung_lu(3000, 2.5, 7.4, seed=41)
    gl = with_labels(g, 4, seed=42)
    qs = query_set(g, 12, lambda i, r: int(r.integers(4, 33)), lambda i, r: bool(i % 2), seed=43)
    qs += query_set(gl, 12, lambda i, r: int(r.integers(4, 33)), lambda i, r: bool(i % 2), seed=44)
    for k, q in enumerate(qs):
        for s in range(8):
            o = sample_order(q, seed=1000 * k + s)
            assert sgm.plan_export(q, o) == O.plan(q, o), (q.name, o)


def test_plan_rejects_disconnected_order():
    q = query_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    with pytest.raises(L.SgmError):
        sgm.plan_export(q, [0, 2, 1, 3])


def test_ctypes_structs_match_the_header(tmp_path):
    """The ctypes mirrors of sgm_stats / sgm_enum_opts have the C header's size and field
    offsets (a C program compiled against include/sgm_c.h prints them)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "sgm_c.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu %zu\\n", sizeof(sgm_stats), sizeof(sgm_enum_opts), '
                   'offsetof(sgm_stats, kernel_ms), offsetof(sgm_stats, busy_ns), '
                   'offsetof(sgm_enum_opts, hub_min_degree)); return 0; }\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == [C.sizeof(L.Stats), C.sizeof(L.EnumOpts), L.Stats.kernel_ms.offset, L.Stats.busy_ns.offset,
                   L.EnumOpts.hub_min_degree.offset]


def test_kcore_and_bfs_match_oracle():
    """The host helpers of the BFS-tree filters and orders (Query::KCore, BFS) equal the
    oracle's on random-walk queries of every size."""
    g = chung_lu(3000, 2.5, 7.4, seed=41)
    qs = query_set(g, 40, lambda i, r: int(r.integers(3, 33)), lambda i, r: bool(i % 2), seed=42)
    for q in qs:
        n = q.getVertexCnt()
        assert sgm.query_kcore(q) == O.kcore(q)
        qoff, qadj = L.u32(q.off), L.u32(q.adj)
        for s in (0, n - 1, n // 2):
            o, par, lev = (np.zeros(n, dtype=np.uint32) for _ in range(3))
            k = L.lib().sgm_query_bfs(L.ptr(qoff), L.ptr(qadj), n, s, L.ptr(o), L.ptr(par), L.ptr(lev))
            ro, rp, rl = O.bfs(q, s)
            assert k == n == len(ro)
            assert o.tolist() == ro and par.tolist() == rp and lev.tolist() == rl
