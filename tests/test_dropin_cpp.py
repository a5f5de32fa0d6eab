"""The drop-in C++ surface (include/sgm/{api,filter,order,explore}.h): the headers compile
against the reference's call pattern, and the row driver built on them
(subgraphmatching_amd/cpp/overall_performance.cpp, the reference harness's rows and CSV
columns) gives the oracle's counts and intersection counts on the GPU."""
import os
import subprocess

import pytest

import oracle as O
from subgraphmatching_amd.generators import chung_lu, query_set, with_labels
from subgraphmatching_amd.graph import save_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "subgraphmatching_amd")
DRIVER = os.path.join(PKG, "sgm_overall")
ORDERS = os.path.join(PKG, "sgm_orders")

# the reference harness's calling sequence (src/overall_performance.cpp:243-262)
USAGE = r'''
#include "sgm/explore.h"
#include "sgm/filter.h"
#include "sgm/order.h"
size_t row(const Graph& data, const Query& query, std::atomic<bool>& stop_flag) {
    double time = 0; size_t ans = 0;
    Order order;
    Order_RI(query, order);
    CandidateParam can(data, query);
    if (Filter_GQL(data, query, can) == false) return 0;
    Config cfg;
    cfg.useFilter = true;
    cfg.can = &can;
    ReuseParam reuse(query, order);
    cfg.useCache = true;
    cfg.reuse = &reuse;
    explore_Intersection_Cache_DAF_Filter(data, query, order, cfg, time, ans, stop_flag, 100000);
    Order o2;
    Order_GQL(data, query, can, o2);
    explore_Intersection_All<false, false, false, false, false>(data, query, o2, cfg, time, ans, stop_flag);
    uint32_t cnt;
    const uint32_t* nb = can.edge_matrix[0][1]->getNeb(0, cnt);
    (void)nb;
    size_t ops = reuse.plan->getTotOp() + reuse.plan->getIterVSetAt(1) + reuse.plan->getSetOps().size();
    // the other filters and orders of the harness rows (src/overall_performance.cpp:66-190)
    CandidateParam c2(data, query);
    if (Filter_CFL(data, query, c2) || Filter_CECI(data, query, c2) || Filter_DAF(data, query, c2)) {
        Order oq, oc, og, oe;
        Order_QSI(data, query, c2, oq);
        Order_CFL(data, query, c2, oc);
        Order_GuP(data, query, c2, og);
        Order_CECI(data, query, oe);
        Config c3;
        c3.useFilter = true;
        c3.can = &c2;
        explore_Intersection_All<true, false, false, false, false>(data, query, oq, c3, time, ans, stop_flag, 100000);
        explore_Intersection_All<false, false, true, false, false>(data, query, oq, c3, time, ans, stop_flag, 100000);
        (void)Profiler::getInst().gup_cd.cnt;
    }
    // the dynamic-order rows of the order comparison (src/order_ablations.cpp:131-172)
    uint32_t** weight_array = new uint32_t*[query.getVertexCnt()];
    for (uint32_t i = 0; i < query.getVertexCnt(); ++i) weight_array[i] = new uint32_t[can.candidates_count[i] + 1];
    Order bfs_order, match_order;
    computeWeightArray(data, query, can, bfs_order, weight_array);
    explore_Intersection_DynamicOrder_DAF(data, query, bfs_order, match_order, cfg, weight_array, time, ans, stop_flag,
                                          100000);
    TreeNode* tree = nullptr;
    BFS(query, DAFStartVertex(data, query), tree, bfs_order);
    explore_Intersection_DynamicOrder_VEQ(data, query, bfs_order, match_order, cfg, time, ans, stop_flag);
    delete[] tree;
    return ans + ops + Profiler::getInst().total_intersection + can.candidates_count[0] + can.candidates_exsist[0][0];
}
'''


def test_headers_compile_with_reference_call_pattern(tmp_path):
    src = tmp_path / "usage.cpp"
    src.write_text(USAGE)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


PLAN_PRINT = r'''
#include <cstdio>
#include "sgm/api.h"
int main(int argc, char** argv) {
    Query q;
    q.Load(argv[1]);
    Order order;
    for (int i = 2; i < argc; i++) order.push_back((VertexID)std::atoi(argv[i]));
    ReuseParam reuse(q, order);
    const PlanIR& p = *reuse.plan;
    for (const auto& op : p.getTotSetOpsArr()) {
        std::printf("op %u %u %u %d", op.getVSetID(), op.getDepth(), op.getVDepth(),
                    op.getParentVSetID() == VertexSetIR::INVALID ? -1 : (int)op.getParentVSetID());
        for (auto e : op.getEdgeIdx()) std::printf(" %u", e);
        std::printf("\n");
    }
    for (uint32_t d = 0; d < order.size(); d++)
        std::printf("it %d\n", p.getIterVSetAt(d) == VertexSetIR::INVALID ? -1 : (int)p.getIterVSetAt(d));
    return 0;
}
'''


def test_reuse_param_plan_matches_oracle(tmp_path):
    """ReuseParam owns a PlanIR with the reference's accessors (inc/api.h:167-182,
    inc/code_gen.h:682-847); its ops and iteration sets equal the oracle's plan.  Host only:
    the plan is built without a device."""
    src = tmp_path / "plan.cpp"
    src.write_text(PLAN_PRINT)
    exe = tmp_path / "plan"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe), "-L", PKG, "-lsgm_hip", f"-Wl,-rpath,{PKG}", "-pthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    g = chung_lu(2000, 2.5, 7.4, seed=81)
    for i, q in enumerate(query_set(g, 8, lambda i, r: int(r.integers(4, 17)), lambda i, r: bool(i % 2), seed=82)):
        qp = tmp_path / f"q{i}.txt"
        save_text(q, str(qp))
        order = O.order_ri(q)
        out = subprocess.run([str(exe), str(qp), *map(str, order)], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
        ref = O.plan(q, order)
        ops, its = [], []
        for ln in out.stdout.splitlines():
            f = ln.split()
            if f[0] == "op":
                ops.append({"depth": int(f[2]), "parent": None if f[4] == "-1" else int(f[4]), "vdepth": int(f[3]),
                            "edges": [int(x) for x in f[5:]]})
            else:
                its.append(None if f[1] == "-1" else int(f[1]))
        assert ops == ref["ops"] and its == ref["iter_set"], q.name


def test_pruning_variants_are_rejected_at_compile_time(tmp_path):
    src = tmp_path / "bad.cpp"
    src.write_text('#include "sgm/explore.h"\n'
                   "void f(const Graph& d, const Query& q, const Order& o, const Config& c, std::atomic<bool>& s) {\n"
                   "  double t; size_t a; explore_Intersection_All<false,true,false,false,false>(d, q, o, c, t, a, s); }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "accelerated path" in r.stderr


def _write_inputs(tmp_path, g, qs):
    dp = tmp_path / "data.txt"
    save_text(g, str(dp))
    paths = []
    for i, q in enumerate(qs):
        qp = tmp_path / f"q{i}.txt"
        save_text(q, str(qp))
        paths.append(str(qp))
    return str(dp), paths


def _run(data, query, limit, rows, timeout=120):
    r = subprocess.run([DRIVER, data, query, str(limit), "300", *rows], capture_output=True, text=True,
                       timeout=timeout)
    return r


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="row driver not built (run __graft_entry__.build())")
def test_driver_reports_missing_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = chung_lu(200, 2.5, 4.0, seed=1)
    qs = query_set(g, 1, 4, False, seed=2)
    data, qp = _write_inputs(tmp_path, g, qs)
    r = _run(data, qp[0], 1000, ["RI"])
    assert r.returncode == 1 and "no usable GPU" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("labels", [0, 3])
def test_driver_rows_match_oracle(tmp_path, labels):
    """The harness rows over reference-format text files (labels in the 'v' lines when
    labels > 0) give the oracle's counts and intersection counts."""
    assert os.path.exists(DRIVER), "row driver not built"
    g = chung_lu(2000, 2.3, 8.0, seed=3)
    if labels:
        g = with_labels(g, labels, seed=72)
    qs = query_set(g, 6, lambda i, r: int(r.integers(4, 10)), lambda i, r: bool(i % 2), seed=71)
    data, paths = _write_inputs(tmp_path, g, qs)
    limit = 100000
    for q, qp in zip(qs, paths):
        print("query", qp, q.getVertexCnt(), flush=True)
        r = _run(data, qp, limit, ["RI", "QSI", "GQL", "CFL", "CECI", "DAF", "RI+cache+filter+fs"])
        assert r.returncode == 0, r.stderr
        lines = r.stdout.strip().splitlines()
        assert lines[0] == "data,query_name,algorithm,time_ms,ans,output_limit,intersection_cnt"
        got = {}
        for ln in lines[1:]:
            f = ln.split(",")
            assert f[0] == "data" and f[5] == str(limit)
            got[f[2]] = (int(f[4]), int(f[6]))
        # RI row: NLF + RI order + plain candidate-space search
        ok, c = O.nlf(g, q)
        ref = O.enum_all(g, q, c, O.order_ri(q), cap=limit) if ok else {"count": 0, "intersections": 0}
        _check(got["RI"], ref, limit)
        # QSI row: NLF + QSI order + plain search
        if ok:
            _check(got["QSI"], O.enum_all(g, q, c, O.order_tree("QSI", g, q, c), cap=limit), limit)
        else:
            assert got["QSI"][0] == 0
        # CFL / CECI rows: their filter and order + plain search; DAF row: DAF filter + RI order
        # + search with failing sets
        for row in ("CFL", "CECI", "DAF"):
            okt, ct = O.filter_tree(row, g, q)
            if not okt:
                assert got[row][0] == 0, row
                continue
            order = O.order_ri(q) if row == "DAF" else O.order_tree(row, g, q, ct)
            _check(got[row], O.enum_all(g, q, ct, order, cap=limit, fs=row == "DAF"), limit)
        # GQL row: GQL filter + GQL order + plain search
        ok, c = O.gql(g, q)
        if ok:
            order = O.order_gql(g, q, [len(x) for x in c])
            _check(got["GQL"], O.enum_all(g, q, c, order, cap=limit), limit)
            _check(got["RI+cache+filter+fs"], O.enum_cache(g, q, c, O.order_ri(q), cap=limit), limit)
        else:
            assert got["GQL"][0] == 0 and got["RI+cache+filter+fs"][0] == 0


def _check(got, ref, limit):
    ans, inter = got
    assert ans == ref["count"]
    if ref["count"] < limit:  # uncapped: the whole search ran, so the work is exact too
        assert inter == ref["intersections"]


@pytest.mark.gpu
def test_order_driver_rows_match_oracle(tmp_path):
    """The order-comparison rows (GQL filter, then a static order with the plain search, or a
    dynamic-order search) over reference-format files give the oracle's counts, and its
    intersection counts where uncapped."""
    assert os.path.exists(ORDERS), "order driver not built"
    g = with_labels(chung_lu(2000, 2.4, 6.0, seed=83), 4, seed=84)
    qs = query_set(g, 6, lambda i, r: int(r.integers(4, 11)), lambda i, r: bool(i % 2), seed=85)
    data, paths = _write_inputs(tmp_path, g, qs)
    limit = 100000
    orders = ["QSI", "RI", "GQL", "CFL", "CECI", "DAF", "VEQ"]
    for q, qp in zip(qs, paths):
        print("query", qp, q.getVertexCnt(), flush=True)
        r = subprocess.run([ORDERS, data, qp, str(limit), "300", *orders], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.strip().splitlines()
        assert lines[0] == "data,query_name,explore,order,order_overhead,time_ms,ans,output_limit,intersection_cnt"
        got = {}
        for ln in lines[1:]:
            f = ln.split(",")
            assert f[0] == "data" and f[2] == "filter" and f[7] == str(limit)
            got[f[3]] = (int(f[6]), int(f[8]))
        ok, c = O.gql(g, q)
        if not ok:
            assert all(v == (0, 0) for v in got.values())
            continue
        static = {"RI": O.order_ri(q), "GQL": O.order_gql(g, q, [len(x) for x in c])}
        for kind in ("QSI", "CFL", "CECI"):
            static[kind] = O.order_tree(kind, g, q, c)
        for kind, order in static.items():
            _check(got[kind], O.enum_all(g, q, c, order, cap=limit), limit)
        for kind in ("DAF", "VEQ"):
            _check(got[kind], O.enum_dynamic(g, q, c, kind, cap=limit), limit)
