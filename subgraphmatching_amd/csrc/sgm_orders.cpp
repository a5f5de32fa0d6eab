// C ABI of the QSI, CFL, CECI and GuP orders (inc/order.h:537-668, 781-976) and of the query
// k-core / BFS helpers.  The order rules run on the host (sgm_tree.cpp, n <= 32); what
// touches the data graph runs on the device: NLF counts for the start vertices, the aux
// tables (sgm_build_tables), and the table-row sums behind CFL's path estimates and GuP's
// path counts (k_row_gather, u64 arithmetic wrapping as the reference's size_t does).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sgm_objects.h"

// NLF counts of every query vertex (NLFByVertex, inc/filter.h:24-58).  Unlabeled graph and
// query: the number of data vertices of degree >= deg_q(u) (the degree histogram ge_cnt);
// otherwise the labelled NLF kernel into a scratch membership array and the count kernel.
extern "C++" int query_nlf_counts(sgm_query* q, uint32_t* out)
{
    sgm_graph* g = q->g;
    sgm_ctx* c = g->ctx;
    const uint32_t n = q->q.n;
    if (q->nlf_ok && q->nlf_version == g->label_version) {
        std::memcpy(out, q->nlf_cnt, sizeof(uint32_t) * n);
        return SGM_OK;
    }
    if (!g->labelled && !q->labelled) {
        for (uint32_t u = 0; u < n; u++) q->nlf_cnt[u] = g->ge_cnt[std::min<uint32_t>(q->q.off[u + 1] - q->q.off[u], 32)];
    } else {
        HIPCHK(c, hipSetDevice(c->device));
        if (!g->has_lab) {
            const int rc = build_label_tables(c, g, nullptr);
            if (rc) return rc;
        }
        int rc = ensure_lreq(c, q);
        if (rc) return rc;
        sgm_qdesc d = q->hq;
        d.lab_on = 1;
        for (uint32_t u = 0; u < n; u++) d.qlab[u] = q->q.label(u);
        d.lreq = q->d_lreq;
        uint32_t *d_vm = nullptr, *d_cnt = nullptr;
        unsigned long long* d_deg = nullptr;
        sgm_qdesc* d_d = nullptr;
        hipError_t e = dalloc(&d_vm, std::max<uint32_t>(g->V, 1));
        if (e == hipSuccess) e = dalloc(&d_cnt, SGM_MAXQ);
        if (e == hipSuccess) e = dalloc(&d_deg, SGM_MAXQ);
        if (e == hipSuccess) e = dalloc(&d_d, 1);
        if (e == hipSuccess) {
            d.vmask = d_vm;
            e = hipMemcpyAsync(d_d, &d, sizeof d, hipMemcpyHostToDevice, c->stream);
        }
        if (e == hipSuccess) {
            sgm::launch_fill_u32(c->stream, d_cnt, SGM_MAXQ, 0u);
            sgm::launch_fill_u32(c->stream, reinterpret_cast<uint32_t*>(d_deg), 2 * SGM_MAXQ, 0u);
            sgm::launch_nlf_lab(c->stream, g->d_off, g->d_lab, g->d_run_off, g->d_run_lab, g->d_run_cnt, g->V, d_d, 1);
            sgm::launch_cand_stats(c->stream, g->d_off, g->V, d_d, 1, d_cnt, d_deg);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(q->nlf_cnt, d_cnt, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(d_vm);
        dfree(d_cnt);
        dfree(d_deg);
        dfree(d_d);
        if (e != hipSuccess)
            return set_err(c, e == hipErrorOutOfMemory ? SGM_ENOMEM : SGM_EHIP, "NLF counts: %s", hipGetErrorString(e));
    }
    q->nlf_ok = true;
    q->nlf_version = g->label_version;
    std::memcpy(out, q->nlf_cnt, sizeof(uint32_t) * n);
    return SGM_OK;
}

// The query's aux tables, built if any directed edge lacks one, and their entry counts as
// the reference's u32 Edges::_e_cnt (inc/graph.h:19), ecnt[a*n + b].
static int tables_and_counts(sgm_query* q, std::vector<uint32_t>& ecnt)
{
    const uint32_t n = q->q.n;
    bool all = true;
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++) all &= q->tables[(size_t)a * n + q->q.adj[p]].built;
    if (!all) {
        const int rc = sgm_build_tables(q);
        if (rc) return rc;
    }
    ecnt.assign((size_t)n * n, 0u);
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++) {
            const uint32_t b = q->q.adj[p];
            ecnt[(size_t)a * n + b] = (uint32_t)q->tables[(size_t)a * n + b].e_cnt;
        }
    return SGM_OK;
}

static int order_check(sgm_query* q, uint32_t* order, const char* what)
{
    if (!q || !order) return SGM_EINVAL;
    int rc = check_filtered(q, what);
    if (rc) return rc;
    if (q->empty) return set_err(q->g->ctx, SGM_EEMPTY, "%s: some candidate set is empty", what);
    if (q->q.n < 2) return set_err(q->g->ctx, SGM_EINVAL, "%s: the query needs 2 or more vertices", what);
    return SGM_OK;
}

// Device buffers of the estimators: two u64 vectors of the largest candidate count, one
// u64 vector per query vertex (GuP's path counts) and 32 totals.
struct EstBuf {
    sgm_ctx* c = nullptr;
    unsigned long long *a = nullptr, *b = nullptr, *per = nullptr, *tot = nullptr;
    std::vector<size_t> off;
    ~EstBuf()
    {
        dfree(a);
        dfree(b);
        dfree(per);
        dfree(tot);
    }
};

static int est_alloc(sgm_query* q, EstBuf& B, bool per_vertex)
{
    sgm_ctx* c = q->g->ctx;
    B.c = c;
    uint32_t mx = 1;
    B.off.assign(q->q.n + 1, 0);
    for (uint32_t u = 0; u < q->q.n; u++) {
        mx = std::max(mx, q->counts[u]);
        B.off[u + 1] = B.off[u] + q->counts[u];
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, dalloc(&B.a, mx));
    HIPCHK(c, dalloc(&B.b, mx));
    HIPCHK(c, dalloc(&B.tot, SGM_MAXQ));
    if (per_vertex) HIPCHK(c, dalloc(&B.per, std::max<size_t>(B.off[q->q.n], 1)));
    return SGM_OK;
}

int sgm_order_qsi(sgm_query* q, uint32_t* order)
{
    int rc = order_check(q, order, "Order_QSI");
    if (rc) return rc;
    std::vector<uint32_t> ecnt;
    if ((rc = tables_and_counts(q, ecnt))) return rc;
    sgm::order_qsi(q->q, ecnt.data(), order);
    return SGM_OK;
}

int sgm_order_ceci(sgm_query* q, uint32_t* order)
{
    if (!q || !order) return SGM_EINVAL;
    if (q->q.n < 2) return set_err(q->g->ctx, SGM_EINVAL, "Order_CECI: the query needs 2 or more vertices");
    int rc;
    uint32_t nlf[SGM_MAXQ];
    if ((rc = query_nlf_counts(q, nlf))) return rc;
    sgm::BfsTree t;
    sgm::bfs_tree(q->q, sgm::ceci_daf_start_vertex(q->q, nlf, q->g->V, false), t);
    if (t.order.size() != q->q.n) return set_err(q->g->ctx, SGM_EINVAL, "Order_CECI: the query is not connected");
    std::copy(t.order.begin(), t.order.end(), order);
    return SGM_OK;
}

int sgm_order_cfl(sgm_query* q, uint32_t* order)
{
    int rc = order_check(q, order, "Order_CFL");
    if (rc) return rc;
    sgm_ctx* c = q->g->ctx;
    uint32_t nlf[SGM_MAXQ];
    std::vector<uint32_t> ecnt;
    if ((rc = query_nlf_counts(q, nlf)) || (rc = tables_and_counts(q, ecnt))) return rc;
    EstBuf B;
    if ((rc = est_alloc(q, B, false))) return rc;
    const uint32_t n = q->q.n;
    // estimatePathEmbeddsingsNum (inc/order.h:197-239): row lengths of the last table, then
    // row sums of the previous tables over the vector below them, back to the path's root
    sgm::PathEstimator est = [&](const std::vector<uint32_t>& path, std::vector<uint64_t>& out) -> int {
        const size_t L = path.size();
        out.assign(L - 1, 0);
        sgm::launch_fill_u64(c->stream, B.tot, SGM_MAXQ, 0ull);
        unsigned long long *child = B.a, *par = B.b;
        {
            const uint32_t s = path[L - 2], e = path[L - 1];
            const sgm_table& t = q->tables[(size_t)s * n + e];
            sgm::launch_row_gather(c->stream, t.d_voff, t.d_pool, q->d_list + q->cand_off[s], q->counts[s], nullptr, child,
                                   0, B.tot + (L - 2));
        }
        for (size_t i = L - 2; i >= 1; i--) {
            const uint32_t s = path[i - 1], e = path[i];
            const sgm_table& t = q->tables[(size_t)s * n + e];
            sgm::launch_row_gather(c->stream, t.d_voff, t.d_pool, q->d_list + q->cand_off[s], q->counts[s], child, par, 0,
                                   B.tot + (i - 1));
            std::swap(child, par);
        }
        HIPCHK(c, hipGetLastError());
        unsigned long long tot[SGM_MAXQ];
        HIPCHK(c, hipMemcpyAsync(tot, B.tot, sizeof(unsigned long long) * (L - 1), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i + 1 < L; i++) out[i] = tot[i];
        return SGM_OK;
    };
    rc = sgm::order_cfl(q->q, nlf, label_stats(q->g), q->counts, est, order);
    if (rc == -1) return set_err(c, SGM_EINVAL, "Order_CFL: the query is not connected");
    return rc;
}

int sgm_order_gup(sgm_query* q, uint32_t* order)
{
    int rc = order_check(q, order, "Order_GuP");
    if (rc) return rc;
    sgm_ctx* c = q->g->ctx;
    const uint32_t n = q->q.n;
    uint32_t nlf[SGM_MAXQ];
    std::vector<uint32_t> ecnt;
    if ((rc = query_nlf_counts(q, nlf)) || (rc = tables_and_counts(q, ecnt))) return rc;
    // Filter_NLF into a fresh CandidateParam (inc/order.h:542-544): _NLF_Filter stops at the
    // first empty set and leaves the later counts 0
    for (uint32_t u = 0; u < n; u++)
        if (nlf[u] == 0) {
            for (uint32_t w = u + 1; w < n; w++) nlf[w] = 0;
            break;
        }
    EstBuf B;
    if ((rc = est_alloc(q, B, true))) return rc;
    // count_embed (inc/order.h:386-410): paths[u][iv] = product over u's unordered tree
    // children c of the row sum of table (u, c) over paths[c]; nembeds[c] = (sum of those row
    // sums) / |C(u)|.  Post-order recursion, one launch per tree edge, the totals read once.
    sgm::CountEmbed ce = [&](uint32_t root, const uint8_t* ordered, const sgm::BfsTree& t, double* nembeds) -> int {
        std::vector<std::pair<uint32_t, uint32_t>> edges;  // (u, child) in completion order
        std::function<void(uint32_t)> rec = [&](uint32_t u) {
            sgm::launch_fill_u64(c->stream, B.per + B.off[u], q->counts[u], 1ull);
            for (uint32_t p = q->q.off[u]; p < q->q.off[u + 1]; p++) {
                const uint32_t ch = q->q.adj[p];
                if (ordered[ch] || t.parent[ch] != u) continue;
                rec(ch);
                edges.emplace_back(u, ch);
                const sgm_table& tb = q->tables[(size_t)u * n + ch];
                sgm::launch_row_gather(c->stream, tb.d_voff, tb.d_pool, q->d_list + q->cand_off[u], q->counts[u],
                                       B.per + B.off[ch],
                                       B.per + B.off[u], 1, B.tot + ch);
            }
        };
        sgm::launch_fill_u64(c->stream, B.tot, SGM_MAXQ, 0ull);
        rec(root);
        HIPCHK(c, hipGetLastError());
        unsigned long long tot[SGM_MAXQ];
        HIPCHK(c, hipMemcpyAsync(tot, B.tot, sizeof tot, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (auto& [u, ch] : edges) nembeds[ch] = static_cast<double>(tot[ch]) / q->counts[u];
        return SGM_OK;
    };
    rc = sgm::order_gup(q->q, nlf, q->counts, ecnt.data(), ce, order);
    if (rc == -2) return set_err(c, SGM_EINVAL, "Order_GuP: every query vertex has degree <= 2 (the reference asserts)");
    if (rc == -1) return set_err(c, SGM_EINVAL, "Order_GuP: the query is not connected");
    return rc;
}

static int csr_of(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, sgm::QueryCSR& q)
{
    if (!qoff || n < 1 || n > SGM_MAXQ || qoff[0] != 0 || (!qadj && qoff[n])) return SGM_EINVAL;
    for (uint32_t u = 0; u < n; u++)
        if (qoff[u + 1] < qoff[u]) return SGM_EINVAL;
    for (uint32_t p = 0; p < qoff[n]; p++)
        if (qadj[p] >= n) return SGM_EINVAL;
    q.n = n;
    q.off.assign(qoff, qoff + n + 1);
    q.adj.assign(qadj, qadj + qoff[n]);
    return SGM_OK;
}

int sgm_query_kcore(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, int32_t* core)
{
    sgm::QueryCSR q;
    if (!core || csr_of(qoff, qadj, n, q)) return SGM_EINVAL;
    sgm::query_kcore(q, core);
    return SGM_OK;
}

int sgm_query_bfs(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint32_t start, uint32_t* order,
                  uint32_t* parent, uint32_t* level)
{
    sgm::QueryCSR q;
    if (!order || !parent || !level || start >= n || csr_of(qoff, qadj, n, q)) return SGM_EINVAL;
    sgm::BfsTree t;
    sgm::bfs_tree(q, start, t);
    std::copy(t.order.begin(), t.order.end(), order);
    for (uint32_t v = 0; v < n; v++) {
        parent[v] = t.parent[v];
        level[v] = t.level[v];
    }
    return (int)t.order.size();
}
