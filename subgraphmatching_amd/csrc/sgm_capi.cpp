// C ABI implementation: device residency, filter / table / enumeration drivers.
#include <array>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sgm_objects.h"

extern "C" {

const char* sgm_version(void) { return "sgm-mi355x 0.1 (gfx950)"; }

int sgm_device_count(int* count)
{
    if (!count) return SGM_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SGM_OK;
}

int sgm_ctx_create(int device, sgm_ctx** out)
{
    if (!out) return SGM_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SGM_ENODEV;
    if (device < 0 || device >= n) return SGM_EINVAL;
    sgm_ctx* c = new sgm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return SGM_EHIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wall_khz > 0)
        c->wall_ns_per_tick = 1e6 / wall_khz;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->ctl, hipStreamNonBlocking) != hipSuccess ||
        dalloc(&c->d_stop, 1) != hipSuccess || dalloc(&c->d_work, 1) != hipSuccess ||
        dalloc(&c->d_stats, 1) != hipSuccess || dalloc(&c->d_auxplan, 1) != hipSuccess ||
        dalloc(&c->d_work64, 1) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->fev[0]) != hipSuccess || hipEventCreate(&c->fev[1]) != hipSuccess ||
        hipEventCreate(&c->fev[2]) != hipSuccess || hipEventCreate(&c->fev[3]) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_stop), 64, hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess) {
        c->h_stop = nullptr;
        sgm_ctx_destroy(c);
        return SGM_EHIP;
    }
    void* dp = nullptr;
    if (hipHostGetDevicePointer(&dp, c->h_stop, 0) != hipSuccess) {
        sgm_ctx_destroy(c);
        return SGM_EHIP;
    }
    c->d_hstop = static_cast<const uint32_t*>(dp);
    __atomic_store_n(c->h_stop, 0u, __ATOMIC_SEQ_CST);
    if (const int rc = noalloc_setup(c)) {  // the no-allocation entry point's scratch array
        sgm_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return SGM_OK;
}

void sgm_ctx_destroy(sgm_ctx* c)
{
    if (!c) return;
    while (!c->batches.empty()) sgm_batch_destroy(c->batches.back());
    while (!c->models.empty()) sgm_costmodel_destroy(c->models.back());
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_fqd);
    dfree(c->d_fcd);
    dfree(c->d_fcnt);
    if (c->h_fcnt) (void)hipHostFree(c->h_fcnt);
    dfree(c->d_fdeg);
    dfree(c->d_rprog);
    if (c->h_rprog) (void)hipHostFree(c->h_rprog);
    dfree(c->d_tb);
    dfree(c->d_ttot);
    dfree(c->d_rrec);
    dfree(c->d_rootord);
    dfree(c->d_rootdesc);
    dfree(c->d_cmpoff);
    dfree(c->d_cmpdesc);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_stop) (void)hipHostFree(c->h_stop);
    dfree(c->d_lastbits);
    dfree(c->d_split);
    dfree(c->d_rec);
    dfree(c->d_stop);
    dfree(c->d_work);
    dfree(c->d_stats);
    dfree(c->d_auxplan);
    dfree(c->d_auxplans);
    dfree(c->d_afirst);
    dfree(c->d_dynplans);
    dfree(c->d_wdesc);
    dfree(c->d_astats);
    dfree(c->d_astop);
    dfree(c->d_anext);
    dfree(c->d_work64);
    dfree(c->d_plans);
    dfree(c->d_eq);
    dfree(c->d_qstats);
    dfree(c->d_qstop);
    dfree(c->d_stack);
    dfree(c->d_gqls);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto& e : c->fev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->ctl) (void)hipStreamDestroy(c->ctl);
    delete c;
}

const char* sgm_ctx_last_error(const sgm_ctx* c) { return c ? c->err.c_str() : "no context"; }

int sgm_ctx_filter_ms(const sgm_ctx* c, double* nlf_ms, double* gql_ms, double* cand_ms)
{
    if (!c) return SGM_EINVAL;
    if (nlf_ms) *nlf_ms = c->filter_ms[0];
    if (gql_ms) *gql_ms = c->filter_ms[1];
    if (cand_ms) *cand_ms = c->filter_ms[2];
    return SGM_OK;
}

int sgm_stop(sgm_ctx* c)
{
    if (!c) return SGM_EINVAL;
    // a plain store: the running kernels poll the word and raise their device stop word
    c->stop_requested = 1;
    __atomic_store_n(c->h_stop, 1u, __ATOMIC_SEQ_CST);
    return SGM_OK;
}

// ------------------------------------------------------------------------------ graph
int sgm_graph_upload(sgm_ctx* c, const uint32_t* off, const uint32_t* adj, uint32_t V, uint64_t adj_len,
                     sgm_graph** out)
{
    if (!c || !off || (!adj && adj_len) || !out) return set_err(c, SGM_EINVAL, "null argument");
    *out = nullptr;
    if (off[0] != 0 || off[V] != adj_len) return set_err(c, SGM_EINVAL, "offsets do not span adj");
    if (adj_len >= 0x80000000ull) return set_err(c, SGM_EINVAL, "adjacency longer than 2^31 entries");
    uint32_t maxdeg = 0;
    std::vector<uint32_t> hist;
    for (uint32_t v = 0; v < V; v++) {
        if (off[v + 1] < off[v]) return set_err(c, SGM_EINVAL, "offsets not monotone at %u", v);
        const uint32_t d = off[v + 1] - off[v];
        maxdeg = std::max(maxdeg, d);
        for (uint32_t p = off[v]; p < off[v + 1]; p++) {
            if (adj[p] >= V) return set_err(c, SGM_EINVAL, "neighbour out of range at vertex %u", v);
            if (adj[p] == v) return set_err(c, SGM_EINVAL, "self loop at vertex %u", v);
            if (p > off[v] && adj[p] <= adj[p - 1])
                return set_err(c, SGM_EINVAL, "adjacency of %u not strictly ascending", v);
        }
    }
    // vertices by degree, descending, ties by id (counting sort)
    hist.assign((size_t)maxdeg + 2, 0);
    for (uint32_t v = 0; v < V; v++) hist[off[v + 1] - off[v]]++;
    std::vector<uint32_t> start((size_t)maxdeg + 2, 0);
    uint32_t acc = 0;
    for (int64_t d = maxdeg; d >= 0; d--) { start[d] = acc; acc += hist[d]; }
    std::vector<uint32_t> bydeg(V ? V : 1);
    for (uint32_t v = 0; v < V; v++) bydeg[start[off[v + 1] - off[v]]++] = v;
    sgm_graph* g = new sgm_graph();
    g->ctx = c;
    g->V = V;
    g->adj_len = adj_len;
    g->max_deg = maxdeg;
    for (uint32_t d = 0; d <= 32; d++) {
        uint64_t s = 0;
        for (uint64_t x = d; x <= maxdeg; x++) s += hist[x];
        g->ge_cnt[d] = (uint32_t)s;
    }
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = dalloc(&g->d_off, (size_t)V + 1);
    if (e == hipSuccess) e = dalloc(&g->d_adj, adj_len);
    if (e == hipSuccess) e = dalloc(&g->d_bydeg, V);
    if (e == hipSuccess) e = dalloc(&g->d_gecnt, 33);
    if (e == hipSuccess) e = hipMemcpy(g->d_off, off, sizeof(uint32_t) * ((size_t)V + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && adj_len) e = hipMemcpy(g->d_adj, adj, sizeof(uint32_t) * adj_len, hipMemcpyHostToDevice);
    if (e == hipSuccess && V) e = hipMemcpy(g->d_bydeg, bydeg.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->d_gecnt, g->ge_cnt, sizeof(g->ge_cnt), hipMemcpyHostToDevice);
    // degree-rank space for GQL: rank of each vertex, CSR of ranks, neighbour-degree table
    if (e == hipSuccess) e = dalloc(&g->d_rank, V);
    if (e == hipSuccess) e = dalloc(&g->d_roff, (size_t)V + 1);
    if (e == hipSuccess) e = dalloc(&g->d_radj, adj_len);
    if (e == hipSuccess) e = dalloc(&g->d_dtab, (size_t)4 * V);
    if (e == hipSuccess && V) {
        std::vector<uint32_t> rank(V), roff((size_t)V + 1, 0);
        for (uint32_t x = 0; x < V; x++) {
            rank[bydeg[x]] = x;
            roff[x + 1] = roff[x] + (off[bydeg[x] + 1] - off[bydeg[x]]);
        }
        e = hipMemcpy(g->d_rank, rank.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(g->d_roff, roff.data(), sizeof(uint32_t) * ((size_t)V + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            sgm::launch_rank_adj(nullptr, g->d_off, g->d_adj, g->d_bydeg, g->d_rank, g->d_roff, V, g->d_radj);
            sgm::launch_nbr_deg(nullptr, g->d_roff, g->d_radj, V, g->d_dtab);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    // hub bitmaps: the highest-degree vertices first (by_deg order), as many as fit
    if (e == hipSuccess && V) {
        g->hub_words = (V + 31) / 32;
        const uint64_t per = (uint64_t)g->hub_words * sizeof(uint32_t);
        uint64_t h = 0;
        while (h < V && off[bydeg[h] + 1] - off[bydeg[h]] >= kHubMinDegree && (h + 1) * per <= kHubBudgetBytes) h++;
        g->n_hubs = (uint32_t)h;
        std::vector<uint32_t> slot(V, SGM_INVALID);
        for (uint32_t i = 0; i < g->n_hubs; i++) {
            slot[bydeg[i]] = i;
            g->hub_vertex.push_back(bydeg[i]);
            g->hub_degree.push_back(off[bydeg[i] + 1] - off[bydeg[i]]);
        }
        e = dalloc(&g->d_hub_slot, V);
        if (e == hipSuccess) e = hipMemcpy(g->d_hub_slot, slot.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice);
        if (e == hipSuccess && g->n_hubs) {
            e = dalloc(&g->d_hub_bits, (size_t)g->n_hubs * g->hub_words);
            if (e == hipSuccess) e = hipMemset(g->d_hub_bits, 0, (size_t)g->n_hubs * per);
            if (e == hipSuccess) {
                sgm::launch_hub_bits(nullptr, g->d_off, g->d_adj, g->d_bydeg, g->n_hubs, g->d_hub_bits, g->hub_words);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
    }
    if (e != hipSuccess) {
        sgm_graph_destroy(g);
        return set_err(c, e == hipErrorOutOfMemory ? SGM_ENOMEM : SGM_EHIP, "graph upload: %s", hipGetErrorString(e));
    }
    *out = g;
    return SGM_OK;
}

// Label tables of the labelled NLF (sgm_graph::d_lab, d_run_*), built on the host from the
// resident CSR: per vertex the labels of its neighbours, sorted, as (label, count) runs.
// labels == nullptr: every label 0 (one run per non-isolated vertex).
extern "C++" int build_label_tables(sgm_ctx* c, sgm_graph* g, const uint32_t* labels)
{
    const uint32_t V = g->V;
    std::vector<uint32_t> off((size_t)V + 1), adj, lab(V, 0u), roff((size_t)V + 1, 0u), rlab, rcnt;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(off.data(), g->d_off, sizeof(uint32_t) * ((size_t)V + 1), hipMemcpyDeviceToHost));
    if (labels) {
        lab.assign(labels, labels + V);
        adj.resize(g->adj_len);
        if (g->adj_len)
            HIPCHK(c, hipMemcpy(adj.data(), g->d_adj, sizeof(uint32_t) * g->adj_len, hipMemcpyDeviceToHost));
    }
    rlab.reserve(V);
    rcnt.reserve(V);
    std::vector<uint32_t> t;
    for (uint32_t v = 0; v < V; v++) {
        const uint32_t d = off[v + 1] - off[v];
        if (!labels) {
            if (d) { rlab.push_back(0); rcnt.push_back(d); }
        } else {
            t.resize(d);
            for (uint32_t k = 0; k < d; k++) t[k] = lab[adj[off[v] + k]];
            std::sort(t.begin(), t.end());
            for (uint32_t k = 0; k < d;) {
                uint32_t e = k;
                while (e < d && t[e] == t[k]) e++;
                rlab.push_back(t[k]);
                rcnt.push_back(e - k);
                k = e;
            }
        }
        roff[v + 1] = (uint32_t)rlab.size();
    }
    for (uint32_t** p : {&g->d_lab, &g->d_run_off, &g->d_run_lab, &g->d_run_cnt}) {
        dfree(*p);
        *p = nullptr;
    }
    g->has_lab = false;
    const size_t R = std::max<size_t>(rlab.size(), 1);
    HIPCHK(c, dalloc(&g->d_lab, std::max<uint32_t>(V, 1)));
    HIPCHK(c, dalloc(&g->d_run_off, (size_t)V + 1));
    HIPCHK(c, dalloc(&g->d_run_lab, R));
    HIPCHK(c, dalloc(&g->d_run_cnt, R));
    HIPCHK(c, hipMemcpy(g->d_lab, lab.data(), sizeof(uint32_t) * V, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(g->d_run_off, roff.data(), sizeof(uint32_t) * ((size_t)V + 1), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(g->d_run_lab, rlab.data(), sizeof(uint32_t) * rlab.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(g->d_run_cnt, rcnt.data(), sizeof(uint32_t) * rcnt.size(), hipMemcpyHostToDevice));
    g->has_lab = true;
    g->lab_gen++;  // captured filter graphs that read the old tables must be recaptured
    g->labelled = false;
    for (uint32_t v = 0; v < V && !g->labelled; v++) g->labelled = lab[v] != 0;
    g->lab_hist.clear();
    g->max_lab_freq = 0;
    if (g->labelled) {
        for (uint32_t v = 0; v < V; v++) g->lab_hist[lab[v]]++;
        for (auto& kv : g->lab_hist) g->max_lab_freq = std::max(g->max_lab_freq, kv.second);
    }
    return SGM_OK;
}

int sgm_graph_set_labels(sgm_graph* g, const uint32_t* labels, uint32_t num_vertices)
{
    if (!g) return SGM_EINVAL;
    sgm_ctx* c = g->ctx;
    if (num_vertices != g->V)
        return set_err(c, SGM_EINVAL, "expected %u labels, got %u", g->V, num_vertices);
    bool any = false;
    for (uint32_t v = 0; labels && v < num_vertices && !any; v++) any = labels[v] != 0;
    g->label_version++;  // queries filtered under the old labels are stale
    if (!any) {  // unlabeled: the tables go; a labelled query rebuilds them with every label 0
        for (uint32_t** p : {&g->d_lab, &g->d_run_off, &g->d_run_lab, &g->d_run_cnt}) {
            dfree(*p);
            *p = nullptr;
        }
        g->has_lab = false;
        g->lab_gen++;
        g->labelled = false;
        g->lab_hist.clear();
        g->max_lab_freq = 0;
        return SGM_OK;
    }
    return build_label_tables(c, g, labels);
}

void sgm_graph_destroy(sgm_graph* g)
{
    if (!g) return;
    dfree(g->d_lab);
    dfree(g->d_run_off);
    dfree(g->d_run_lab);
    dfree(g->d_run_cnt);
    dfree(g->d_off);
    dfree(g->d_adj);
    dfree(g->d_bydeg);
    dfree(g->d_gecnt);
    dfree(g->d_hub_slot);
    dfree(g->d_hub_bits);
    dfree(g->d_rank);
    dfree(g->d_roff);
    dfree(g->d_radj);
    dfree(g->d_dtab);
    for (auto& m : g->hub_maps) dfree(m.second);
    delete g;
}

uint32_t sgm_graph_max_degree(const sgm_graph* g) { return g ? g->max_deg : 0; }

// ------------------------------------------------------------------------------ query
// Tables are views into the query's arenas: dropping them frees nothing.
static void free_tables(sgm_query* q)
{
    for (auto& t : q->tables) t = sgm_table();
    q->weights_ok = false;
    q->guards_ok = false;
}

// Grows a u32 arena to at least `need` words (contents are not kept).
static hipError_t grow_arena(uint32_t** p, size_t& cap, size_t need)
{
    if (cap >= need) return hipSuccess;
    dfree(*p);
    *p = nullptr;
    cap = 0;
    const hipError_t e = dalloc(p, std::max<size_t>(need, 1));
    if (e == hipSuccess) cap = need;
    return e;
}

void sgm_query_destroy(sgm_query* q)
{
    if (!q) return;
    free_tables(q);
    dfree(q->d_tvoff);
    dfree(q->d_tpool);
    dfree(q->d_weight);
    dfree(q->d_guard);
    dfree(q->d_vmask);
    dfree(q->d_removed);
    dfree(q->d_bm);
    dfree(q->d_rank);
    dfree(q->d_cnt);
    dfree(q->d_candoff);
    dfree(q->d_list);
    dfree(q->d_lreq);
    dfree(q->d_qdesc);
    dfree(q->d_cdesc);
    delete q;
}

int sgm_query_create(sgm_graph* g, const uint32_t* qoff, const uint32_t* qadj, uint32_t n, sgm_query** out)
{
    if (!g || !qoff || !out) return SGM_EINVAL;
    sgm_ctx* c = g->ctx;
    *out = nullptr;
    if (n < 1 || n > SGM_MAXQ) return set_err(c, SGM_EINVAL, "query must have 1..32 vertices (got %u)", n);
    if (qoff[0] != 0) return set_err(c, SGM_EINVAL, "query offsets must start at 0");
    sgm_query* q = new sgm_query();
    q->g = g;
    q->q.n = n;
    q->q.off.assign(qoff, qoff + n + 1);
    q->q.adj.assign(qadj, qadj + qoff[n]);
    for (uint32_t u = 0; u < n; u++) {
        if (qoff[u + 1] < qoff[u]) { sgm_query_destroy(q); return set_err(c, SGM_EINVAL, "query offsets not monotone"); }
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
            const uint32_t w = qadj[p];
            if (w >= n || w == u || (p > qoff[u] && w <= qadj[p - 1])) {
                sgm_query_destroy(q);
                return set_err(c, SGM_EINVAL, "query adjacency of %u malformed", u);
            }
            bool back = false;
            for (uint32_t r = qoff[w]; r < qoff[w + 1]; r++) back |= qadj[r] == u;
            if (!back) { sgm_query_destroy(q); return set_err(c, SGM_EINVAL, "query graph not symmetric"); }
            q->nbr[u] |= 1u << w;
        }
    }
    const uint32_t V = g->V, W = (V + 63) / 64;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = dalloc(&q->d_vmask, V);
    if (e == hipSuccess) e = dalloc(&q->d_removed, SGM_GQL_FLAG_WORDS);
    if (e == hipSuccess) e = dalloc(&q->d_bm, (size_t)n * W);
    if (e == hipSuccess) e = dalloc(&q->d_rank, (size_t)n * W);
    if (e == hipSuccess) e = dalloc(&q->d_cnt, SGM_MAXQ);
    if (e == hipSuccess) e = dalloc(&q->d_candoff, SGM_MAXQ + 1);
    if (e == hipSuccess) e = dalloc(&q->d_qdesc, 1);
    if (e == hipSuccess) e = dalloc(&q->d_cdesc, 1);
    if (e == hipSuccess) {
        sgm_qdesc qd;
        std::memset(&qd, 0, sizeof qd);
        qd.n = n;
        for (uint32_t u = 0; u < n; u++) { qd.qdeg[u] = qoff[u + 1] - qoff[u]; qd.nbr[u] = q->nbr[u]; }
        for (uint32_t u = 0; u < n; u++) {  // neighbour degrees, descending (GQL degree-table test)
            std::vector<uint32_t> r;
            for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) r.push_back(qd.qdeg[qadj[p]]);
            std::sort(r.begin(), r.end(), [](uint32_t a, uint32_t b) { return a > b; });
            for (size_t j = 0; j < r.size(); j++) reinterpret_cast<uint8_t*>(qd.req[u])[j] = (uint8_t)r[j];
        }
        qd.vmask = q->d_vmask;
        qd.removed = q->d_removed;
        sgm_cdesc cd;
        std::memset(&cd, 0, sizeof cd);
        cd.n = n;
        cd.vmask = q->d_vmask;
        cd.bm = q->d_bm;
        cd.rank = q->d_rank;
        cd.cnt = q->d_cnt;
        cd.cand_off = q->d_candoff;
        cd.list = nullptr;
        q->hq = qd;
        q->hc = cd;
        e = hipMemcpy(q->d_qdesc, &qd, sizeof qd, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(q->d_cdesc, &cd, sizeof cd, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        sgm_query_destroy(q);
        return set_err(c, e == hipErrorOutOfMemory ? SGM_ENOMEM : SGM_EHIP, "query alloc: %s", hipGetErrorString(e));
    }
    q->tables.resize((size_t)n * n);
    *out = q;
    return SGM_OK;
}

int sgm_query_set_labels(sgm_query* q, const uint32_t* labels, uint32_t n)
{
    if (!q) return SGM_EINVAL;
    if (n != q->q.n) return set_err(q->g->ctx, SGM_EINVAL, "expected %u labels, got %u", q->q.n, n);
    bool any = false;
    for (uint32_t u = 0; labels && u < n && !any; u++) any = labels[u] != 0;
    if (any) q->q.lab.assign(labels, labels + n);
    else q->q.lab.clear();
    q->labelled = any;
    dfree(q->d_lreq);
    q->d_lreq = nullptr;
    q->filtered = 0;  // candidates and tables of the old labels are stale
    q->nlf_ok = false;
    return SGM_OK;
}

// The lreq block of a query (sgm_qdesc::lreq): per u its neighbours' labels with
// multiplicities, ascending by label.
extern "C++" std::vector<uint32_t> lreq_block(const sgm::QueryCSR& Q)
{
    const uint32_t n = Q.n;
    std::vector<uint32_t> first(n + 1, 0), plab, pcnt;
    for (uint32_t u = 0; u < n; u++) {
        std::vector<uint32_t> t;
        for (uint32_t p = Q.off[u]; p < Q.off[u + 1]; p++) t.push_back(Q.label(Q.adj[p]));
        std::sort(t.begin(), t.end());
        for (size_t k = 0; k < t.size();) {
            size_t e = k;
            while (e < t.size() && t[e] == t[k]) e++;
            plab.push_back(t[k]);
            pcnt.push_back((uint32_t)(e - k));
            k = e;
        }
        first[u + 1] = (uint32_t)plab.size();
    }
    std::vector<uint32_t> blk(first);
    blk.insert(blk.end(), plab.begin(), plab.end());
    blk.insert(blk.end(), pcnt.begin(), pcnt.end());
    return blk;
}

extern "C++" int ensure_lreq(sgm_ctx* c, sgm_query* q)
{
    if (q->d_lreq) return SGM_OK;
    const std::vector<uint32_t> blk = lreq_block(q->q);
    HIPCHK(c, dalloc(&q->d_lreq, blk.size()));
    HIPCHK(c, hipMemcpy(q->d_lreq, blk.data(), sizeof(uint32_t) * blk.size(), hipMemcpyHostToDevice));
    return SGM_OK;
}

// ------------------------------------------------------------------------------ filter
// The refinement steps of the BFS-tree filters (CFL / CECI / DAF) for a chunk whose vmask
// words hold the NLF sets: the NLF counts come back to the host (they pick each query's start
// vertex), the host writes one refinement program per query, and the steps run as one
// launch each for the whole chunk (k_refine).
static int tree_refine(sgm_ctx* c, sgm_graph* g, sgm_query* const* qs, uint32_t nq, int kind, const sgm_qdesc* d_qd)
{
    const uint32_t V = g->V;
    if (!c->d_fcnt || !c->d_fdeg || c->fcap < nq) return set_err(c, SGM_ESTATE, "filter: count buffers missing");
    sgm::launch_fill_u32(c->stream, c->d_fcnt, (size_t)nq * SGM_MAXQ, 0u);
    sgm::launch_fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->d_fdeg), (size_t)nq * SGM_MAXQ * 2, 0u);
    sgm::launch_cand_stats(c->stream, g->d_off, V, d_qd, nq, c->d_fcnt, c->d_fdeg);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_fcnt, c->d_fcnt, sizeof(uint32_t) * SGM_MAXQ * nq, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->rprog_cap < nq) {
        dfree(c->d_rprog);
        if (c->h_rprog) (void)hipHostFree(c->h_rprog);
        c->d_rprog = nullptr;
        c->h_rprog = nullptr;
        c->rprog_cap = 0;
        HIPCHK(c, dalloc(&c->d_rprog, (size_t)nq * SGM_RPROG_WORDS));
        HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_rprog), sizeof(uint32_t) * SGM_RPROG_WORDS * nq));
        c->rprog_cap = nq;
    }
    const sgm::LabelStats ls = label_stats(g);
    uint32_t steps = 0;
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        std::memcpy(q->nlf_cnt, c->h_fcnt + (size_t)i * SGM_MAXQ, sizeof(uint32_t) * q->q.n);
        q->nlf_ok = true;
        q->nlf_version = g->label_version;
        uint32_t* prog = c->h_rprog + (size_t)i * SGM_RPROG_WORDS;
        const int k = sgm::refine_program(kind, q->q, q->nlf_cnt, ls, prog);
        if (k < 0) return set_err(c, SGM_EINVAL, "filter: query %u is not connected", i);
        steps = std::max(steps, (uint32_t)k);
    }
    HIPCHK(c, hipMemcpyAsync(c->d_rprog, c->h_rprog, sizeof(uint32_t) * SGM_RPROG_WORDS * nq, hipMemcpyHostToDevice,
                             c->stream));
    if (kind == SGM_FILTER_CECI) sgm::launch_copy_vmask(c->stream, V, d_qd, nq);
    for (uint32_t k = 0; k < steps; k++) sgm::launch_refine(c->stream, g->d_off, g->d_adj, V, d_qd, nq, c->d_rprog, k);
    HIPCHK(c, hipGetLastError());
    return SGM_OK;
}

// One chunk of a filter batch: NLF (+ GQL) for nq queries in one set of launches, then the
// candidate bitmaps and counts.  Adds the device time of the three stages to ms[].
static int filter_chunk(sgm_ctx* c, sgm_graph* g, sgm_query* const* qs, uint32_t nq, int kind, double ms[3])
{
    const uint32_t V = g->V;
    // descriptors are built in pinned memory (a plain vector if none can be had)
    const size_t qd_bytes = stage_align(sizeof(sgm_qdesc) * nq);
    std::vector<sgm_qdesc> qd_v;
    std::vector<sgm_cdesc> cd_v;
    sgm_qdesc* qd = nullptr;
    sgm_cdesc* cd = nullptr;
    if (uint8_t* st = static_cast<uint8_t*>(host_stage(c, qd_bytes + sizeof(sgm_cdesc) * nq))) {
        qd = reinterpret_cast<sgm_qdesc*>(st);
        cd = reinterpret_cast<sgm_cdesc*>(st + qd_bytes);
    } else {
        qd_v.resize(nq);
        cd_v.resize(nq);
        qd = qd_v.data();
        cd = cd_v.data();
    }
    uint32_t max_n = 0;
    bool any_lab = false;
    for (uint32_t i = 0; i < nq; i++) {
        qd[i] = qs[i]->hq;
        cd[i] = qs[i]->hc;
        // labelled NLF / GQL when the graph or the query carries labels other than 0
        qd[i].lab_on = (g->labelled || qs[i]->labelled) ? 1u : 0u;
        if (qd[i].lab_on) {
            if (!g->has_lab) {
                const int rc = build_label_tables(c, g, nullptr);
                if (rc) return rc;
            }
            const int rc = ensure_lreq(c, qs[i]);
            if (rc) return rc;
            for (uint32_t u = 0; u < qs[i]->q.n; u++) qd[i].qlab[u] = qs[i]->q.label(u);
            qd[i].lreq = qs[i]->d_lreq;
            any_lab = true;
        }
        if (kind != SGM_FILTER_NLF) {  // GQL scratch; the BFS-tree filters use its vmr words
            sgm_gql_layout(qd[i], c->d_gqls + sgm_gql_scratch_words(V) * i, V);
        }
        max_n = std::max(max_n, qs[i]->q.n);
        free_tables(qs[i]);
        qs[i]->compacted = false;  // the list arena is kept: a free would synchronise the device
    }
    if (c->fcap < nq) {  // descriptor and count buffers, kept for later batches
        dfree(c->d_fqd);
        dfree(c->d_fcd);
        dfree(c->d_fcnt);
        dfree(c->d_fdeg);
        if (c->h_fcnt) (void)hipHostFree(c->h_fcnt);
        c->d_fqd = nullptr;
        c->d_fcd = nullptr;
        c->d_fcnt = nullptr;
        c->d_fdeg = nullptr;
        c->h_fcnt = nullptr;
        c->fcap = 0;
        HIPCHK(c, dalloc(&c->d_fqd, nq));
        HIPCHK(c, dalloc(&c->d_fcd, nq));
        HIPCHK(c, dalloc(&c->d_fcnt, (size_t)nq * SGM_MAXQ));
        HIPCHK(c, dalloc(&c->d_fdeg, (size_t)nq * SGM_MAXQ));
        HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_fcnt), sizeof(uint32_t) * SGM_MAXQ * nq));
        c->fcap = nq;
    }
    sgm_qdesc* d_qd = c->d_fqd;
    sgm_cdesc* d_cd = c->d_fcd;
    HIPCHK(c, hipMemcpyAsync(d_qd, qd, sizeof(sgm_qdesc) * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_cd, cd, sizeof(sgm_cdesc) * nq, hipMemcpyHostToDevice, c->stream));
    sgm::launch_clear_flags(c->stream, d_qd, nq);
    if (kind == SGM_FILTER_GQL && V)  // touch and dirty words start clear; the rest is written before it is read
        sgm::launch_fill_strided(c->stream, c->d_gqls + V, sgm_gql_clear_words(V), sgm_gql_scratch_words(V), nq, 0);
    HIPCHK(c, hipEventRecord(c->fev[0], c->stream));
    if (kind == SGM_FILTER_GQL && V) {
        sgm::launch_nlf_rank(c->stream, g->d_gecnt, V, d_qd, nq);
        if (any_lab)
            sgm::launch_nlf_rank_lab(c->stream, g->d_bydeg, g->d_off, g->d_lab, g->d_run_off, g->d_run_lab,
                                     g->d_run_cnt, V, d_qd, nq);
    } else {
        sgm::launch_nlf(c->stream, g->d_off, V, d_qd, nq);
        if (any_lab && V)
            sgm::launch_nlf_lab(c->stream, g->d_off, g->d_lab, g->d_run_off, g->d_run_lab, g->d_run_cnt, V, d_qd,
                                nq);
    }
    HIPCHK(c, hipEventRecord(c->fev[1], c->stream));
    if (kind == SGM_FILTER_GQL && V) {
        sgm::launch_gql_prepass(c->stream, g->d_dtab, V, g->d_gecnt, d_qd, nq);
        for (uint32_t ph = 0; ph < 2 * max_n; ph++)
            sgm::launch_gql_phase(c->stream, g->d_roff, g->d_radj, V, g->d_gecnt, d_qd, nq, ph);
        sgm::launch_unrank(c->stream, g->d_rank, V, d_qd, nq);
    }
    if (kind >= SGM_FILTER_CFL && V) {
        const int rc = tree_refine(c, g, qs, nq, kind, d_qd);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->fev[2], c->stream));
    sgm::launch_bitmaps(c->stream, d_cd, nq, V);
    sgm::launch_rank_scan(c->stream, d_cd, nq, V, max_n);
    HIPCHK(c, hipEventRecord(c->fev[3], c->stream));
    sgm::launch_gather_counts(c->stream, d_cd, nq, c->d_fcnt);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_fcnt, c->d_fcnt, sizeof(uint32_t) * SGM_MAXQ * nq, hipMemcpyDeviceToHost, c->stream));
    hipError_t se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) return set_err(c, SGM_EHIP, "filter: %s", hipGetErrorString(se));
    for (uint32_t i = 0; i < nq; i++)
        std::memcpy(qs[i]->counts, c->h_fcnt + (size_t)i * SGM_MAXQ, sizeof(uint32_t) * qs[i]->q.n);
    for (int k = 0; k < 3; k++) {
        float t = 0;
        HIPCHK(c, hipEventElapsedTime(&t, c->fev[k], c->fev[k + 1]));
        ms[k] += t;
    }
    return SGM_OK;
}

int sgm_filter_batch(sgm_query* const* qs, uint32_t nq, int kind, int32_t* status)
{
    if (!qs || nq == 0) return SGM_EINVAL;
    sgm_graph* g = qs[0]->g;
    sgm_ctx* c = g->ctx;
    if (kind < SGM_FILTER_NLF || kind > SGM_FILTER_DAF) return set_err(c, SGM_EINVAL, "unknown filter kind %d", kind);
    for (uint32_t i = 0; i < nq; i++)
        if (!qs[i] || qs[i]->g != g) return set_err(c, SGM_EINVAL, "queries of a batch must share one graph");
    HIPCHK(c, hipSetDevice(c->device));
    const uint64_t per_q = (uint64_t)sgm_gql_scratch_words(g->V) * sizeof(uint32_t);
    uint32_t chunk = nq;
    if (kind != SGM_FILTER_NLF && per_q) {
        chunk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nq, kGqlScratchBytes / per_q));
        const size_t words = sgm_gql_scratch_words(g->V) * chunk;
        if (c->gqls_words < words) {
            dfree(c->d_gqls);
            c->d_gqls = nullptr;
            c->gqls_words = 0;
            HIPCHK(c, dalloc(&c->d_gqls, words));
            c->gqls_words = words;
        }
    }
    double ms[3] = {0, 0, 0};
    for (uint32_t s0 = 0; s0 < nq; s0 += chunk) {
        const int rc = filter_chunk(c, g, qs + s0, std::min(chunk, nq - s0), kind, ms);
        if (rc) return rc;
    }
    for (int k = 0; k < 3; k++) c->filter_ms[k] = ms[k];
    int rc = SGM_OK;
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        q->filtered = kind;
        q->filter_label_version = g->label_version;
        q->empty = false;
        for (uint32_t u = 0; u < q->q.n; u++) q->empty |= q->counts[u] == 0;
        if (status) status[i] = q->empty ? SGM_EEMPTY : SGM_OK;
        if (q->empty) rc = SGM_EEMPTY;
    }
    return nq == 1 ? rc : SGM_OK;
}

int sgm_gql_work(const sgm_query* q, uint64_t* work)
{
    if (!q || !work) return SGM_EINVAL;
    uint32_t w[4];
    HIPCHK(q->g->ctx, hipMemcpy(w, q->d_removed + SGM_GQL_WORK, sizeof w, hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; i++) work[i] = w[i];
    return SGM_OK;
}

int sgm_filter(sgm_query* q, int kind)
{
    if (!q) return SGM_EINVAL;
    int32_t st = 0;
    int rc = sgm_filter_batch(&q, 1, kind, &st);
    return rc;
}

int sgm_cand_counts(const sgm_query* q, uint32_t* counts)
{
    if (!q || !counts) return SGM_EINVAL;
    if (!q->filtered) return set_err(q->g->ctx, SGM_ESTATE, "filter the query first");
    std::memcpy(counts, q->counts, sizeof(uint32_t) * q->q.n);
    return SGM_OK;
}

static int ensure_compacted(sgm_query* q)
{
    sgm_ctx* c = q->g->ctx;
    if (!q->filtered) return set_err(c, SGM_ESTATE, "filter the query first");
    if (q->compacted) return SGM_OK;
    const uint32_t n = q->q.n;
    q->cand_off[0] = 0;
    for (uint32_t u = 0; u < n; u++) q->cand_off[u + 1] = q->cand_off[u] + q->counts[u];
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, grow_arena(&q->d_list, q->list_cap, std::max<size_t>(q->cand_off[n], 1)));
    q->hc.list = q->d_list;
    HIPCHK(c, hipMemcpy(q->d_cdesc, &q->hc, sizeof q->hc, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(q->d_candoff, q->cand_off, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    sgm::launch_compact(c->stream, q->d_cdesc, 1, q->g->V);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    q->compacted = true;
    return SGM_OK;
}

int sgm_cand_download(sgm_query* q, uint32_t u, uint32_t* out, uint32_t out_cap)
{
    if (!q || !out) return SGM_EINVAL;
    sgm_ctx* c = q->g->ctx;
    if (u >= q->q.n) return set_err(c, SGM_EINVAL, "query vertex out of range");
    int rc = ensure_compacted(q);
    if (rc) return rc;
    if (out_cap < q->counts[u]) return set_err(c, SGM_EINVAL, "output buffer too small (%u < %u)", out_cap, q->counts[u]);
    if (q->counts[u])
        HIPCHK(c, hipMemcpy(out, q->d_list + q->cand_off[u], sizeof(uint32_t) * q->counts[u], hipMemcpyDeviceToHost));
    return SGM_OK;
}

int sgm_cand_exist_download(sgm_query* q, uint32_t* words, uint32_t num_words)
{
    if (!q || !words) return SGM_EINVAL;
    sgm_ctx* c = q->g->ctx;
    if (!q->filtered) return set_err(c, SGM_ESTATE, "filter the query first");
    if (num_words < q->g->V) return set_err(c, SGM_EINVAL, "need %u words", q->g->V);
    HIPCHK(c, hipMemcpy(words, q->d_vmask, sizeof(uint32_t) * q->g->V, hipMemcpyDeviceToHost));
    return SGM_OK;
}

// ------------------------------------------------------------------------------ tables
// buildTables (inc/table.h:7-137) for every directed edge of many queries at once.  The row
// of edge (a, b) at candidate v of a is N(v) n C(b) as positions in C(b): it depends on
// (b, v) only, so each query's rows are made once per (b, v) in a row pool (v in U_b, the
// union of C(a) over b's query neighbours a), with offsets by data vertex id, and every edge
// into b reads b's pool at its own candidates' rows (rows of vertices outside U_b are empty).
// Per pass of queries (whose rank records share one scratch of ceil(n/4)*4 V words per
// query): the pool row lengths (one launch), a scan per (query, b), the edges' entry totals
// and longest rows, one read-back of the totals, then the rank records and the pool rows --
// two host synchronisations per pass.  Queries whose candidate sets are empty are skipped
// (no search reads their tables).
static inline uint32_t rank_stride(uint32_t n) { return (n + 3) & ~3u; }

static int build_tables_pass(sgm_ctx* c, sgm_graph* g, sgm_query* const* qs, const uint32_t* idx, uint32_t nqp)
{
    const uint32_t V = g->V;
    const size_t per_b = (size_t)V + 1;
    std::vector<size_t> rbase(nqp);  // each query's block of rank records
    size_t rw = 0;
    for (uint32_t k = 0; k < nqp; k++) {
        rbase[k] = rw;
        rw += (size_t)rank_stride(qs[idx[k]]->q.n) * V;
    }
    HIPCHK(c, grow_arena(&c->d_rrec, c->rrec_cap, std::max<size_t>(rw, 1)));
    struct Edge {
        uint32_t k, a, b;
    };
    std::vector<Edge> edges;
    std::vector<uint32_t> pfirst(nqp + 1, 0);  // each query's first pool scan
    uint32_t kmax = 0, max_rows = 0;
    for (uint32_t k = 0; k < nqp; k++) {
        sgm_query* q = qs[idx[k]];
        pfirst[k + 1] = pfirst[k] + q->q.n;
        kmax = std::max(kmax, q->q.n);
        for (uint32_t a = 0; a < q->q.n; a++)
            for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++) {
                edges.push_back(Edge{k, a, q->q.adj[p]});
                max_rows = std::max(max_rows, q->counts[a]);
            }
    }
    const uint32_t M = (uint32_t)edges.size(), NP = pfirst[nqp];
    if (M == 0) return SGM_OK;
    // device descriptors [pools nqp][edges M][pool scans NP]; pinned staging holds the same
    // plus the totals and maxima read back (pool scans, then edges)
    const size_t pool_b = stage_align(sizeof(sgm_pooldesc) * nqp), edge_b = stage_align(sizeof(sgm_edgedesc) * M);
    const size_t scan_b = stage_align(sizeof(sgm_scandesc) * NP), dev_b = pool_b + edge_b + scan_b;
    const uint32_t NT = NP + M;
    if (c->tb_bytes < dev_b) {
        dfree(c->d_tb);
        c->d_tb = nullptr;
        c->tb_bytes = 0;
        HIPCHK(c, dalloc(&c->d_tb, dev_b));
        c->tb_bytes = dev_b;
    }
    if (c->ttot_cap < 2 * NT) {
        dfree(c->d_ttot);
        c->d_ttot = nullptr;
        c->ttot_cap = 0;
        HIPCHK(c, dalloc(&c->d_ttot, 2 * (size_t)NT));
        c->ttot_cap = 2 * NT;
    }
    std::vector<uint8_t> fallback;
    uint8_t* st = static_cast<uint8_t*>(host_stage(c, dev_b + sizeof(uint32_t) * 2 * NT));
    if (!st) {
        fallback.resize(dev_b + sizeof(uint32_t) * 2 * NT);
        st = fallback.data();
    }
    sgm_pooldesc* pd = reinterpret_cast<sgm_pooldesc*>(st);
    sgm_edgedesc* ed = reinterpret_cast<sgm_edgedesc*>(st + pool_b);
    sgm_scandesc* sd = reinterpret_cast<sgm_scandesc*>(st + pool_b + edge_b);
    uint32_t* h = reinterpret_cast<uint32_t*>(st + dev_b);  // totals[NT], maxima[NT]
    uint8_t* dev = reinterpret_cast<uint8_t*>(c->d_tb);
    for (uint32_t k = 0; k < nqp; k++) {
        sgm_query* q = qs[idx[k]];
        const uint32_t n = q->q.n;
        sgm_pooldesc& P = pd[k];
        std::memset(&P, 0, sizeof P);
        P.vmask = q->d_vmask;
        P.poff = q->d_tvoff;
        P.rrec = c->d_rrec + rbase[k];
        P.bm = q->d_bm;
        P.rank = q->d_rank;
        P.n = n;
        P.rstride = rank_stride(n);
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t p = q->q.off[u]; p < q->q.off[u + 1]; p++) P.nbr[u] |= 1u << q->q.adj[p];
        for (uint32_t b = 0; b < n; b++) {
            const uint32_t j = pfirst[k] + b;
            sd[j] = sgm_scandesc{P.poff + b * per_b, P.poff + b * per_b, c->d_ttot + j, c->d_ttot + NT + j, V, 0};
        }
    }
    for (uint32_t e = 0; e < M; e++) {
        const Edge& E = edges[e];
        sgm_query* q = qs[idx[E.k]];
        ed[e] = sgm_edgedesc{q->d_list + q->cand_off[E.a], pd[E.k].poff + E.b * per_b, q->counts[E.a], NP + e};
    }
    HIPCHK(c, hipMemcpyAsync(dev, st, dev_b, hipMemcpyHostToDevice, c->stream));
    const sgm_pooldesc* dpd = reinterpret_cast<const sgm_pooldesc*>(dev);
    const sgm_edgedesc* ded = reinterpret_cast<const sgm_edgedesc*>(dev + pool_b);
    const sgm_scandesc* dsd = reinterpret_cast<const sgm_scandesc*>(dev + pool_b + edge_b);
    sgm::launch_fill_u32(c->stream, c->d_ttot, 2 * (size_t)NT, 0);
    sgm::launch_pool_count(c->stream, g->d_off, g->d_adj, V, dpd, nqp, kmax);
    sgm::launch_scan_segments(c->stream, dsd, NP);
    sgm::launch_edge_stats(c->stream, ded, M, max_rows, c->d_ttot, c->d_ttot + NT);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(h, c->d_ttot, sizeof(uint32_t) * 2 * NT, hipMemcpyDeviceToHost, c->stream));
    hipError_t se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) return set_err(c, SGM_EHIP, "table count: %s", hipGetErrorString(se));
    // the pools into each query's arena
    for (uint32_t k = 0; k < nqp; k++) {
        sgm_query* q = qs[idx[k]];
        const uint32_t n = q->q.n;
        sgm_pooldesc& P = pd[k];
        size_t pool = 0;
        for (uint32_t b = 0; b < n; b++) {
            P.pbase[b] = pool;
            pool += h[pfirst[k] + b];
        }
        if (grow_arena(&q->d_tpool, q->tpool_cap, std::max<size_t>(pool, 1)) != hipSuccess)
            return set_err(c, SGM_ENOMEM, "table row pool of %llu entries", (unsigned long long)pool);
        P.pool = q->d_tpool;
    }
    for (uint32_t e = 0; e < M; e++) {
        const Edge& E = edges[e];
        sgm_query* q = qs[idx[E.k]];
        sgm_table& t = q->tables[(size_t)E.a * q->q.n + E.b];
        t.e_cnt = h[NP + e];
        t.max_deg = h[NT + NP + e];
        t.d_voff = pd[E.k].poff + E.b * per_b;
        t.d_pool = q->d_tpool + pd[E.k].pbase[E.b];
        t.built = true;
    }
    HIPCHK(c, hipMemcpyAsync(dev, st, pool_b, hipMemcpyHostToDevice, c->stream));
    sgm::launch_pool_fill(c->stream, g->d_off, g->d_adj, V, dpd, nqp);
    HIPCHK(c, hipGetLastError());
    se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) return set_err(c, SGM_EHIP, "tables: %s", hipGetErrorString(se));
    return SGM_OK;
}

extern "C++" int build_tables_many(sgm_query* const* qs, uint32_t nq)
{
    if (!nq) return SGM_OK;
    sgm_graph* g = qs[0]->g;
    sgm_ctx* c = g->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<uint32_t> todo;
    // candidate lists of the queries not compacted yet, all in one launch (the descriptors
    // and prefix arrays stay alive until the passes below synchronise)
    std::vector<sgm_cdesc> cds;
    std::vector<uint32_t> coffs;
    std::vector<size_t> cat;
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        if (q->g != g) return set_err(c, SGM_EINVAL, "tables: queries must share one graph");
        int rc = check_filtered(q, "tables");
        if (rc) return rc;
        free_tables(q);
        if (q->empty) continue;
        if (!q->compacted) {
            const uint32_t n = q->q.n;
            q->cand_off[0] = 0;
            for (uint32_t u = 0; u < n; u++) q->cand_off[u + 1] = q->cand_off[u] + q->counts[u];
            HIPCHK(c, grow_arena(&q->d_list, q->list_cap, std::max<size_t>(q->cand_off[n], 1)));
            q->hc.list = q->d_list;
            cds.push_back(q->hc);
            cat.push_back(coffs.size());
            coffs.insert(coffs.end(), q->cand_off, q->cand_off + n + 1);
            q->compacted = true;
        }
        HIPCHK(c, grow_arena(&q->d_tvoff, q->tvoff_cap, (size_t)q->q.n * ((size_t)g->V + 1)));
        todo.push_back(i);
    }
    if (!cds.empty()) {
        HIPCHK(c, grow_arena(&c->d_cmpoff, c->cmpoff_cap, coffs.size()));
        if (const int grc = grow(c, &c->d_cmpdesc, c->cmpdesc_cap, (uint32_t)cds.size())) return grc;
        for (size_t k = 0; k < cds.size(); k++) cds[k].cand_off = c->d_cmpoff + cat[k];
        HIPCHK(c, hipMemcpyAsync(c->d_cmpoff, coffs.data(), sizeof(uint32_t) * coffs.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_cmpdesc, cds.data(), sizeof(sgm_cdesc) * cds.size(), hipMemcpyHostToDevice,
                                 c->stream));
        sgm::launch_compact(c->stream, c->d_cmpdesc, (uint32_t)cds.size(), g->V);
        HIPCHK(c, hipGetLastError());
    }
    // passes of queries whose rank records fit a scratch of at most 2^30 words (a single
    // larger query takes a pass of its own)
    auto words_of = [&](const sgm_query* q) { return (size_t)rank_stride(q->q.n) * g->V; };
    const size_t limit = (size_t)1 << 30;
    int rc = SGM_OK;
    for (size_t k0 = 0; k0 < todo.size() && rc == SGM_OK;) {
        size_t k1 = k0, words = 0;
        while (k1 < todo.size() && (k1 == k0 || words + words_of(qs[todo[k1]]) <= limit)) words += words_of(qs[todo[k1++]]);
        rc = build_tables_pass(c, g, qs, todo.data() + k0, (uint32_t)(k1 - k0));
        k0 = k1;
    }
    if (rc == SGM_OK) {  // (the passes end synchronised; this covers none, and the compaction)
        const hipError_t se = hipStreamSynchronize(c->stream);
        if (se != hipSuccess) rc = set_err(c, SGM_EHIP, "tables: %s", hipGetErrorString(se));
    }
    if (rc != SGM_OK)
        for (uint32_t i = 0; i < nq; i++) free_tables(qs[i]);
    return rc;
}

extern "C++" int root_orders(sgm_ctx* c, sgm_query* const* qs, const uint32_t* us, uint32_t nq,
                            std::vector<const uint32_t*>& root_order)
{
    root_order.assign(nq, nullptr);
    if (!nq) return SGM_OK;
    sgm_graph* g = qs[0]->g;
    const uint32_t W = (g->V + 63) / 64;
    size_t total = 0;
    std::vector<sgm_rootdesc> ds;
    std::vector<size_t> at(nq, 0);
    for (uint32_t i = 0; i < nq; i++) {
        const sgm_query* q = qs[i];
        if (q->empty || !q->compacted) continue;
        at[i] = total;
        total += q->counts[us[i]];
    }
    HIPCHK(c, grow_arena(&c->d_rootord, c->rootord_cap, std::max<size_t>(total, 1)));
    for (uint32_t i = 0; i < nq; i++) {
        const sgm_query* q = qs[i];
        if (q->empty || !q->compacted) continue;
        root_order[i] = c->d_rootord + at[i];
        ds.push_back(sgm_rootdesc{q->d_vmask, q->d_bm + (size_t)us[i] * W, q->d_rank + (size_t)us[i] * W,
                                  c->d_rootord + at[i], us[i], 0});
    }
    if (ds.empty()) return SGM_OK;
    if (const int rc = grow(c, &c->d_rootdesc, c->rootdesc_cap, (uint32_t)ds.size())) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_rootdesc, ds.data(), sizeof(sgm_rootdesc) * ds.size(), hipMemcpyHostToDevice,
                             c->stream));
    sgm::launch_root_order(c->stream, g->d_bydeg, g->V, c->d_rootdesc, (uint32_t)ds.size());
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));  // ds is read by the copy above
    return SGM_OK;
}

// Edge (a, b)'s table in the reference's layout (offsets per candidate row of a, then the
// positions in C(b)), gathered on the host from b's pool at a's candidates' rows.
extern "C++" int table_fetch(sgm_query* q, uint32_t a, uint32_t b, std::vector<uint32_t>& off, std::vector<uint32_t>& idx)
{
    sgm_ctx* c = q->g->ctx;
    const uint32_t V = q->g->V;
    const sgm_table& t = q->tables[(size_t)a * q->q.n + b];
    if (!t.built) return set_err(c, SGM_ESTATE, "no table for (%u,%u)", a, b);
    std::vector<uint32_t> ca(q->counts[a]), voff((size_t)V + 1);
    if (!ca.empty())
        HIPCHK(c, hipMemcpy(ca.data(), q->d_list + q->cand_off[a], sizeof(uint32_t) * ca.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(voff.data(), t.d_voff, sizeof(uint32_t) * voff.size(), hipMemcpyDeviceToHost));
    std::vector<uint32_t> pool(voff[V]);
    if (!pool.empty())
        HIPCHK(c, hipMemcpy(pool.data(), t.d_pool, sizeof(uint32_t) * pool.size(), hipMemcpyDeviceToHost));
    off.assign(ca.size() + 1, 0);
    idx.clear();
    idx.reserve(t.e_cnt);
    for (size_t r = 0; r < ca.size(); r++) {
        const uint32_t v = ca[r];
        idx.insert(idx.end(), pool.begin() + voff[v], pool.begin() + voff[v + 1]);
        off[r + 1] = (uint32_t)idx.size();
    }
    return SGM_OK;
}

int sgm_build_tables(sgm_query* q)
{
    if (!q) return SGM_EINVAL;
    return build_tables_many(&q, 1);
}

int sgm_table_info(sgm_query* q, uint32_t a, uint32_t b, uint64_t* e_cnt, uint32_t* max_degree)
{
    if (!q || a >= q->q.n || b >= q->q.n) return SGM_EINVAL;
    const sgm_table& t = q->tables[(size_t)a * q->q.n + b];
    if (!t.built) return set_err(q->g->ctx, SGM_ESTATE, "no table for (%u,%u)", a, b);
    if (e_cnt) *e_cnt = t.e_cnt;
    if (max_degree) *max_degree = t.max_deg;
    return SGM_OK;
}

int sgm_table_download(sgm_query* q, uint32_t a, uint32_t b, uint32_t* offsets, uint32_t* idx, uint32_t* vtx)
{
    if (!q || a >= q->q.n || b >= q->q.n) return SGM_EINVAL;
    sgm_ctx* c = q->g->ctx;
    std::vector<uint32_t> off, pos;
    if (const int rc = table_fetch(q, a, b, off, pos)) return rc;
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    if (idx) std::copy(pos.begin(), pos.end(), idx);
    if (vtx && !pos.empty()) {
        // _edge_v is the candidate at each position, C(b)[_edge[i]]: read from C(b) here
        // rather than kept on the device beside the positions
        std::vector<uint32_t> cb(q->counts[b]);
        HIPCHK(c, hipMemcpy(cb.data(), q->d_list + q->cand_off[b], sizeof(uint32_t) * cb.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < pos.size(); i++) vtx[i] = cb[pos[i]];
    }
    return SGM_OK;
}

// ------------------------------------------------------------------------------ orders
static int csr_from(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, sgm::QueryCSR& q)
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

int sgm_order_ri(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint32_t* order)
{
    sgm::QueryCSR q;
    if (!order || csr_from(qoff, qadj, n, q)) return SGM_EINVAL;
    sgm::order_ri(q, order);
    return SGM_OK;
}

int sgm_order_sample(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint64_t seed, uint32_t* order)
{
    sgm::QueryCSR q;
    if (!order || csr_from(qoff, qadj, n, q)) return SGM_EINVAL;
    // splitmix64 stream; uniform choice among k by the high half of a 64x32 product
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto pick = [&](uint32_t k) { return (uint32_t)(((next() >> 32) * (uint64_t)k) >> 32); };
    uint32_t visited = 0, frontier = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t u;
        if (i == 0 || !frontier) {  // start (or a new component of a disconnected query)
            uint32_t cand[SGM_MAXQ], k = 0;
            for (uint32_t x = 0; x < n; x++)
                if (!((visited >> x) & 1u)) cand[k++] = x;
            u = cand[pick(k)];
        } else {
            uint32_t cand[SGM_MAXQ], k = 0;
            for (uint32_t x = 0; x < n; x++)
                if ((frontier >> x) & 1u) cand[k++] = x;
            u = cand[pick(k)];
        }
        order[i] = u;
        visited |= 1u << u;
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) frontier |= 1u << q.adj[p];
        frontier &= ~visited;
    }
    return SGM_OK;
}

int sgm_order_gql(const sgm_query* q, uint32_t* order)
{
    if (!q || !order) return SGM_EINVAL;
    if (!q->filtered) return set_err(q->g->ctx, SGM_ESTATE, "filter the query first");
    sgm::order_gql(q->q, q->counts, q->g->V, order);
    return SGM_OK;
}

int sgm_plan_export(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, const uint32_t* order, uint32_t* n_ops,
                    uint32_t* depth, uint32_t* parent, uint32_t* vdepth, uint32_t* emask, uint32_t* iter_set)
{
    return sgm_plan_export_labeled(qoff, qadj, nullptr, n, order, n_ops, depth, parent, vdepth, emask, iter_set);
}

int sgm_plan_export_labeled(const uint32_t* qoff, const uint32_t* qadj, const uint32_t* labels, uint32_t n,
                            const uint32_t* order, uint32_t* n_ops, uint32_t* depth, uint32_t* parent,
                            uint32_t* vdepth, uint32_t* emask, uint32_t* iter_set)
{
    sgm::QueryCSR q;
    if (!order || !n_ops || csr_from(qoff, qadj, n, q)) return SGM_EINVAL;
    if (labels) q.lab.assign(labels, labels + n);
    sgm::Plan pl;
    const int rc = sgm::build_plan(q, order, pl);
    if (rc == -2) return SGM_EORDER;
    if (rc) return SGM_EINVAL;
    *n_ops = (uint32_t)pl.ops.size();
    for (size_t i = 0; i < pl.ops.size(); i++) {
        if (depth) depth[i] = pl.ops[i].depth;
        if (parent) parent[i] = pl.ops[i].parent;
        if (vdepth) vdepth[i] = pl.ops[i].vdepth;
        if (emask) emask[i] = pl.ops[i].emask;
    }
    if (iter_set)
        for (uint32_t d = 0; d < n; d++) iter_set[d] = pl.iter_set[d];
    return SGM_OK;
}

// ------------------------------------------------------------------------------ enumerate
}  // extern "C"

int ensure_stack(sgm_ctx* c, size_t bytes)
{
    if (c->stack_bytes >= bytes) return SGM_OK;
    dfree(c->d_stack);
    c->d_stack = nullptr;
    c->stack_bytes = 0;
    HIPCHK(c, hipMalloc((void**)&c->d_stack, bytes));
    c->stack_bytes = bytes;
    return SGM_OK;
}

int hub_slot_map(sgm_graph* g, const sgm_enum_opts& o, const uint32_t** out)
{
    *out = nullptr;
    if ((o.flags & SGM_ENUM_F_NO_BITMAPS) || !g->n_hubs) return SGM_OK;
    if (o.hub_min_degree <= kHubMinDegree) { *out = g->d_hub_slot; return SGM_OK; }
    for (auto& m : g->hub_maps)
        if (m.first == o.hub_min_degree) { *out = m.second; return SGM_OK; }
    // the hubs are in degree order, so the threshold keeps a prefix of the slots
    std::vector<uint32_t> slot(g->V, SGM_INVALID);
    for (uint32_t i = 0; i < g->n_hubs && g->hub_degree[i] >= o.hub_min_degree; i++) slot[g->hub_vertex[i]] = i;
    uint32_t* d = nullptr;
    HIPCHK(g->ctx, dalloc(&d, g->V));
    HIPCHK(g->ctx, hipMemcpy(d, slot.data(), sizeof(uint32_t) * g->V, hipMemcpyHostToDevice));
    g->hub_maps.emplace_back(o.hub_min_degree, d);
    *out = d;
    return SGM_OK;
}

sgm_enum_opts normalize_opts(const sgm_enum_opts* opts)
{
    sgm_enum_opts o;
    std::memset(&o, 0, sizeof o);
    o.cap = UINT64_MAX;
    if (opts) o = *opts;
    if (o.cap == 0) o.cap = 1;  // the reference stops at the first embedding when ans >= 0
    return o;
}

uint32_t default_blocks(const sgm_ctx* c, const sgm_enum_opts& o, uint32_t nq)
{
    // A single capped query: only its first embeddings matter and extra waves add overshoot.
    // A batch (or an uncapped search) wants the wave slots busy.
    const uint32_t per_cu = o.cap == UINT64_MAX ? 8u : (nq > 1 ? 4u : 2u);
    uint32_t b = o.blocks ? o.blocks : (uint32_t)c->num_cus * per_cu;
    return std::max<uint32_t>(1, std::min<uint32_t>(b, 1u << 16));
}

unsigned long long budget_ticks(const sgm_ctx* c, const sgm_enum_opts& o)
{
    if (!o.time_limit_ms) return 0;
    const double t = (double)o.time_limit_ms * 1e6 / c->wall_ns_per_tick;
    return t >= 1.8e19 ? ~0ull : std::max<unsigned long long>(1, (unsigned long long)t);
}

void stats_out(const sgm_stats_dev& st, uint64_t cap, sgm_stats* out, double ns_per_tick)
{
    out->busy_ns = (uint64_t)((double)st.busy_ticks * ns_per_tick);
    out->count = st.count < cap ? st.count : cap;
    out->capped = st.count >= cap;
    // a query that reached its cap has its result, whatever its clock says
    out->timed_out = (st.timed_out && !out->capped) ? 1 : 0;
    out->intersections = st.intersections;
    for (uint32_t d = 0; d < SGM_MAXQ; d++) out->nodes[d] = st.nodes[d];
    out->fs_prunes = st.fs_prunes;
    out->elems_read = st.elems_read;
    out->elems_out = st.elems_out;
    out->elems_scanned = st.elems_scanned;
    out->probe_words = st.probe_words;
}

extern "C" {

// explore_Intersection_All<fs,false,false,false,false> for many queries in one launch over
// their device tables (built here when missing); orders[i] is query i's order.
static int enumerate_aux_batch(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq,
                               const sgm_enum_opts& o, sgm_stats* out)
{
    sgm_graph* g = qs[0]->g;
    sgm_ctx* c = g->ctx;
    std::vector<sgm_auxplan_dev> aps(nq);
    std::vector<sgm_query*> missing;
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        const uint32_t n = q->q.n;
        bool have = true;
        for (uint32_t a = 0; a < n && have; a++)
            for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++)
                have &= q->tables[(size_t)a * n + q->q.adj[p]].built;
        if (!have && !q->empty) missing.push_back(q);
    }
    if (!missing.empty()) {  // every missing table of the launch in one build
        const int trc = build_tables_many(missing.data(), (uint32_t)missing.size());
        if (trc) return trc;
    }
    const bool gup_cd = o.method == SGM_ENUM_AUX_TABLES_GUPCD || o.method == SGM_ENUM_AUX_TABLES_FS_GUPCD;
    if (gup_cd)  // GuP's reservation guards for each query's order (kept while order and tables stay)
        for (uint32_t i = 0; i < nq; i++) {
            sgm_query* q = qs[i];
            if (q->empty || (q->guards_ok && std::equal(orders[i], orders[i] + q->q.n, q->guard_order))) continue;
            const int grc = gup_guards_build(q, orders[i]);
            if (grc) return grc;
        }
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        const uint32_t* order = orders[i];
        const uint32_t n = q->q.n;
        std::vector<std::vector<uint32_t>> prev;
        sgm::previous_neighbours(q->q, order, prev);
        std::vector<uint32_t> pos(n);
        for (uint32_t d = 0; d < n; d++) pos[order[d]] = d;
        sgm_auxplan_dev& ap = aps[i];
        std::memset(&ap, 0, sizeof ap);
        ap.n = n;
        uint32_t e = 0;
        for (uint32_t d = 0; d < n; d++) {
            const uint32_t u = order[d];
            ap.order[d] = u;
            ap.prev_first[d] = e;
            ap.cand[d] = q->d_list ? q->d_list + q->cand_off[u] : nullptr;
            ap.guard[d] = gup_cd && q->d_guard ? q->d_guard + 4 * (size_t)q->guard_off[d] : nullptr;
            for (uint32_t p : prev[u]) {
                const sgm_table& t = q->tables[(size_t)p * n + u];
                ap.prev_depth[e] = pos[p];
                ap.tab_off[e] = t.d_voff;
                ap.tab_idx[e] = t.d_pool;
                e++;
            }
        }
        ap.prev_first[n] = e;
        // DAF ancestors over depths (inc/prune.h:100-104): anc[d] = {d} and the ancestors of
        // every backward neighbour
        for (uint32_t d = 0; d < n; d++) {
            uint64_t a = 1ull << d;
            for (uint32_t p : prev[order[d]]) a |= ap.anc[pos[p]];
            ap.anc[d] = a;
        }
    }
    const int use_fs = o.method == SGM_ENUM_AUX_TABLES_FS || o.method == SGM_ENUM_AUX_TABLES_FS_GUPCD ? 1 : 0;
    const uint32_t* hub_slot = nullptr;  // adjacency bitmaps of the hubs, for the row probes
    if (const int hrc = hub_slot_map(g, o, &hub_slot)) return hrc;
    {  // each query's roots in degree order
        std::vector<uint32_t> us(nq);
        for (uint32_t i = 0; i < nq; i++) us[i] = orders[i][0];
        std::vector<const uint32_t*> ro;
        if (const int rrc = root_orders(c, qs, us.data(), nq, ro)) return rrc;
        for (uint32_t i = 0; i < nq; i++) aps[i].root_order = ro[i];
    }
    uint32_t rlo, rhi;
    root_range(o, &rlo, &rhi);
    std::vector<uint32_t> live;
    for (uint32_t i = 0; i < nq; i++)
        if (!qs[i]->empty) live.push_back(i);  // the reference's driver reports 0 without enumerating
    const uint32_t blocks = default_blocks(c, o, (uint32_t)std::max<size_t>(live.size(), 1));
    const uint32_t waves = blocks * 4;
    uint64_t words = o.stack_words ? o.stack_words : (uint64_t)8 * g->max_deg + 1024;
    for (uint32_t attempt = 0; !live.empty(); attempt++) {
        if (attempt > 3) return set_err(c, SGM_EDEVICE, "set-op stack overflow after %u retries", attempt - 1);
        if (words > 0x7FFFFFFFull) return set_err(c, SGM_EDEVICE, "set-op stack would exceed 2^31 words");
        int rc = ensure_stack(c, (size_t)waves * words * sizeof(uint32_t));
        if (rc) return rc;
        const uint32_t m = (uint32_t)live.size();
        if ((rc = grow(c, &c->d_auxplans, c->aux_cap, m))) return rc;
        if ((rc = grow(c, &c->d_afirst, c->afirst_cap, m + 1))) return rc;
        if ((rc = grow(c, &c->d_astats, c->astats_cap, m))) return rc;
        if ((rc = grow(c, &c->d_astop, c->astop_cap, m))) return rc;
        if ((rc = grow(c, &c->d_anext, c->anext_cap, m))) return rc;
        std::vector<sgm_auxplan_dev> sub(m);
        std::vector<uint32_t> first(m + 1, 0);
        for (uint32_t k = 0; k < m; k++) {
            const uint32_t i = live[k];
            sub[k] = aps[i];
            first[k + 1] = first[k] + qs[i]->counts[orders[i][0]];
        }
        HIPCHK(c, hipMemcpyAsync(c->d_auxplans, sub.data(), sizeof(sgm_auxplan_dev) * m, hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_afirst, first.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_astats, 0, sizeof(sgm_stats_dev) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_astop, 0, sizeof(uint32_t) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_anext, 0, sizeof(uint32_t) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_stop, 0, sizeof(uint32_t), c->stream));
        clear_host_stop(c);
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        sgm::launch_enum_aux(c->stream, blocks, c->d_auxplans, c->d_afirst, m, c->d_stack, (uint32_t)words, c->d_anext,
                             c->d_astats, c->d_astop, c->d_stop, c->d_hstop, (unsigned long long)o.cap,
                             budget_ticks(c, o), rlo, rhi, use_fs, hub_slot, g->d_hub_bits, g->hub_words);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        std::vector<sgm_stats_dev> got(m);
        HIPCHK(c, hipMemcpyAsync(got.data(), c->d_astats, sizeof(sgm_stats_dev) * m, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        std::vector<uint32_t> again;
        for (uint32_t k = 0; k < m; k++) {
            const uint32_t i = live[k];
            out[i].kernel_ms += ms;
            if (got[k].overflow) {
                out[i].retries++;
                again.push_back(i);
                continue;
            }
            stats_out(got[k], o.cap, &out[i], c->wall_ns_per_tick);
            if (c->stop_requested && !out[i].capped) out[i].stopped = 1;
        }
        live.swap(again);
        words *= 4;
    }
    if (c->stop_requested) {
        for (uint32_t i = 0; i < nq; i++)
            if (out[i].stopped) return set_err(c, SGM_ESTOPPED, "stopped by sgm_stop");
    }
    return SGM_OK;
}

int sgm_enumerate_batch(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq, const sgm_enum_opts* opts,
                        sgm_stats* out)
{
    if (!qs || !out || nq == 0) return SGM_EINVAL;
    for (uint32_t i = 0; i < nq; i++)
        if (!qs[i]) return SGM_EINVAL;
    sgm_graph* g = qs[0]->g;
    sgm_ctx* c = g->ctx;
    const sgm_enum_opts o = normalize_opts(opts);
    std::memset(out, 0, sizeof(sgm_stats) * nq);
    if (is_aux_method(o.method)) {
        // the aux-table search over many queries: orders (given or RI) checked like plans
        std::vector<std::array<uint32_t, SGM_MAXQ>> ords(nq);
        std::vector<const uint32_t*> optr(nq);
        for (uint32_t i = 0; i < nq; i++) {
            sgm_query* q = qs[i];
            if (q->g != g) return set_err(c, SGM_EINVAL, "queries of a batch must share one graph");
            if (const int frc = check_filtered(q, "enumerate")) return frc;
            if (q->q.n < 2) return set_err(c, SGM_EINVAL, "enumeration needs at least 2 query vertices (query %u)", i);
            if (orders && orders[i]) std::copy(orders[i], orders[i] + q->q.n, ords[i].begin());
            else sgm::order_ri(q->q, ords[i].data());
            if (sgm::check_order(q->q, ords[i].data()))
                return set_err(c, SGM_EORDER, "order of query %u is not a connected permutation", i);
            optr[i] = ords[i].data();
        }
        HIPCHK(c, hipSetDevice(c->device));
        return enumerate_aux_batch(qs, optr.data(), nq, o, out);
    }
    if (o.method == SGM_ENUM_DYN_DAF || o.method == SGM_ENUM_DYN_VEQ) {
        // start orders: given, or BFS from DAFStartVertex (the reference driver's)
        std::vector<std::array<uint32_t, SGM_MAXQ>> ords(nq);
        std::vector<const uint32_t*> optr(nq);
        HIPCHK(c, hipSetDevice(c->device));
        for (uint32_t i = 0; i < nq; i++) {
            sgm_query* q = qs[i];
            if (q->g != g) return set_err(c, SGM_EINVAL, "queries of a batch must share one graph");
            if (const int frc = check_filtered(q, "enumerate")) return frc;
            if (q->q.n < 2) return set_err(c, SGM_EINVAL, "enumeration needs at least 2 query vertices (query %u)", i);
            if (orders && orders[i]) std::copy(orders[i], orders[i] + q->q.n, ords[i].begin());
            else if (const int orc = dyn_default_order(q, ords[i].data())) return orc;
            if (sgm::check_order(q->q, ords[i].data()))
                return set_err(c, SGM_EORDER, "order of query %u is not a connected permutation", i);
            optr[i] = ords[i].data();
        }
        return enumerate_dyn_batch(qs, optr.data(), nq, o, out);
    }
    if (o.method != SGM_ENUM_CACHE_FILTER_FS && o.method != SGM_ENUM_CACHE_FILTER)
        return set_err(c, SGM_EINVAL, "unknown enumeration method %d", o.method);
    // plans (host), validated before anything runs; plans, query descriptors and statistics
    // live in pinned staging (plain vectors if none can be had)
    const size_t pl_bytes = stage_align(sizeof(sgm_plan_dev) * nq);
    const size_t eq_bytes = stage_align(sizeof(sgm_equery) * nq);
    std::vector<sgm_plan_dev> plans_v;
    std::vector<sgm_equery> eq_v;
    std::vector<sgm_stats_dev> got_v;
    sgm_plan_dev* plans = nullptr;
    sgm_equery* eqs = nullptr;
    sgm_stats_dev* got = nullptr;
    if (uint8_t* st = static_cast<uint8_t*>(host_stage(c, pl_bytes + eq_bytes + sizeof(sgm_stats_dev) * nq))) {
        plans = reinterpret_cast<sgm_plan_dev*>(st);
        eqs = reinterpret_cast<sgm_equery*>(st + pl_bytes);
        got = reinterpret_cast<sgm_stats_dev*>(st + pl_bytes + eq_bytes);
    } else {
        plans_v.resize(nq);
        eq_v.resize(nq);
        got_v.resize(nq);
        plans = plans_v.data();
        eqs = eq_v.data();
        got = got_v.data();
    }
    std::vector<uint32_t> live;
    std::vector<std::array<uint32_t, SGM_MAXQ>> ord(nq);  // the orders used (given or RI)
    uint32_t max_ops = 1;
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        if (q->g != g) return set_err(c, SGM_EINVAL, "queries of a batch must share one graph");
        if (const int frc = check_filtered(q, "enumerate")) return frc;
        if (q->q.n < 2) return set_err(c, SGM_EINVAL, "enumeration needs at least 2 query vertices (query %u)", i);
    }
    // orders and plans of large launches on host threads; errors reported in query order
    std::vector<int> prcs(nq, 0);
    sgm::parallel_for(nq, 32, [&](uint32_t i) {
        sgm_query* q = qs[i];
        sgm::Plan pl;
        if (orders && orders[i]) std::copy(orders[i], orders[i] + q->q.n, ord[i].begin());
        else sgm::order_ri(q->q, ord[i].data());
        prcs[i] = sgm::build_plan(q->q, ord[i].data(), pl);
        if (!prcs[i]) sgm::plan_to_device(pl, plans[i]);
    });
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        const int prc = prcs[i];
        if (prc == -2) return set_err(c, SGM_EORDER, "order of query %u is not a connected permutation", i);
        if (prc) return set_err(c, SGM_EINVAL, "plan construction failed for query %u (%d)", i, prc);
        max_ops = std::max<uint32_t>(max_ops, plans[i].n_ops);
        if (!q->empty) live.push_back(i);  // the reference's driver reports 0 without enumerating
    }
    if (live.empty()) return SGM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = grow(c, &c->d_plans, c->batch_cap, nq);
    if (rc) return rc;
    uint32_t cap2 = c->batch_cap2;
    if ((rc = grow(c, &c->d_eq, cap2, nq))) return rc;
    uint32_t cap3 = c->batch_cap2;
    if ((rc = grow(c, &c->d_qstats, cap3, nq))) return rc;
    uint32_t cap4 = 3 * c->batch_cap2;
    if ((rc = grow(c, &c->d_qstop, cap4, 3 * nq))) return rc;  // stop words, item and chunk counters
    c->batch_cap2 = std::max(c->batch_cap2, nq);
    HIPCHK(c, hipMemcpyAsync(c->d_plans, plans, sizeof(sgm_plan_dev) * nq, hipMemcpyHostToDevice, c->stream));
    const uint32_t blocks = default_blocks(c, o, (uint32_t)live.size());
    const uint32_t waves = blocks * 4;
    uint64_t words = o.stack_words ? o.stack_words : (uint64_t)8 * g->max_deg + 1024;
    const int use_fs = o.method == SGM_ENUM_CACHE_FILTER_FS ? 1 : 0;
    const int serial = (o.flags & SGM_ENUM_F_SERIAL_LAST) ? 1 : 0;
    const uint32_t* hub_slot = nullptr;
    if ((rc = hub_slot_map(g, o, &hub_slot))) return rc;
    std::vector<sgm_stats_dev> st(nq);
    const size_t lb_words = ((size_t)g->V + 31) / 32;
    if (c->lastbits_words < lb_words * live.size()) {
        dfree(c->d_lastbits);
        c->d_lastbits = nullptr;
        c->lastbits_words = 0;
        HIPCHK(c, dalloc(&c->d_lastbits, lb_words * live.size()));
        c->lastbits_words = lb_words * live.size();
    }
    // depth-1 splitting for queries with few candidate roots (or all, or none)
    std::vector<char> want_split(nq, 0);
    uint32_t n_split = 0;
    for (uint32_t i : live) {
        const uint32_t roots = qs[i]->counts[ord[i][0]];
        want_split[i] = !(o.flags & SGM_ENUM_F_NO_SPLIT) &&
                        ((o.flags & SGM_ENUM_F_SPLIT_ALL) || roots <= SGM_SPLIT_MAX_ROOTS || narrow_roots(o));
        n_split += want_split[i] ? 1u : 0u;
    }
    const bool split = n_split > 0;
    if (c->split_cap < n_split) {  // split blocks and records for the split queries only
        dfree(c->d_split);
        dfree(c->d_rec);
        c->d_split = nullptr;
        c->d_rec = nullptr;
        c->split_cap = 0;
        HIPCHK(c, dalloc(&c->d_split, (size_t)SGM_SPLIT_WORDS * n_split));
        HIPCHK(c, dalloc(&c->d_rec, (size_t)SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS * n_split));
        c->split_cap = n_split;
    }
    for (uint32_t attempt = 0; !live.empty(); attempt++) {
        if (attempt > 3) return set_err(c, SGM_EDEVICE, "set-op stack overflow after %u retries", attempt - 1);
        if (words > 0x7FFFFFFFull) return set_err(c, SGM_EDEVICE, "set-op stack would exceed 2^31 words");
        rc = ensure_stack(c, (size_t)waves * words * sizeof(uint32_t));
        if (rc) return rc;
        sgm_equery* eq = eqs;
        std::fill(eq, eq + live.size(), sgm_equery());
        const uint32_t m = (uint32_t)live.size();
        uint32_t js = 0;  // next split block
        for (size_t k = 0; k < live.size(); k++) {
            const uint32_t i = live[k];
            sgm_query* q = qs[i];
            const uint32_t u0 = ord[i][0];
            eq[k].plan = c->d_plans + i;
            eq[k].vmask = q->d_vmask;
            eq[k].lastbits = c->d_lastbits + lb_words * k;
            eq[k].stop = c->d_qstop + k;       // stop words: [0, m)
            eq[k].next = c->d_qstop + m + k;   // item counters: [m, 2m)
            eq[k].snext = c->d_qstop + 2 * m + k;  // split chunk counters: [2m, 3m)
            if (want_split[i]) {
                eq[k].split = c->d_split + (size_t)SGM_SPLIT_WORDS * js;
                eq[k].rec = c->d_rec + (size_t)SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS * js;
                js++;
            }
            eq[k].st = c->d_qstats + k;
            eq[k].cap = (unsigned long long)o.cap;
            eq[k].budget = budget_ticks(c, o);
            root_range(o, &eq[k].root_lo, &eq[k].root_hi);
            eq[k].n_items = g->ge_cnt[std::min<uint32_t>(q->q.off[u0 + 1] - q->q.off[u0], 32)];
            eq[k].root_u = u0;
        }
        HIPCHK(c, hipMemcpyAsync(c->d_eq, eq, sizeof(sgm_equery) * m, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_qstats, 0, sizeof(sgm_stats_dev) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_qstop, 0, sizeof(uint32_t) * 3 * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_stop, 0, sizeof(uint32_t), c->stream));
        clear_host_stop(c);
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        sgm::launch_last_bits(c->stream, c->d_eq, m, g->V);
        if (split) sgm::launch_split_roots(c->stream, g->d_off, g->d_bydeg, c->d_eq, m, g->V, o.split_chunks);
        sgm::launch_enum_cache(c->stream, blocks, g->d_off, g->d_adj, g->d_bydeg, c->d_eq, m, max_ops, c->d_stack,
                               (uint32_t)words, c->d_stop, c->d_hstop, use_fs, hub_slot, g->d_hub_bits, g->hub_words,
                               serial, split);
        if (split) sgm::launch_split_reduce(c->stream, c->d_eq, m);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipMemcpyAsync(got, c->d_qstats, sizeof(sgm_stats_dev) * m, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        std::vector<uint32_t> again;
        for (size_t k = 0; k < live.size(); k++) {
            const uint32_t i = live[k];
            out[i].kernel_ms += ms;
            if (got[k].overflow) {
                out[i].retries++;
                again.push_back(i);
                continue;
            }
            stats_out(got[k], o.cap, &out[i], c->wall_ns_per_tick);
            if (c->stop_requested && !out[i].capped) out[i].stopped = 1;
        }
        live.swap(again);
        words *= 4;
    }
    if (c->stop_requested) {
        for (uint32_t i = 0; i < nq; i++)
            if (out[i].stopped) return set_err(c, SGM_ESTOPPED, "stopped by sgm_stop");
    }
    return SGM_OK;
}

int sgm_enumerate(sgm_query* q, const uint32_t* order, const sgm_enum_opts* opts, sgm_stats* out)
{
    if (!q || !order || !out) return SGM_EINVAL;
    sgm_ctx* c = q->g->ctx;
    const sgm_enum_opts o = normalize_opts(opts);
    if (is_aux_method(o.method)) {
        std::memset(out, 0, sizeof *out);
        if (const int frc = check_filtered(q, "enumerate")) return frc;
        if (q->q.n < 2) return set_err(c, SGM_EINVAL, "enumeration needs at least 2 query vertices");
        sgm::Plan pl;
        const int prc = sgm::build_plan(q->q, order, pl);
        if (prc == -2) return set_err(c, SGM_EORDER, "order is not a connected permutation");
        if (prc) return set_err(c, SGM_EINVAL, "plan construction failed (%d)", prc);
        if (q->empty) return SGM_OK;
        HIPCHK(c, hipSetDevice(c->device));
        return enumerate_aux_batch(&q, &order, 1, o, out);
    }
    if (o.method == SGM_ENUM_DYN_DAF || o.method == SGM_ENUM_DYN_VEQ) return sgm_enumerate_batch(&q, &order, 1, &o, out);
    if (o.method != SGM_ENUM_CACHE_FILTER_FS && o.method != SGM_ENUM_CACHE_FILTER)
        return set_err(c, SGM_EINVAL, "unknown enumeration method %d", o.method);
    return sgm_enumerate_batch(&q, &order, 1, &o, out);
}

}  // extern "C"
