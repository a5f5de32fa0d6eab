// Library objects behind the C ABI's opaque handles (context, resident graph, query with
// its candidate space) and the helpers the driver sources share.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sgm_c.h"
#include "sgm_host.h"
#include "sgm_internal.h"
#include "sgm_kernels.h"

struct sgm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t ctl = nullptr;       // non-blocking stream for sgm_stop
    std::string err;
    uint32_t* d_stop = nullptr;      // stop word polled by the enumeration kernels
    // the host's stop request: one word of pinned, coherent host memory (h_stop as the host
    // writes it, d_hstop as the kernels read it), relayed into d_stop by the polling waves
    uint32_t* h_stop = nullptr;
    const uint32_t* d_hstop = nullptr;
    uint32_t* d_work = nullptr;      // work counter
    sgm_stats_dev* d_stats = nullptr;
    sgm_auxplan_dev* d_auxplan = nullptr;   // (unused since the batched aux search; kept allocated)
    // batched aux-table search: per-query plans, item prefix, statistics, stop words
    sgm_auxplan_dev* d_auxplans = nullptr;
    uint32_t* d_afirst = nullptr;
    sgm_stats_dev* d_astats = nullptr;
    uint32_t* d_astop = nullptr;
    uint32_t* d_anext = nullptr;  // per-query item counters
    uint32_t aux_cap = 0, afirst_cap = 0, astats_cap = 0, astop_cap = 0, anext_cap = 0;
    sgm_dynplan_dev* d_dynplans = nullptr;   // dynamic-order search plans (shares the buffers above)
    uint32_t dyn_cap = 0;
    sgm_wdesc* d_wdesc = nullptr;            // DAF weight row passes of a batch
    uint32_t wdesc_cap = 0;
    unsigned long long* d_work64 = nullptr;  // item counter of the batched search
    sgm_plan_dev* d_plans = nullptr;          // per-query plans of a batch
    sgm_equery* d_eq = nullptr;
    sgm_stats_dev* d_qstats = nullptr;
    uint32_t* d_qstop = nullptr;
    uint32_t batch_cap = 0, batch_cap2 = 0;
    uint32_t* d_stack = nullptr;
    size_t stack_bytes = 0;
    uint32_t* d_gqls = nullptr;      // GQL filter scratch: touch + dirty marks, 2*V words per query
    size_t gqls_words = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t fev[4] = {nullptr, nullptr, nullptr, nullptr};
    double filter_ms[3] = {0, 0, 0};
    int num_cus = 256;
    double wall_ns_per_tick = 10.0;  // device wall clock (s_memrealtime) period
    volatile int stop_requested = 0;
    int enum_inflight = 0;           // batch enumerations launched and not yet waited for
    // filter batches: descriptors and the gathered candidate counts (device, pinned host)
    sgm_qdesc* d_fqd = nullptr;
    sgm_cdesc* d_fcd = nullptr;
    uint32_t* d_fcnt = nullptr;
    uint32_t* h_fcnt = nullptr;
    uint32_t fcap = 0;
    // BFS-tree filters: NLF degree sums (the statistics kernel's second output) and the
    // per-query refinement programs (SGM_RPROG_WORDS words each)
    unsigned long long* d_fdeg = nullptr;
    uint32_t* d_rprog = nullptr;
    uint32_t* h_rprog = nullptr;
    uint32_t rprog_cap = 0;
    // aux-table builds of many queries: descriptors (count / scan / fill) and the totals and
    // maxima of every directed edge's scan
    uint8_t* d_tb = nullptr;
    size_t tb_bytes = 0;
    uint32_t* d_ttot = nullptr;
    uint32_t ttot_cap = 0;
    uint32_t* d_rrec = nullptr;  // rank records of the queries of one table build pass
    size_t rrec_cap = 0;
    // the depth-0 candidates of an aux-table / dynamic-order launch's queries in degree order
    uint32_t* d_rootord = nullptr;
    size_t rootord_cap = 0;
    sgm_rootdesc* d_rootdesc = nullptr;
    uint32_t rootdesc_cap = 0;
    // candidate-list compaction of many queries in one launch: descriptors, prefix arrays
    sgm_cdesc* d_cmpdesc = nullptr;
    uint32_t cmpdesc_cap = 0;
    uint32_t* d_cmpoff = nullptr;
    size_t cmpoff_cap = 0;
    // batched enumeration: bitmap of C(u_L) per query; depth-1 split blocks and chunk records
    uint32_t* d_lastbits = nullptr;
    size_t lastbits_words = 0;
    uint32_t* d_split = nullptr;
    unsigned long long* d_rec = nullptr;
    uint32_t split_cap = 0;  // queries the split buffers hold
    // pinned host staging of one call's descriptor uploads and statistics downloads (every
    // call synchronises its stream before it returns, so the next call may reuse it)
    void* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    // objects that hold this context (destroyed with it if the caller has not done so)
    std::vector<sgm_batch*> batches;
    std::vector<sgm_costmodel*> models;
};

template <class T>
inline void forget(std::vector<T*>& v, T* x)
{
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == x) { v[i] = v.back(); v.pop_back(); return; }
}

struct sgm_graph {
    sgm_ctx* ctx = nullptr;
    uint32_t V = 0;
    uint64_t adj_len = 0;
    uint32_t max_deg = 0;
    uint32_t* d_off = nullptr;
    uint32_t* d_adj = nullptr;
    uint32_t* d_bydeg = nullptr;
    uint32_t* d_gecnt = nullptr;
    uint32_t ge_cnt[33] = {0};
    // adjacency bitmaps of the highest-degree vertices (degree >= kHubMinDegree, within
    // kHubBudgetBytes); hub_slot[v] = bitmap index or SGM_INVALID
    uint32_t n_hubs = 0;
    uint32_t hub_words = 0;
    uint32_t* d_hub_slot = nullptr;
    uint32_t* d_hub_bits = nullptr;
    // degree-rank space of the GQL filter: rank[v] (position in by_deg), the CSR of ranks
    // (roff[V+1], radj[2E]) and the neighbour-degree table (4 planes of V words)
    uint32_t* d_rank = nullptr;
    uint32_t* d_roff = nullptr;
    uint32_t* d_radj = nullptr;
    uint64_t* d_dtab = nullptr;
    // vertex labels (labelled NLF / GQL): lab[V] by id, and each vertex's neighbour-label
    // counts as runs ascending by label: run_off[V+1], run_lab[R], run_cnt[R].  Built by
    // sgm_graph_set_labels, or with every label 0 for the first labelled query.
    bool labelled = false;  // some label is not 0
    bool has_lab = false;   // the label tables are on the device
    uint64_t lab_gen = 0;        // bumped whenever the label tables are rebuilt or dropped
    uint64_t label_version = 0;  // bumped whenever sgm_graph_set_labels changes the labels
    uint32_t* d_lab = nullptr;
    uint32_t* d_run_off = nullptr;
    uint32_t* d_run_lab = nullptr;
    uint32_t* d_run_cnt = nullptr;
    // label histogram of a labelled graph (Graph::getLabelFreq / getMaxLabelFreq,
    // inc/graph.h:339-345); empty = every label 0
    std::unordered_map<uint32_t, uint32_t> lab_hist;
    uint32_t max_lab_freq = 0;
    std::vector<uint32_t> hub_vertex;                   // bitmap slot -> vertex (degree descending)
    std::vector<uint32_t> hub_degree;
    std::vector<std::pair<uint32_t, uint32_t*>> hub_maps;  // (min degree, slot map) for higher thresholds
};

// GQL scratch (3 V words per query: rank-space membership, touch and dirty marks) is
// allocated per filter batch; larger batches run in chunks of this size.
static const uint64_t kGqlScratchBytes = 8ull << 30;

static const uint32_t kHubMinDegree = 64;
static const uint64_t kHubBudgetBytes = 1ull << 30;

// Table of directed query edge (a, b): row r (the r-th candidate v of a) is
// d_pool[d_voff[v] .. d_voff[v + 1]), positions in C(b).  d_voff (V + 1 offsets by data
// vertex id) and d_pool are b's row pool, shared by every edge into b; vertices outside the
// pool have empty rows.
struct sgm_table {
    bool built = false;
    uint64_t e_cnt = 0;
    uint32_t max_deg = 0;
    const uint32_t* d_voff = nullptr;
    const uint32_t* d_pool = nullptr;
};

struct sgm_query {
    sgm_graph* g = nullptr;
    sgm::QueryCSR q;
    uint32_t nbr[SGM_MAXQ] = {0};
    int filtered = 0;                 // SGM_FILTER_* after a successful or empty filter
    uint64_t filter_label_version = 0; // the graph's label_version the candidates were filtered under
    bool empty = false;
    bool compacted = false;
    uint32_t counts[SGM_MAXQ] = {0};
    uint32_t cand_off[SGM_MAXQ + 1] = {0};
    uint32_t* d_vmask = nullptr;
    uint32_t* d_removed = nullptr;
    uint64_t* d_bm = nullptr;
    uint32_t* d_rank = nullptr;
    uint32_t* d_cnt = nullptr;
    uint32_t* d_candoff = nullptr;
    uint32_t* d_list = nullptr;
    size_t list_cap = 0;  // d_list's capacity (kept across filters)
    uint32_t* d_lreq = nullptr;       // labelled NLF requirements (sgm_qdesc::lreq), built on demand
    bool labelled = false;            // some query label is not 0 (labels in q.lab)
    // NLF candidate counts (NLFByVertex, inc/filter.h:24-58) for the start vertices of the
    // BFS-tree filters and orders, valid for the graph's label_version nlf_version
    uint32_t nlf_cnt[SGM_MAXQ] = {0};
    bool nlf_ok = false;
    uint64_t nlf_version = 0;
    sgm_qdesc* d_qdesc = nullptr;
    sgm_cdesc* d_cdesc = nullptr;
    sgm_qdesc hq;                     // host copies of the descriptors
    sgm_cdesc hc;
    std::vector<sgm_table> tables;    // n*n, directed (a,b): views into the arenas below
    // table arenas, grown as needed and kept across filters (no allocation when the query is
    // filtered and tabled again): the row pools' offsets by vertex (n blocks of V + 1) and
    // entries (positions in C(b))
    uint32_t* d_tvoff = nullptr;
    uint32_t* d_tpool = nullptr;
    size_t tvoff_cap = 0, tpool_cap = 0;
    // DAF weights of the dynamic-order search (computeWeightArray), per candidate position in
    // the cand_off layout, for the start order weight_order; dropped with the tables
    uint32_t* d_weight = nullptr;
    size_t weight_cap = 0;
    bool weights_ok = false;
    uint32_t weight_order[SGM_MAXQ] = {0};
    // GuP conflict-detection guards (4 words per (depth, candidate position), depth d's from
    // guard_off[d]) for the order guard_order; dropped with the tables
    uint32_t* d_guard = nullptr;
    size_t guard_cap = 0;
    bool guards_ok = false;
    uint32_t guard_order[SGM_MAXQ] = {0};
    uint32_t guard_off[SGM_MAXQ + 1] = {0};
};

inline int set_err(sgm_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return set_err((ctx), e_ == hipErrorOutOfMemory ? SGM_ENOMEM : SGM_EHIP, "%s: %s (%s:%d)", \
                           #call, hipGetErrorString(e_), __FILE__, __LINE__);                      \
    } while (0)

template <class T>
inline hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

inline void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// The context's pinned staging area, grown to `bytes` (nullptr when pinned memory runs out).
inline void* host_stage(sgm_ctx* c, size_t bytes)
{
    if (c->h_stage_bytes < bytes) {
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = nullptr;
        c->h_stage_bytes = 0;
        if (hipHostMalloc(&c->h_stage, bytes) != hipSuccess) {
            c->h_stage = nullptr;
            return nullptr;
        }
        c->h_stage_bytes = bytes;
    }
    return c->h_stage;
}

inline size_t stage_align(size_t x) { return (x + 255) & ~(size_t)255; }

// SGM_ESTATE unless q's candidates are filtered under the graph's current labels.
inline int check_filtered(sgm_query* q, const char* what)
{
    if (!q->filtered) return set_err(q->g->ctx, SGM_ESTATE, "%s: filter the query first", what);
    if (q->filter_label_version != q->g->label_version)
        return set_err(q->g->ctx, SGM_ESTATE, "%s: the graph's labels changed since the query was filtered", what);
    return SGM_OK;
}

// (Re)allocate a context buffer so that it holds at least `need` elements.
template <class T>
inline int grow(sgm_ctx* c, T** p, uint32_t& cap_now, uint32_t need)
{
    if (cap_now >= need) return SGM_OK;
    dfree(*p);
    *p = nullptr;
    cap_now = 0;
    HIPCHK(c, dalloc(p, need));
    cap_now = need;
    return SGM_OK;
}

// Label tables of the labelled filters (labels == nullptr: every label 0), and the lreq block
// (sgm_qdesc::lreq) of a query; shared by the filter drivers (sgm_capi.cpp, sgm_batch.cpp).
int build_label_tables(sgm_ctx* c, sgm_graph* g, const uint32_t* labels);
std::vector<uint32_t> lreq_block(const sgm::QueryCSR& q);
int ensure_lreq(sgm_ctx* c, sgm_query* q);
// The data graph's label statistics for the start-vertex rules.
inline sgm::LabelStats label_stats(const sgm_graph* g)
{
    sgm::LabelStats ls;
    ls.V = g->V;
    ls.hist = g->lab_hist.empty() ? nullptr : &g->lab_hist;
    ls.max_freq = g->lab_hist.empty() ? g->V : g->max_lab_freq;
    return ls;
}
// NLF counts of every query vertex (cached in the query; computed on the device when the
// graph or the query carries labels, from the degree histogram otherwise).
int query_nlf_counts(sgm_query* q, uint32_t* out);

// The no-allocation entry point's static scratch array (sgm_noalloc.hip): looks up its
// address on the context's device, once per device (the device must be current).
int noalloc_setup(sgm_ctx* c);

// Shared by the enumeration drivers (sgm_capi.cpp, sgm_batch.cpp).
// Cost-model work buffers filled on the device, and a forward pass over them (ranking).
int costmodel_work_inputs(sgm_costmodel* m, uint32_t S, uint16_t** F, uint8_t** valid, uint32_t** back, uint32_t** fwd);
// forward pass over the work buffers on stream s, the stream their inputs were built on
int costmodel_forward_work(sgm_costmodel* m, uint32_t S, const float** y, hipStream_t s);
// Slot map of the hub bitmaps to use for opts (nullptr = binary search everywhere).
int hub_slot_map(sgm_graph* g, const sgm_enum_opts& o, const uint32_t** out);
int ensure_stack(sgm_ctx* c, size_t bytes);
// Aux tables (buildTables) of every directed edge of the queries, in one set of launches.
int build_tables_many(sgm_query* const* qs, uint32_t nq);
// root_order[i] (device, queued on the context's stream) = the candidates of query vertex
// us[i] of qs[i] in degree order as positions in its candidate list; null for an empty query
int root_orders(sgm_ctx* c, sgm_query* const* qs, const uint32_t* us, uint32_t nq,
                std::vector<const uint32_t*>& root_order);
// edge (a, b)'s table in the reference's layout, gathered on the host from b's row pool
int table_fetch(sgm_query* q, uint32_t a, uint32_t b, std::vector<uint32_t>& off, std::vector<uint32_t>& idx);
// The dynamic-order searches over many queries (sgm_dynamic.cpp): orders[i] = start order
// (checked); and their default start order, BFS from DAFStartVertex.
int enumerate_dyn_batch(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq, const sgm_enum_opts& o,
                        sgm_stats* out);
// GuP conflict-detection guards of q for `order` (sgm_gup.cpp; tables built), to q->d_guard.
int gup_guards_build(sgm_query* q, const uint32_t* order);
int dyn_default_order(sgm_query* q, uint32_t* order);
sgm_enum_opts normalize_opts(const sgm_enum_opts* opts);
uint32_t default_blocks(const sgm_ctx* c, const sgm_enum_opts& o, uint32_t nq);
void stats_out(const sgm_stats_dev& st, uint64_t cap, sgm_stats* out, double ns_per_tick);
// Device-clock ticks of a per-query time budget (sgm_enum_opts.time_limit_ms; 0 = none).
unsigned long long budget_ticks(const sgm_ctx* c, const sgm_enum_opts& o);
// The aux-table search methods (plain, DAF failing sets, GuP conflict detection, both).
inline bool is_aux_method(int m)
{
    return m == SGM_ENUM_AUX_TABLES || m == SGM_ENUM_AUX_TABLES_FS || m == SGM_ENUM_AUX_TABLES_GUPCD ||
           m == SGM_ENUM_AUX_TABLES_FS_GUPCD;
}
// The root range of opts as [lo, hi) (root_end == 0: every vertex).
inline void root_range(const sgm_enum_opts& o, uint32_t* lo, uint32_t* hi)
{
    *lo = o.root_begin;
    *hi = o.root_end ? o.root_end : 0xFFFFFFFFu;
}
// A root range of at most SGM_SPLIT_ROOTS ids: every candidate in it gets its depth-1 loop
// split (k_split_roots reads the range directly), so a heavy root is not left to one wave.
inline bool narrow_roots(const sgm_enum_opts& o)
{
    return o.root_end && o.root_end - o.root_begin <= SGM_SPLIT_ROOTS;
}
// Clears the host's stop request before a launch (only when no launch is in flight: a
// request then ends every one of them).
inline void clear_host_stop(sgm_ctx* c)
{
    __atomic_store_n(c->h_stop, 0u, __ATOMIC_SEQ_CST);
    c->stop_requested = 0;
}
