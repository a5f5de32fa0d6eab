// Batched runner: up to `slots` queries of one resident data graph per batch, with
// per-slot device state reused across batches, and the filter and enumeration launches of
// a batch captured once as hipGraphs and replayed for every later batch.
//
// Filter graph (per filter kind): clear flags / GQL marks / statistics with fill kernels,
// NLF, the 64 GQL refinement phases (a phase with nothing to do for a slot returns at
// once), candidate counts and degree sums.  Enumeration graph (per launch shape): clear the
// per-query statistics, stop words and item counters, then one persistent cache + filter
// (+ failing-set) launch over all slots.  Unused slots have n = 0 (filter) and no items
// (enumeration).  Orders are chosen on the host between the two graphs, from the query
// alone (RI) or from the candidate statistics (GQL order, a learned cost model).
//
// Every runner has its own stream, and an enumeration can be launched without waiting
// (sgm_batch_enumerate_launch / _wait): with two runners the host filters batch i+1 while
// batch i enumerates, and batch i+1's persistent grid takes over the GPU as batch i's
// waves run out of work, instead of the GPU idling through each batch's tail.
#include "sgm_costmodel.h"
#include "sgm_objects.h"

namespace {

struct EnumGraph {
    uint32_t max_ops, use_fs, blocks, words, serial, split, chunks;
    const uint32_t* hub;
    hipGraphExec_t exec;
};

}  // namespace

struct sgm_batch {
    sgm_graph* g = nullptr;
    sgm_ctx* c = nullptr;
    hipStream_t stream = nullptr;            // all of this runner's work (default priority)
    uint32_t slots = 0;
    uint32_t* d_vmask = nullptr;             // slots * V
    uint32_t* d_flags = nullptr;             // slots * SGM_GQL_FLAG_WORDS
    uint32_t* d_scratch = nullptr;           // slots * sgm_gql_scratch_words(V): GQL scratch
    sgm_qdesc* d_qd = nullptr;               // slots
    uint32_t* d_cnt = nullptr;               // slots * 32
    unsigned long long* d_degsum = nullptr;  // slots * 32
    sgm_plan_dev* d_plans = nullptr;         // slots
    sgm_equery* d_eq = nullptr;              // slots
    sgm_stats_dev* d_st = nullptr;           // slots
    uint32_t* d_qstop = nullptr;             // 3 * slots: stop words, item counters, split chunk counters
    uint32_t* d_split = nullptr;             // slots * SGM_SPLIT_WORDS: depth-1 split blocks
    unsigned long long* d_rec = nullptr;     // slots * SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS: chunk records
    uint32_t* d_lastbits = nullptr;          // slots * ceil(V/32): bitmap of C(u_L) per slot
    uint32_t* d_stack = nullptr;
    size_t stack_words_total = 0;
    // learned-order ranking: given RI / GQL orders, candidate orders, the choice per query
    uint8_t* d_rank_given = nullptr;         // slots * 2 * 32
    uint8_t* d_rank_orders = nullptr;        // rank_cap * 32
    uint32_t rank_cap = 0;
    uint32_t* d_rank_out = nullptr;          // slots * 32
    float* d_rank_pred = nullptr;            // 2 * slots: chosen, RI
    int32_t* d_rank_flag = nullptr;          // slots
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // the loaded batch
    uint32_t nq = 0;
    int kind = 0;
    std::vector<sgm::QueryCSR> q;
    std::vector<sgm_qdesc> hqd;
    std::vector<uint32_t> counts;
    std::vector<unsigned long long> degsum;
    std::vector<int32_t> status;
    double filter_ms = 0, enum_ms = 0;
    hipGraphExec_t gf_nlf[2] = {nullptr, nullptr};            // unlabeled / labelled batches
    std::vector<std::pair<uint32_t, hipGraphExec_t>> gf_gql;  // by 2 * (number of GQL phases) + labelled
    uint32_t* d_lreq = nullptr;                               // slots * kLreqWords: labelled NLF requirements
    uint64_t lab_gen = ~0ull;                                 // the graph's lab_gen the labelled graphs captured
    std::vector<EnumGraph> ge;
    // an enumeration launched and not yet waited for
    bool pending = false;
    sgm_enum_opts pend_o{};
    std::vector<sgm_equery> pend_eq;
    uint32_t pend_bucket = 0, pend_blocks = 0;
    uint64_t pend_words = 0;
    int pend_use_fs = 0, pend_serial = 0;
    bool pend_split = false;
    const uint32_t* pend_hub = nullptr;
    sgm_stats_dev* h_got = nullptr;          // pinned: per-slot statistics copied back
};

namespace {

// lreq words of the largest query (32 vertices, 32 * 31 adjacency entries)
const size_t kLreqWords = sgm_lreq_words(SGM_MAXQ, (size_t)SGM_MAXQ * (SGM_MAXQ - 1));

int load_query(sgm_ctx* c, const uint32_t* qoff, const uint32_t* qadj, uint32_t n, sgm::QueryCSR& q, sgm_qdesc& qd)
{
    if (!qoff || (!qadj && qoff[n]) || n < 2 || n > SGM_MAXQ)
        return set_err(c, SGM_EINVAL, "batch queries need 2..32 vertices");
    if (qoff[0] != 0) return set_err(c, SGM_EINVAL, "query offsets must start at 0");
    q.n = n;
    q.off.assign(qoff, qoff + n + 1);
    q.adj.assign(qadj, qadj + qoff[n]);
    std::memset(&qd.qdeg, 0, sizeof qd.qdeg);
    std::memset(&qd.nbr, 0, sizeof qd.nbr);
    std::memset(&qd.req, 0, sizeof qd.req);
    qd.n = n;
    for (uint32_t u = 0; u < n; u++) {
        if (qoff[u + 1] < qoff[u]) return set_err(c, SGM_EINVAL, "query offsets not monotone");
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
            const uint32_t w = qadj[p];
            if (w >= n || w == u || (p > qoff[u] && w <= qadj[p - 1]))
                return set_err(c, SGM_EINVAL, "query adjacency of %u malformed", u);
            bool back = false;
            for (uint32_t r = qoff[w]; r < qoff[w + 1]; r++) back |= qadj[r] == u;
            if (!back) return set_err(c, SGM_EINVAL, "query graph not symmetric");
            qd.nbr[u] |= 1u << w;
        }
        qd.qdeg[u] = qoff[u + 1] - qoff[u];
    }
    for (uint32_t u = 0; u < n; u++) {
        std::vector<uint32_t> r;
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) r.push_back(qd.qdeg[qadj[p]]);
        std::sort(r.begin(), r.end(), [](uint32_t a, uint32_t b) { return a > b; });
        for (size_t j = 0; j < r.size(); j++) reinterpret_cast<uint8_t*>(qd.req[u])[j] = (uint8_t)r[j];
    }
    return SGM_OK;
}

// Records `body` (launches on stream s) into an executable graph.
template <class F>
int capture(sgm_ctx* c, hipStream_t s, F&& body, hipGraphExec_t* out)
{
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    body();
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(s, &graph);
    if (le != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_err(c, SGM_EHIP, "graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ee));
    }
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return set_err(c, SGM_EHIP, "graph instantiate: %s", hipGetErrorString(ie));
    return SGM_OK;
}

uint32_t ops_bucket(uint32_t max_ops)
{
    uint32_t b = 64;
    while (b < max_ops) b <<= 1;
    return std::min<uint32_t>(b, SGM_MAXOPS);
}

}  // namespace

extern "C" {

void sgm_batch_destroy(sgm_batch* b)
{
    if (!b) return;
    if (b->c) {
        forget(b->c->batches, b);
        (void)hipSetDevice(b->c->device);
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        if (b->pending) b->c->enum_inflight--;
    }
    for (auto x : b->gf_nlf)
        if (x) (void)hipGraphExecDestroy(x);
    dfree(b->d_lreq);
    for (auto& x : b->gf_gql) (void)hipGraphExecDestroy(x.second);
    for (auto& x : b->ge) (void)hipGraphExecDestroy(x.exec);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    dfree(b->d_vmask);
    dfree(b->d_flags);
    dfree(b->d_scratch);
    dfree(b->d_qd);
    dfree(b->d_cnt);
    dfree(b->d_degsum);
    dfree(b->d_plans);
    dfree(b->d_eq);
    dfree(b->d_st);
    dfree(b->d_qstop);
    dfree(b->d_split);
    dfree(b->d_rec);
    dfree(b->d_lastbits);
    dfree(b->d_stack);
    dfree(b->d_rank_given);
    dfree(b->d_rank_orders);
    dfree(b->d_rank_out);
    dfree(b->d_rank_pred);
    dfree(b->d_rank_flag);
    if (b->h_got) (void)hipHostFree(b->h_got);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int sgm_batch_create(sgm_graph* g, uint32_t slots, sgm_batch** out)
{
    if (!g || !out || slots == 0 || slots > 4096) return SGM_EINVAL;
    *out = nullptr;
    sgm_ctx* c = g->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    sgm_batch* b = new sgm_batch();
    b->g = g;
    b->c = c;
    b->slots = slots;
    const size_t V = g->V;
    hipError_t e = dalloc(&b->d_vmask, slots * V);
    if (e == hipSuccess) e = dalloc(&b->d_flags, (size_t)slots * SGM_GQL_FLAG_WORDS);
    if (e == hipSuccess) e = dalloc(&b->d_scratch, sgm_gql_scratch_words(g->V) * slots);
    if (e == hipSuccess) e = dalloc(&b->d_qd, slots);
    if (e == hipSuccess) e = dalloc(&b->d_cnt, (size_t)slots * SGM_MAXQ);
    if (e == hipSuccess) e = dalloc(&b->d_degsum, (size_t)slots * SGM_MAXQ);
    if (e == hipSuccess) e = dalloc(&b->d_plans, slots);
    if (e == hipSuccess) e = dalloc(&b->d_eq, slots);
    if (e == hipSuccess) e = dalloc(&b->d_st, slots);
    if (e == hipSuccess) e = dalloc(&b->d_qstop, 3 * (size_t)slots);
    if (e == hipSuccess) e = dalloc(&b->d_split, (size_t)SGM_SPLIT_WORDS * slots);
    if (e == hipSuccess) e = dalloc(&b->d_lreq, kLreqWords * slots);
    if (e == hipSuccess) e = dalloc(&b->d_rec, (size_t)SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS * slots);
    if (e == hipSuccess) e = dalloc(&b->d_lastbits, (size_t)slots * ((V + 31) / 32));
    if (e == hipSuccess) e = dalloc(&b->d_rank_given, (size_t)slots * 2 * CM_ROWS);
    if (e == hipSuccess) e = dalloc(&b->d_rank_out, (size_t)slots * CM_ROWS);
    if (e == hipSuccess) e = dalloc(&b->d_rank_pred, 2 * (size_t)slots);
    if (e == hipSuccess) e = dalloc(&b->d_rank_flag, slots);
    for (auto& ev : b->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->h_got), sizeof(sgm_stats_dev) * slots);
    if (e != hipSuccess) {
        sgm_batch_destroy(b);
        return set_err(c, e == hipErrorOutOfMemory ? SGM_ENOMEM : SGM_EHIP, "batch alloc: %s", hipGetErrorString(e));
    }
    b->hqd.assign(slots, sgm_qdesc());
    for (uint32_t i = 0; i < slots; i++) {
        sgm_qdesc& d = b->hqd[i];
        std::memset(&d, 0, sizeof d);
        d.vmask = b->d_vmask + i * V;
        d.removed = b->d_flags + (size_t)i * SGM_GQL_FLAG_WORDS;
        sgm_gql_layout(d, b->d_scratch + sgm_gql_scratch_words(g->V) * i, g->V);
    }
    b->counts.assign((size_t)slots * SGM_MAXQ, 0);
    b->degsum.assign((size_t)slots * SGM_MAXQ, 0);
    c->batches.push_back(b);
    *out = b;
    return SGM_OK;
}

int sgm_batch_filter(sgm_batch* b, uint32_t nq, const uint32_t* const* qoffs, const uint32_t* const* qadjs,
                     const uint32_t* ns, int kind, int32_t* status, uint32_t* cand_counts, uint64_t* cand_degsum)
{
    return sgm_batch_filter_labeled(b, nq, qoffs, qadjs, nullptr, ns, kind, status, cand_counts, cand_degsum);
}

int sgm_batch_filter_labeled(sgm_batch* b, uint32_t nq, const uint32_t* const* qoffs, const uint32_t* const* qadjs,
                             const uint32_t* const* qlabs, const uint32_t* ns, int kind, int32_t* status,
                             uint32_t* cand_counts, uint64_t* cand_degsum)
{
    if (!b || !qoffs || !qadjs || !ns) return SGM_EINVAL;
    sgm_ctx* c = b->c;
    sgm_graph* g = b->g;
    if (nq == 0 || nq > b->slots) return set_err(c, SGM_EINVAL, "batch of %u queries for %u slots", nq, b->slots);
    if (kind != SGM_FILTER_NLF && kind != SGM_FILTER_GQL) return set_err(c, SGM_EINVAL, "unknown filter kind %d", kind);
    if (b->pending) return set_err(c, SGM_ESTATE, "wait for the batch's enumeration before loading another batch");
    HIPCHK(c, hipSetDevice(c->device));
    b->nq = 0;
    b->q.assign(nq, sgm::QueryCSR());
    bool lab = false;  // some slot runs the labelled filter (then every slot's descriptor says which)
    std::vector<uint32_t> lreq;
    for (uint32_t i = 0; i < b->slots; i++) {
        sgm_qdesc& d = b->hqd[i];
        d.lab_on = 0;
        d.lreq = nullptr;
        if (i < nq) {
            const int rc = load_query(c, qoffs[i], qadjs[i], ns[i], b->q[i], d);
            if (rc) return rc;
            sgm::QueryCSR& q = b->q[i];
            bool ql = false;
            for (uint32_t u = 0; qlabs && qlabs[i] && u < q.n && !ql; u++) ql = qlabs[i][u] != 0;
            if (ql) q.lab.assign(qlabs[i], qlabs[i] + q.n);
            if (ql || g->labelled) {
                if (!lab && !g->has_lab) {
                    const int lrc = build_label_tables(c, g, nullptr);
                    if (lrc) return lrc;
                }
                lab = true;
                d.lab_on = 1;
                for (uint32_t u = 0; u < q.n; u++) d.qlab[u] = q.label(u);
                d.lreq = b->d_lreq + kLreqWords * i;
                std::vector<uint32_t> blk = lreq_block(q);
                blk.resize(kLreqWords, 0u);
                lreq.insert(lreq.end(), blk.begin(), blk.end());
            } else {
                lreq.resize(lreq.size() + kLreqWords, 0u);
            }
        } else {
            d.n = 0;
        }
    }
    const uint32_t slots = b->slots, V = g->V;
    // GQL runs 2n phases for a query of n vertices: one graph per phase count, sized for the
    // batch's largest query (phases past a smaller query's own return at once)
    uint32_t max_n = 0;
    for (uint32_t i = 0; i < nq; i++) max_n = std::max(max_n, b->q[i].n);
    const uint32_t n_phases = kind == SGM_FILTER_GQL ? 2 * max_n : 0;
    if (lab && b->lab_gen != g->lab_gen) {  // label tables rebuilt since capture: recapture
        if (b->gf_nlf[1]) (void)hipGraphExecDestroy(b->gf_nlf[1]);
        b->gf_nlf[1] = nullptr;
        std::vector<std::pair<uint32_t, hipGraphExec_t>> keep;
        for (auto& x : b->gf_gql) {
            if (x.first & 1u) (void)hipGraphExecDestroy(x.second);
            else keep.push_back(x);
        }
        b->gf_gql.swap(keep);
        b->lab_gen = g->lab_gen;
    }
    hipGraphExec_t* gexec = &b->gf_nlf[lab ? 1 : 0];
    if (kind == SGM_FILTER_GQL) {
        const uint32_t key = 2 * n_phases + (lab ? 1u : 0u);
        gexec = nullptr;
        for (auto& x : b->gf_gql)
            if (x.first == key) gexec = &x.second;
        if (!gexec) {
            b->gf_gql.emplace_back(key, nullptr);
            gexec = &b->gf_gql.back().second;
        }
    }
    if (!*gexec) {
        const int rc = capture(c, b->stream, [&] {
            sgm::launch_fill_u32(b->stream, b->d_flags, (size_t)slots * SGM_GQL_FLAG_WORDS, 0);
            if (kind == SGM_FILTER_GQL)  // touch and dirty words of every slot
                sgm::launch_fill_strided(b->stream, b->d_scratch + V, sgm_gql_clear_words(V),
                                         sgm_gql_scratch_words(V), slots, 0);
            sgm::launch_fill_u32(b->stream, b->d_cnt, (size_t)slots * SGM_MAXQ, 0);
            sgm::launch_fill_u32(b->stream, reinterpret_cast<uint32_t*>(b->d_degsum), 2 * (size_t)slots * SGM_MAXQ, 0);
            if (kind == SGM_FILTER_GQL && V) {
                sgm::launch_nlf_rank(b->stream, g->d_gecnt, V, b->d_qd, slots);
                if (lab)
                    sgm::launch_nlf_rank_lab(b->stream, g->d_bydeg, g->d_off, g->d_lab, g->d_run_off, g->d_run_lab,
                                             g->d_run_cnt, V, b->d_qd, slots);
                sgm::launch_gql_prepass(b->stream, g->d_dtab, V, g->d_gecnt, b->d_qd, slots);
                for (uint32_t ph = 0; ph < n_phases; ph++)
                    sgm::launch_gql_phase(b->stream, g->d_roff, g->d_radj, V, g->d_gecnt, b->d_qd, slots, ph);
                sgm::launch_unrank(b->stream, g->d_rank, V, b->d_qd, slots);
            } else {
                sgm::launch_nlf(b->stream, g->d_off, V, b->d_qd, slots);
                if (lab && V)
                    sgm::launch_nlf_lab(b->stream, g->d_off, g->d_lab, g->d_run_off, g->d_run_lab, g->d_run_cnt, V,
                                        b->d_qd, slots);
            }
            sgm::launch_cand_stats(b->stream, g->d_off, V, b->d_qd, slots, b->d_cnt, b->d_degsum);
        }, gexec);
        if (rc) return rc;
    }
    if (lab)
        HIPCHK(c, hipMemcpyAsync(b->d_lreq, lreq.data(), sizeof(uint32_t) * lreq.size(), hipMemcpyHostToDevice,
                                 b->stream));
    HIPCHK(c, hipMemcpyAsync(b->d_qd, b->hqd.data(), sizeof(sgm_qdesc) * slots, hipMemcpyHostToDevice, b->stream));
    HIPCHK(c, hipEventRecord(b->ev[0], b->stream));
    HIPCHK(c, hipGraphLaunch(*gexec, b->stream));
    HIPCHK(c, hipEventRecord(b->ev[1], b->stream));
    HIPCHK(c, hipMemcpyAsync(b->counts.data(), b->d_cnt, sizeof(uint32_t) * SGM_MAXQ * nq, hipMemcpyDeviceToHost,
                             b->stream));
    HIPCHK(c, hipMemcpyAsync(b->degsum.data(), b->d_degsum, sizeof(unsigned long long) * SGM_MAXQ * nq,
                             hipMemcpyDeviceToHost, b->stream));
    HIPCHK(c, hipStreamSynchronize(b->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->filter_ms = ms;
    b->status.assign(nq, SGM_OK);
    for (uint32_t i = 0; i < nq; i++) {
        for (uint32_t u = 0; u < b->q[i].n; u++)
            if (b->counts[(size_t)i * SGM_MAXQ + u] == 0) b->status[i] = SGM_EEMPTY;
        if (status) status[i] = b->status[i];
    }
    if (cand_counts) std::memcpy(cand_counts, b->counts.data(), sizeof(uint32_t) * SGM_MAXQ * nq);
    if (cand_degsum) std::memcpy(cand_degsum, b->degsum.data(), sizeof(uint64_t) * SGM_MAXQ * nq);
    b->nq = nq;
    b->kind = kind;
    return SGM_OK;
}

int sgm_batch_order_gql(const sgm_batch* b, uint32_t i, uint32_t* order)
{
    if (!b || !order || i >= b->nq) return SGM_EINVAL;
    sgm::order_gql(b->q[i], b->counts.data() + (size_t)i * SGM_MAXQ, b->g->V, order);
    return SGM_OK;
}

int sgm_batch_enumerate_launch(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* opts)
{
    if (!b) return SGM_EINVAL;
    sgm_ctx* c = b->c;
    sgm_graph* g = b->g;
    if (b->nq == 0) return set_err(c, SGM_ESTATE, "filter a batch first");
    if (b->pending) return set_err(c, SGM_ESTATE, "the batch's previous enumeration was not waited for");
    const sgm_enum_opts o = normalize_opts(opts);
    if (o.method != SGM_ENUM_CACHE_FILTER_FS && o.method != SGM_ENUM_CACHE_FILTER)
        return set_err(c, SGM_EINVAL, "the batch runner supports the cache + filter methods (got %d)", o.method);
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t nq = b->nq, slots = b->slots, V = g->V;
    std::vector<sgm_plan_dev> plans(nq);
    std::vector<sgm_equery>& eq = b->pend_eq;
    eq.assign(slots, sgm_equery());
    bool split = false;  // some slot is split (the launch then runs the split variant)
    uint32_t max_ops = 1;
    // orders and plans on host threads (errors reported in slot order below)
    std::vector<std::array<uint32_t, SGM_MAXQ>> ords(nq);
    std::vector<int> prcs(nq, 0);
    sgm::parallel_for(nq, 16, [&](uint32_t i) {
        const sgm::QueryCSR& q = b->q[i];
        if (orders && orders[i]) std::copy(orders[i], orders[i] + q.n, ords[i].begin());
        else sgm::order_ri(q, ords[i].data());
        sgm::Plan pl;
        prcs[i] = sgm::build_plan(q, ords[i].data(), pl);
        if (!prcs[i]) sgm::plan_to_device(pl, plans[i]);
    });
    for (uint32_t i = 0; i < slots; i++) {
        sgm_equery& e = eq[i];
        std::memset(&e, 0, sizeof e);
        e.plan = b->d_plans + std::min(i, nq - 1);
        e.vmask = b->d_vmask + (size_t)i * V;
        e.lastbits = b->d_lastbits + (size_t)i * ((V + 31) / 32);
        e.stop = b->d_qstop + i;
        e.next = b->d_qstop + slots + i;
        e.snext = b->d_qstop + 2 * slots + i;
        e.split = nullptr;
        e.rec = nullptr;
        e.st = b->d_st + i;
        e.cap = (unsigned long long)o.cap;
        e.budget = budget_ticks(c, o);
        root_range(o, &e.root_lo, &e.root_hi);
        if (i >= nq) continue;
        const sgm::QueryCSR& q = b->q[i];
        const std::array<uint32_t, SGM_MAXQ>& ord = ords[i];
        const int prc = prcs[i];
        if (prc == -2) return set_err(c, SGM_EORDER, "order of query %u is not a connected permutation", i);
        if (prc) return set_err(c, SGM_EINVAL, "plan construction failed for query %u (%d)", i, prc);
        max_ops = std::max<uint32_t>(max_ops, plans[i].n_ops);
        e.root_u = ord[0];
        if (b->status[i] == SGM_OK) {  // the reference's driver reports 0 without enumerating
            e.n_items = g->ge_cnt[std::min<uint32_t>(q.off[ord[0] + 1] - q.off[ord[0]], 32)];
            const uint32_t roots = b->counts[(size_t)i * SGM_MAXQ + ord[0]];
            if (!(o.flags & SGM_ENUM_F_NO_SPLIT) &&
                ((o.flags & SGM_ENUM_F_SPLIT_ALL) || roots <= SGM_SPLIT_MAX_ROOTS || narrow_roots(o))) {
                e.split = b->d_split + (size_t)SGM_SPLIT_WORDS * i;
                e.rec = b->d_rec + (size_t)SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS * i;
                split = true;
            }
        }
    }
    uint32_t live = 0;
    for (uint32_t i = 0; i < nq; i++) live += eq[i].n_items ? 1u : 0u;
    const uint32_t bucket = ops_bucket(max_ops);
    const uint32_t blocks = default_blocks(c, o, std::max<uint32_t>(live, 1));
    const uint64_t words = o.stack_words ? o.stack_words : std::min<uint64_t>((uint64_t)8 * g->max_deg + 1024, 1u << 20);
    const int use_fs = o.method == SGM_ENUM_CACHE_FILTER_FS ? 1 : 0;
    const int serial = (o.flags & SGM_ENUM_F_SERIAL_LAST) ? 1 : 0;
    const uint32_t* hub = nullptr;
    int rc = hub_slot_map(g, o, &hub);
    if (rc) return rc;
    b->pend_o = o;
    b->pend_bucket = bucket;
    b->pend_blocks = blocks;
    b->pend_words = words;
    b->pend_use_fs = use_fs;
    b->pend_serial = serial;
    b->pend_split = split;
    b->pend_hub = hub;
    if (!live) {  // nothing to enumerate: the wait reports zero counts
        std::memset(b->h_got, 0, sizeof(sgm_stats_dev) * slots);
        b->enum_ms = 0;
        b->pending = true;
        c->enum_inflight++;
        return SGM_OK;
    }
    HIPCHK(c, hipMemcpyAsync(b->d_plans, plans.data(), sizeof(sgm_plan_dev) * nq, hipMemcpyHostToDevice, b->stream));
    HIPCHK(c, hipMemcpyAsync(b->d_eq, eq.data(), sizeof(sgm_equery) * slots, hipMemcpyHostToDevice, b->stream));
    const size_t need = (size_t)blocks * 4 * words;
    if (b->stack_words_total < need) {
        dfree(b->d_stack);
        b->d_stack = nullptr;
        b->stack_words_total = 0;
        HIPCHK(c, dalloc(&b->d_stack, need));
        b->stack_words_total = need;
        for (auto& x : b->ge) (void)hipGraphExecDestroy(x.exec);  // captured with the old stack
        b->ge.clear();
    }
    hipGraphExec_t exec = nullptr;
    for (auto& x : b->ge)
        if (x.max_ops == bucket && x.use_fs == (uint32_t)use_fs && x.blocks == blocks && x.words == words && x.hub == hub &&
            x.serial == (uint32_t)serial && x.split == (uint32_t)split && x.chunks == o.split_chunks)
            exec = x.exec;
    const size_t st_words = sizeof(sgm_stats_dev) / sizeof(uint32_t) * slots;
    if (!exec) {
        rc = capture(c, b->stream, [&] {
            sgm::launch_fill_u32(b->stream, reinterpret_cast<uint32_t*>(b->d_st), st_words, 0);
            sgm::launch_fill_u32(b->stream, b->d_qstop, 3 * (size_t)slots, 0);
            sgm::launch_last_bits(b->stream, b->d_eq, slots, V);
            if (split) sgm::launch_split_roots(b->stream, g->d_off, g->d_bydeg, b->d_eq, slots, V, o.split_chunks);
            sgm::launch_enum_cache(b->stream, blocks, g->d_off, g->d_adj, g->d_bydeg, b->d_eq, slots, bucket, b->d_stack,
                                   (uint32_t)words, c->d_stop, c->d_hstop, use_fs, hub, g->d_hub_bits, g->hub_words,
                                   serial, split);
            if (split) sgm::launch_split_reduce(b->stream, b->d_eq, slots);
        }, &exec);
        if (rc) return rc;
        b->ge.push_back(EnumGraph{bucket, (uint32_t)use_fs, blocks, (uint32_t)words, (uint32_t)serial, (uint32_t)split,
                                  o.split_chunks, hub, exec});
    }
    // the context's stop word is shared by every launch in flight: clear it only when none is
    // (a stop request then ends all of them)
    if (c->enum_inflight == 0) {
        sgm::launch_fill_u32(b->stream, c->d_stop, 1, 0);
        clear_host_stop(c);
    }
    HIPCHK(c, hipEventRecord(b->ev[2], b->stream));
    HIPCHK(c, hipGraphLaunch(exec, b->stream));
    HIPCHK(c, hipEventRecord(b->ev[3], b->stream));
    HIPCHK(c, hipMemcpyAsync(b->h_got, b->d_st, sizeof(sgm_stats_dev) * nq, hipMemcpyDeviceToHost, b->stream));
    b->pending = true;
    c->enum_inflight++;
    b->enum_ms = -1;
    return SGM_OK;
}

int sgm_batch_enumerate_wait(sgm_batch* b, sgm_stats* stats)
{
    if (!b || !stats) return SGM_EINVAL;
    sgm_ctx* c = b->c;
    sgm_graph* g = b->g;
    if (!b->pending) return set_err(c, SGM_ESTATE, "no enumeration launched on this batch");
    HIPCHK(c, hipSetDevice(c->device));
    const hipError_t se = hipStreamSynchronize(b->stream);
    b->pending = false;
    c->enum_inflight--;
    if (se != hipSuccess) return set_err(c, SGM_EHIP, "batch enumeration: %s", hipGetErrorString(se));
    const uint32_t nq = b->nq, slots = b->slots;
    const sgm_enum_opts o = b->pend_o;
    const std::vector<sgm_equery>& eq = b->pend_eq;
    std::memset(stats, 0, sizeof(sgm_stats) * nq);
    float ms = 0;
    if (b->enum_ms < 0) {
        HIPCHK(c, hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
        b->enum_ms = ms;
    }
    std::vector<uint32_t> again;
    for (uint32_t i = 0; i < nq; i++) {
        stats[i].kernel_ms = ms;
        if (b->h_got[i].overflow) { again.push_back(i); continue; }
        stats_out(b->h_got[i], o.cap, &stats[i], c->wall_ns_per_tick);
        if (c->stop_requested && !stats[i].capped && b->status[i] == SGM_OK) stats[i].stopped = 1;
    }
    // set-op stack overflow (rare): rerun those queries alone with a larger stack, outside
    // the graph (this allocates)
    const uint32_t blocks = b->pend_blocks, bucket = b->pend_bucket;
    uint64_t words = b->pend_words;
    const size_t st_words = sizeof(sgm_stats_dev) / sizeof(uint32_t) * slots;
    for (uint32_t attempt = 0; !again.empty(); attempt++) {
        if (attempt >= 3) return set_err(c, SGM_EDEVICE, "set-op stack overflow after 3 retries");
        words *= 4;
        if (words > 0x7FFFFFFFull) return set_err(c, SGM_EDEVICE, "set-op stack would exceed 2^31 words");
        int rc = ensure_stack(c, (size_t)blocks * 4 * words * sizeof(uint32_t));
        if (rc) return rc;
        std::vector<sgm_equery> sub;
        for (uint32_t i : again) {
            sgm_equery e = eq[i];
            const uint32_t k = (uint32_t)sub.size();
            e.stop = b->d_qstop + k;
            e.next = b->d_qstop + slots + k;
            e.snext = b->d_qstop + 2 * slots + k;
            e.st = b->d_st + k;
            sub.push_back(e);
        }
        HIPCHK(c, hipMemcpyAsync(b->d_eq, sub.data(), sizeof(sgm_equery) * sub.size(), hipMemcpyHostToDevice,
                                 b->stream));
        sgm::launch_fill_u32(b->stream, reinterpret_cast<uint32_t*>(b->d_st), st_words, 0);
        sgm::launch_fill_u32(b->stream, b->d_qstop, 3 * (size_t)slots, 0);
        HIPCHK(c, hipEventRecord(b->ev[2], b->stream));
        const uint32_t ns = (uint32_t)sub.size();
        sgm::launch_split_roots(b->stream, g->d_off, g->d_bydeg, b->d_eq, ns, g->V, o.split_chunks);  // no-op without split blocks
        sgm::launch_enum_cache(b->stream, blocks, g->d_off, g->d_adj, g->d_bydeg, b->d_eq, ns, bucket, c->d_stack,
                               (uint32_t)words, c->d_stop, c->d_hstop, b->pend_use_fs, b->pend_hub, g->d_hub_bits,
                               g->hub_words, b->pend_serial, b->pend_split);
        sgm::launch_split_reduce(b->stream, b->d_eq, ns);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(b->ev[3], b->stream));
        HIPCHK(c, hipMemcpyAsync(b->h_got, b->d_st, sizeof(sgm_stats_dev) * sub.size(), hipMemcpyDeviceToHost,
                                 b->stream));
        HIPCHK(c, hipStreamSynchronize(b->stream));
        HIPCHK(c, hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
        b->enum_ms += ms;
        std::vector<uint32_t> next;
        for (size_t k = 0; k < again.size(); k++) {
            const uint32_t i = again[k];
            stats[i].kernel_ms += ms;
            stats[i].retries = attempt + 1;
            if (b->h_got[k].overflow) { next.push_back(i); continue; }
            stats_out(b->h_got[k], o.cap, &stats[i], c->wall_ns_per_tick);
        }
        again.swap(next);
        // the slot layout of d_eq was replaced: restore it for the captured graph
        HIPCHK(c, hipMemcpyAsync(b->d_eq, eq.data(), sizeof(sgm_equery) * slots, hipMemcpyHostToDevice, b->stream));
        HIPCHK(c, hipStreamSynchronize(b->stream));
    }
    if (c->stop_requested)
        for (uint32_t i = 0; i < nq; i++)
            if (stats[i].stopped) return set_err(c, SGM_ESTOPPED, "stopped by sgm_stop");
    return SGM_OK;
}

int sgm_batch_enumerate(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* opts, sgm_stats* stats)
{
    if (!b || !stats) return SGM_EINVAL;
    const int rc = sgm_batch_enumerate_launch(b, orders, opts);
    if (rc) return rc;
    return sgm_batch_enumerate_wait(b, stats);
}

int sgm_batch_rank_orders(sgm_batch* b, sgm_costmodel* m, uint32_t n_samples, uint64_t seed, float margin,
                          uint32_t* orders, float* pred_best, float* pred_ri, int32_t* learned)
{
    if (!b || !m || !orders) return SGM_EINVAL;
    sgm_ctx* c = b->c;
    if (b->nq == 0) return set_err(c, SGM_ESTATE, "filter a batch first");
    if (b->pending) return set_err(c, SGM_ESTATE, "wait for the batch's enumeration first");
    HIPCHK(c, hipSetDevice(c->device));
    const uint32_t nq = b->nq, K = 2 + n_samples, S = nq * K;
    // RI and GQL orders on the host (cheap, and the GQL order needs the host's counts);
    // the sampled candidates, features, scores and the choice are built on the device
    std::vector<uint8_t> given((size_t)nq * 2 * CM_ROWS, 0);
    std::vector<uint32_t> o(SGM_MAXQ);
    for (uint32_t i = 0; i < nq; i++) {
        const sgm::QueryCSR& q = b->q[i];
        sgm::order_ri(q, o.data());
        for (uint32_t j = 0; j < q.n; j++) given[((size_t)i * 2) * CM_ROWS + j] = (uint8_t)o[j];
        sgm::order_gql(q, b->counts.data() + (size_t)i * SGM_MAXQ, b->g->V, o.data());
        for (uint32_t j = 0; j < q.n; j++) given[((size_t)i * 2 + 1) * CM_ROWS + j] = (uint8_t)o[j];
    }
    if (b->rank_cap < S) {
        dfree(b->d_rank_orders);
        b->d_rank_orders = nullptr;
        b->rank_cap = 0;
        HIPCHK(c, dalloc(&b->d_rank_orders, (size_t)S * CM_ROWS));
        b->rank_cap = S;
    }
    uint16_t* F;
    uint8_t* valid;
    uint32_t *back, *fwd;
    int rc = costmodel_work_inputs(m, S, &F, &valid, &back, &fwd);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(b->d_rank_given, given.data(), given.size(), hipMemcpyHostToDevice, b->stream));
    sgm::cm_rank_inputs(b->stream, b->d_qd, b->d_cnt, b->d_degsum, b->d_rank_given, nq, K, seed, b->d_rank_orders, F,
                        valid, back, fwd);
    const float* y = nullptr;
    rc = costmodel_forward_work(m, S, &y, b->stream);  // same stream as the inputs and the choice
    if (rc) return rc;
    sgm::cm_choose(b->stream, y, b->d_rank_orders, nq, K, margin, b->d_rank_out, b->d_rank_pred,
                   b->d_rank_pred + b->slots, b->d_rank_flag);
    HIPCHK(c, hipGetLastError());
    std::vector<uint32_t> out((size_t)nq * CM_ROWS);
    std::vector<float> pb(nq), pr(nq);
    std::vector<int32_t> fl(nq);
    HIPCHK(c, hipMemcpyAsync(out.data(), b->d_rank_out, sizeof(uint32_t) * out.size(), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(c, hipMemcpyAsync(pb.data(), b->d_rank_pred, sizeof(float) * nq, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(c, hipMemcpyAsync(pr.data(), b->d_rank_pred + b->slots, sizeof(float) * nq, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(c, hipMemcpyAsync(fl.data(), b->d_rank_flag, sizeof(int32_t) * nq, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(c, hipStreamSynchronize(b->stream));
    for (uint32_t i = 0; i < nq; i++) {
        std::copy(out.begin() + (size_t)i * CM_ROWS, out.begin() + (size_t)i * CM_ROWS + SGM_MAXQ,
                  orders + (size_t)i * SGM_MAXQ);
        if (pred_best) pred_best[i] = pb[i];
        if (pred_ri) pred_ri[i] = pr[i];
        if (learned) learned[i] = fl[i];
    }
    return SGM_OK;
}

int sgm_batch_last_ms(const sgm_batch* b, double* filter_ms, double* enum_ms)
{
    if (!b) return SGM_EINVAL;
    if (filter_ms) *filter_ms = b->filter_ms;
    if (enum_ms) *enum_ms = b->enum_ms;
    return SGM_OK;
}

int sgm_batch_autotune(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* base, sgm_enum_opts* best,
                       double* best_ms)
{
    if (!b || !best) return SGM_EINVAL;
    sgm_ctx* c = b->c;
    if (b->nq == 0) return set_err(c, SGM_ESTATE, "filter a batch first");
    if (b->pending) return set_err(c, SGM_ESTATE, "wait for the batch's enumeration first");
    const sgm_enum_opts o0 = normalize_opts(base);
    std::vector<sgm_stats> st(b->nq), ref;
    double best_t = 1e300;
    sgm_enum_opts winner = o0;
    auto trial = [&](const sgm_enum_opts& o) -> int {
        const int rc = sgm_batch_enumerate(b, orders, &o, st.data());
        if (rc) return rc;
        if (ref.empty()) {
            ref = st;
        } else {
            for (uint32_t i = 0; i < b->nq; i++)
                if (st[i].count != ref[i].count)
                    return set_err(c, SGM_EDEVICE, "autotune: counts differ between settings (query %u)", i);
        }
        if (b->enum_ms < best_t) { best_t = b->enum_ms; winner = o; }
        return SGM_OK;
    };
    // 1. persistent grid (blocks of 4 waves per CU) x intersection strategy (adjacency bitmaps
    // for lists of at least T vertices; the rest by binary search)
    static const uint32_t per_cu[] = {2, 4, 8};
    static const uint32_t hub_min[] = {64, 256, 1024, 0};  // 0: no bitmaps
    for (uint32_t pc : per_cu)
        for (uint32_t hm : hub_min) {
            sgm_enum_opts o = o0;
            o.blocks = (uint32_t)c->num_cus * pc;
            o.hub_min_degree = hm;
            o.flags = hm ? (o0.flags & ~SGM_ENUM_F_NO_BITMAPS) : (o0.flags | SGM_ENUM_F_NO_BITMAPS);
            if (const int rc = trial(o)) return rc;
        }
    // 2. from the best of those, the work-chunk size: how many chunks the depth-1 loop of each
    // of a query's heaviest roots is cut into (smaller chunks spread a heavy root over more
    // waves, larger ones cost fewer records)
    {
        const sgm_enum_opts w = winner;
        for (uint32_t ch : {8u, 16u}) {
            sgm_enum_opts o = w;
            o.split_chunks = ch;
            if (const int rc = trial(o)) return rc;
        }
    }
    // 3. the per-wave set-op stack: half and twice the automatic size (a stack that overflows
    // re-runs its queries with a larger one, which the timing charges)
    {
        const sgm_enum_opts w = winner;
        const uint32_t auto_words =
            (uint32_t)std::min<uint64_t>((uint64_t)8 * b->g->max_deg + 1024, 1u << 20);
        for (uint32_t words : {auto_words / 2, auto_words * 2}) {
            sgm_enum_opts o = w;
            o.stack_words = std::max<uint32_t>(words, 1024);
            if (const int rc = trial(o)) return rc;
        }
    }
    *best = winner;
    if (best_ms) *best_ms = best_t;
    return SGM_OK;
}

}  // extern "C"
