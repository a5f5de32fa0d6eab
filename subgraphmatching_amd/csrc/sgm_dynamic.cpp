// The dynamic-order searches (explore_Intersection_DynamicOrder_DAF / _VEQ,
// inc/explore.h:1674-2162) and the DAF weight array they rank with (computeWeightArray,
// inc/order.h:978-1031): host side.  The weights are computed on the device (k_weight_step,
// one launch per reverse-BFS position for a whole batch) and stay resident with the query for
// the enumeration; the search itself is k_enum_dyn.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "sgm_objects.h"

namespace {

// Previous / following neighbours of every query vertex in order-position order
// (GetPreviousNeb / GetFollowingNeb, inc/common_ops.h:259-313).
void neighbours_in_order(const sgm::QueryCSR& q, const uint32_t* order, std::vector<std::vector<uint32_t>>& prev,
                         std::vector<std::vector<uint32_t>>& fol)
{
    const uint32_t n = q.n;
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; i++) pos[order[i]] = i;
    prev.assign(n, {});
    fol.assign(n, {});
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
            const uint32_t w = q.adj[p];
            (pos[w] < pos[u] ? prev[u] : fol[u]).push_back(w);
        }
        auto by_pos = [&](uint32_t a, uint32_t b) { return pos[a] < pos[b]; };
        std::sort(prev[u].begin(), prev[u].end(), by_pos);
        std::sort(fol[u].begin(), fol[u].end(), by_pos);
    }
}

bool tables_built(const sgm_query* q)
{
    const uint32_t n = q->q.n;
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++)
            if (!q->tables[(size_t)a * n + q->q.adj[p]].built) return false;
    return true;
}

int ensure_tables(sgm_query* q)
{
    if (q->empty || tables_built(q)) return SGM_OK;
    return build_tables_many(&q, 1);
}

int ensure_weight_buffer(sgm_query* q)
{
    sgm_ctx* c = q->g->ctx;
    const size_t need = std::max<size_t>(q->cand_off[q->q.n], 1);
    if (q->weight_cap >= need) return SGM_OK;
    dfree(q->d_weight);
    q->d_weight = nullptr;
    q->weight_cap = 0;
    HIPCHK(c, dalloc(&q->d_weight, need));
    q->weight_cap = need;
    return SGM_OK;
}

bool same_order(const sgm_query* q, const uint32_t* order)
{
    return q->weights_ok && std::equal(order, order + q->q.n, q->weight_order);
}

// computeWeightArray (inc/order.h:997-1029) on the device, for many queries and their start
// orders at once: from the last vertex of each order back, a vertex's weight per candidate is
// the minimum, over its tree children whose only previous neighbour it is, of the child
// weights summed along the candidate's table row; 1 everywhere when it has no such child;
// weights start at UINT32_MAX (src/order_ablations.cpp:134-138).  The BFS-tree children of x
// with one previous neighbour are exactly the following neighbours of x whose one previous
// neighbour is x, so the rule reads the order alone.  Launches: one fill, then one per
// reverse position, each over every query's (vertex, child) passes at that position.  The
// queries' tables must be built.
int compute_weights_many(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq)
{
    if (!nq) return SGM_OK;
    sgm_ctx* c = qs[0]->g->ctx;
    std::vector<std::vector<sgm_wdesc>> by_step(SGM_MAXQ + 1);
    std::vector<uint32_t> max_rows(SGM_MAXQ + 1, 0);
    auto add = [&](uint32_t k, const sgm_wdesc& d) {
        by_step[k].push_back(d);
        max_rows[k] = std::max(max_rows[k], d.rows);
    };
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        const uint32_t* order = orders[i];
        const uint32_t n = q->q.n;
        if (const int rc = ensure_weight_buffer(q)) return rc;
        add(0, sgm_wdesc{nullptr, nullptr, nullptr, nullptr, q->d_weight, q->cand_off[n], 0xFFFFFFFFu});
        std::vector<std::vector<uint32_t>> prev, fol;
        neighbours_in_order(q->q, order, prev, fol);
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t x = order[n - 1 - k];
            uint32_t* out = q->d_weight + q->cand_off[x];
            bool one = true;
            for (uint32_t ch : fol[x]) {
                if (prev[ch].size() != 1) continue;
                one = false;
                const sgm_table& t = q->tables[(size_t)x * n + ch];
                add(k + 1, sgm_wdesc{t.d_voff, t.d_pool, q->d_list + q->cand_off[x], q->d_weight + q->cand_off[ch], out,
                                     q->counts[x], 0u});
            }
            if (one) add(k + 1, sgm_wdesc{nullptr, nullptr, nullptr, nullptr, out, q->counts[x], 1u});
        }
    }
    std::vector<sgm_wdesc> all;
    std::vector<uint32_t> first(by_step.size() + 1, 0);
    for (size_t k = 0; k < by_step.size(); k++) {
        first[k] = (uint32_t)all.size();
        all.insert(all.end(), by_step[k].begin(), by_step[k].end());
    }
    first[by_step.size()] = (uint32_t)all.size();
    if (const int rc = grow(c, &c->d_wdesc, c->wdesc_cap, (uint32_t)all.size())) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_wdesc, all.data(), sizeof(sgm_wdesc) * all.size(), hipMemcpyHostToDevice, c->stream));
    for (size_t k = 0; k < by_step.size(); k++)
        sgm::launch_weight_step(c->stream, c->d_wdesc + first[k], first[k + 1] - first[k], max_rows[k]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));  // `all` is the copy's source
    for (uint32_t i = 0; i < nq; i++) {
        std::copy(orders[i], orders[i] + qs[i]->q.n, qs[i]->weight_order);
        qs[i]->weights_ok = true;
    }
    return SGM_OK;
}

// the reference driver's start order: BFS from DAFStartVertex (inc/filter.h:922-940,
// inc/common_ops.h:318-350; src/order_ablations.cpp:158-160, inc/order.h:984-987)
int daf_bfs_order(sgm_query* q, uint32_t* order)
{
    uint32_t nlf[SGM_MAXQ];
    const int rc = query_nlf_counts(q, nlf);
    if (rc) return rc;
    sgm::BfsTree t;
    sgm::bfs_tree(q->q, sgm::ceci_daf_start_vertex(q->q, nlf, q->g->V, true), t);
    if (t.order.size() != q->q.n) return set_err(q->g->ctx, SGM_EINVAL, "the query is not connected");
    std::copy(t.order.begin(), t.order.end(), order);
    return SGM_OK;
}

int dyn_checks(sgm_query* q, const char* what)
{
    if (const int rc = check_filtered(q, what)) return rc;
    if (q->q.n < 2) return set_err(q->g->ctx, SGM_EINVAL, "%s: the query needs 2 or more vertices", what);
    return SGM_OK;
}

}  // namespace

// Both dynamic-order searches for many queries of one graph in one launch; orders[i] is
// query i's start order (connected permutations, checked by the caller).
extern "C++" int enumerate_dyn_batch(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq,
                                     const sgm_enum_opts& o, sgm_stats* out)
{
    sgm_graph* g = qs[0]->g;
    sgm_ctx* c = g->ctx;
    const bool daf = o.method == SGM_ENUM_DYN_DAF;
    std::vector<sgm_query*> missing;
    for (uint32_t i = 0; i < nq; i++)
        if (!qs[i]->empty && !tables_built(qs[i])) missing.push_back(qs[i]);
    if (!missing.empty()) {
        const int trc = build_tables_many(missing.data(), (uint32_t)missing.size());
        if (trc) return trc;
    }
    if (daf) {  // the weights of every query without resident ones for its order, in one pass
        std::vector<sgm_query*> wq;
        std::vector<const uint32_t*> wo;
        for (uint32_t i = 0; i < nq; i++)
            if (!qs[i]->empty && !same_order(qs[i], orders[i])) {
                wq.push_back(qs[i]);
                wo.push_back(orders[i]);
            }
        if (const int rc = compute_weights_many(wq.data(), wo.data(), (uint32_t)wq.size())) return rc;
    }
    std::vector<sgm_dynplan_dev> plans(nq);
    uint64_t words_need = 1024;  // stack words that hold every intersection along one search path
    for (uint32_t i = 0; i < nq; i++) {
        sgm_query* q = qs[i];
        const uint32_t* order = orders[i];
        const uint32_t n = q->q.n;
        if (q->empty) continue;
        std::vector<std::vector<uint32_t>> prev, fol;
        neighbours_in_order(q->q, order, prev, fol);
        sgm_dynplan_dev& P = plans[i];
        std::memset(&P, 0, sizeof P);
        P.n = n;
        P.root = order[0];
        uint32_t e = 0, f = 0;
        uint64_t path_words = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (q->q.off[u + 1] - q->q.off[u] == 1) P.deg1 |= 1u << u;
            P.prev_first[u] = e;
            // intersection k (k >= 1) holds at most the shortest longest-row of tables 0..k
            uint32_t shortest = 0xFFFFFFFFu;
            for (uint32_t p : prev[u]) {
                const sgm_table& t = q->tables[(size_t)p * n + u];
                P.prev_vtx[e] = (uint8_t)p;
                P.tab_off[e] = t.d_voff;
                P.tab_idx[e] = t.d_pool;
                shortest = std::min(shortest, t.max_deg);
                if (e > P.prev_first[u]) path_words += shortest;
                e++;
            }
            P.fol_first[u] = f;
            for (uint32_t w : fol[u]) P.fol[f++] = (uint8_t)w;
            P.cand[u] = q->d_list + q->cand_off[u];
            P.weight[u] = daf ? q->d_weight + q->cand_off[u] : nullptr;
        }
        P.prev_first[n] = e;
        P.fol_first[n] = f;
        words_need = std::max(words_need, path_words + 64);
    }
    uint32_t rlo, rhi;
    root_range(o, &rlo, &rhi);
    const uint32_t* hub_slot = nullptr;  // adjacency bitmaps of the hubs, for row probes
    if (const int hrc = hub_slot_map(g, o, &hub_slot)) return hrc;
    {  // each query's roots in degree order
        std::vector<uint32_t> us(nq);
        for (uint32_t i = 0; i < nq; i++) us[i] = plans[i].root;
        std::vector<const uint32_t*> ro;
        if (const int rrc = root_orders(c, qs, us.data(), nq, ro)) return rrc;
        for (uint32_t i = 0; i < nq; i++) plans[i].root_order = ro[i];
    }
    std::vector<uint32_t> live;
    for (uint32_t i = 0; i < nq; i++)
        if (!qs[i]->empty) live.push_back(i);
    const uint32_t blocks = default_blocks(c, o, (uint32_t)std::max<size_t>(live.size(), 1));
    const uint32_t waves = blocks * 4;
    // The bound above cannot overflow; it is used while the stacks of all waves stay within
    // 2 GiB, else the aux search's default with re-runs on overflow.
    uint64_t words = o.stack_words;
    if (!words) {
        words = words_need;
        if (words * waves * 4 > (2ull << 30)) words = std::min<uint64_t>(words, (uint64_t)8 * g->max_deg + 1024);
    }
    for (uint32_t attempt = 0; !live.empty(); attempt++) {
        if (attempt > 3) return set_err(c, SGM_EDEVICE, "set-op stack overflow after %u retries", attempt - 1);
        if (words > 0x7FFFFFFFull) return set_err(c, SGM_EDEVICE, "set-op stack would exceed 2^31 words");
        int rc = ensure_stack(c, (size_t)waves * words * sizeof(uint32_t));
        if (rc) return rc;
        const uint32_t m = (uint32_t)live.size();
        if ((rc = grow(c, &c->d_dynplans, c->dyn_cap, m))) return rc;
        if ((rc = grow(c, &c->d_afirst, c->afirst_cap, m + 1))) return rc;
        if ((rc = grow(c, &c->d_astats, c->astats_cap, m))) return rc;
        if ((rc = grow(c, &c->d_astop, c->astop_cap, m))) return rc;
        if ((rc = grow(c, &c->d_anext, c->anext_cap, m))) return rc;
        std::vector<sgm_dynplan_dev> sub(m);
        std::vector<uint32_t> first(m + 1, 0);
        for (uint32_t k = 0; k < m; k++) {
            const uint32_t i = live[k];
            sub[k] = plans[i];
            first[k + 1] = first[k] + qs[i]->counts[plans[i].root];
        }
        HIPCHK(c, hipMemcpyAsync(c->d_dynplans, sub.data(), sizeof(sgm_dynplan_dev) * m, hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_afirst, first.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_astats, 0, sizeof(sgm_stats_dev) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_astop, 0, sizeof(uint32_t) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_anext, 0, sizeof(uint32_t) * m, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_stop, 0, sizeof(uint32_t), c->stream));
        clear_host_stop(c);
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        sgm::launch_enum_dyn(c->stream, blocks, c->d_dynplans, c->d_afirst, m, c->d_stack, (uint32_t)words, c->d_anext,
                             c->d_astats, c->d_astop, c->d_stop, c->d_hstop, (unsigned long long)o.cap,
                             budget_ticks(c, o), rlo, rhi, hub_slot, g->d_hub_bits, g->hub_words);
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

// The default start order of the dynamic searches (orders == NULL).
extern "C++" int dyn_default_order(sgm_query* q, uint32_t* order) { return daf_bfs_order(q, order); }

extern "C" {

int sgm_daf_start_vertex(sgm_query* q, uint32_t* start)
{
    if (!q || !start) return SGM_EINVAL;
    uint32_t nlf[SGM_MAXQ];
    const int rc = query_nlf_counts(q, nlf);
    if (rc) return rc;
    *start = sgm::ceci_daf_start_vertex(q->q, nlf, q->g->V, true);
    return SGM_OK;
}

int sgm_daf_weights(sgm_query* q, uint32_t* bfs_order, uint32_t* const* weights)
{
    if (!q || !bfs_order) return SGM_EINVAL;
    int rc;
    if ((rc = dyn_checks(q, "computeWeightArray"))) return rc;
    sgm_ctx* c = q->g->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = daf_bfs_order(q, bfs_order))) return rc;
    if (q->empty) return set_err(c, SGM_EEMPTY, "computeWeightArray: some candidate set is empty");
    const uint32_t* bo = bfs_order;
    if ((rc = ensure_tables(q)) || (rc = compute_weights_many(&q, &bo, 1))) return rc;
    if (weights) {
        for (uint32_t u = 0; u < q->q.n; u++)
            if (q->counts[u] && !weights[u]) return set_err(c, SGM_EINVAL, "weights[%u] is null", u);
        std::vector<uint32_t> h(q->cand_off[q->q.n]);
        HIPCHK(c, hipMemcpyAsync(h.data(), q->d_weight, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (uint32_t u = 0; u < q->q.n; u++)
            std::copy(h.begin() + q->cand_off[u], h.begin() + q->cand_off[u + 1], weights[u]);
    }
    return SGM_OK;
}

int sgm_set_daf_weights(sgm_query* q, const uint32_t* order, const uint32_t* const* weights)
{
    if (!q || !order || !weights) return SGM_EINVAL;
    int rc;
    if ((rc = dyn_checks(q, "set weights"))) return rc;
    sgm_ctx* c = q->g->ctx;
    if (sgm::check_order(q->q, order)) return set_err(c, SGM_EORDER, "order is not a connected permutation");
    if (q->empty) return set_err(c, SGM_EEMPTY, "set weights: some candidate set is empty");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = ensure_tables(q)) || (rc = ensure_weight_buffer(q))) return rc;
    std::vector<uint32_t> h(q->cand_off[q->q.n]);
    for (uint32_t u = 0; u < q->q.n; u++) {
        if (q->counts[u] && !weights[u]) return set_err(c, SGM_EINVAL, "weights[%u] is null", u);
        std::copy(weights[u], weights[u] + q->counts[u], h.begin() + q->cand_off[u]);
    }
    HIPCHK(c, hipMemcpyAsync(q->d_weight, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::copy(order, order + q->q.n, q->weight_order);
    q->weights_ok = true;
    return SGM_OK;
}

}  // extern "C"
