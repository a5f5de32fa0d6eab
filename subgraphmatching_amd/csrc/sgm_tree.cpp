// Host-side planning of the BFS-tree filters (CFL, CECI, DAF) and of the QSI, CFL, CECI and
// GuP orders: query k-core, BFS tree, previous / following neighbours, start vertices, the
// per-query refinement programs run by k_refine, and the order rules.  All of it is
// O(n^2..n^3) over queries of at most 32 vertices; the data-graph work (NLF counts, aux
// tables, path estimates) is done on the device by the callers.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <tuple>
#include <utility>
#include <vector>

#include "sgm_host.h"
#include "sgm_internal.h"

// double arithmetic below must round exactly as the reference's does (no fused multiply-add)
#pragma clang fp contract(off)

namespace sgm {

static uint32_t deg_of(const QueryCSR& q, uint32_t u) { return q.off[u + 1] - q.off[u]; }

// Query::KCore (inc/graph.h:387-458): bin-sorted peeling; returns #vertices with core > 1.
uint32_t query_kcore(const QueryCSR& q, int32_t* core)
{
    const uint32_t n = q.n;
    int32_t md = 0;
    for (uint32_t v = 0; v < n; v++) md = std::max(md, (int32_t)deg_of(q, v));
    std::vector<int32_t> vert(n), pos(n), bin(md + 1, 0), start(md + 1, 0);
    for (uint32_t v = 0; v < n; v++) {
        core[v] = (int32_t)deg_of(q, v);
        bin[core[v]]++;
    }
    for (int32_t d = 0, s = 0; d <= md; d++) {
        start[d] = s;
        s += bin[d];
    }
    for (uint32_t v = 0; v < n; v++) {
        pos[v] = start[core[v]];
        vert[pos[v]] = (int32_t)v;
        start[core[v]]++;
    }
    for (int32_t d = md; d > 0; d--) start[d] = start[d - 1];
    start[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        const int32_t v = vert[i];
        for (uint32_t p = q.off[v]; p < q.off[v + 1]; p++) {
            const int32_t u = (int32_t)q.adj[p];
            if (core[u] > core[v]) {
                const int32_t du = core[u], pu = pos[u], pw = start[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    pos[w] = pu;
                    vert[pu] = w;
                    vert[pw] = u;
                }
                start[du]++;
                core[u]--;
            }
        }
    }
    uint32_t len = 0;
    for (uint32_t v = 0; v < n; v++) len += core[v] > 1;
    return len;
}

// BFS (inc/common_ops.h:318-350): children in adjacency order, the start's parent 0.
void bfs_tree(const QueryCSR& q, uint32_t s, BfsTree& t)
{
    t = BfsTree();
    t.order.push_back(s);
    uint32_t seen = 1u << s;
    for (size_t head = 0; head < t.order.size(); head++) {
        const uint32_t u = t.order[head];
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
            const uint32_t w = q.adj[p];
            if ((seen >> w) & 1u) continue;
            seen |= 1u << w;
            t.order.push_back(w);
            t.parent[w] = u;
            t.level[w] = t.level[u] + 1;
            t.children[u].push_back(w);
        }
    }
}

// GetPreviousNeb / GetFollowingNeb (inc/common_ops.h:259-316): neighbours earlier / later in
// the order, by position.
void order_neighbours(const QueryCSR& q, const std::vector<uint32_t>& order, bool following,
                      std::vector<uint32_t> (&nb)[SGM_MAXQ])
{
    uint32_t pos[SGM_MAXQ];
    for (uint32_t i = 0; i < order.size(); i++) pos[order[i]] = i;
    for (uint32_t v = 0; v < q.n; v++) {
        nb[v].clear();
        uint32_t m = 0;
        for (uint32_t p = q.off[v]; p < q.off[v + 1]; p++) m |= 1u << q.adj[p];
        for (uint32_t i = 0; i < order.size(); i++) {
            const uint32_t w = order[i];
            if (((m >> w) & 1u) && (following ? i > pos[v] : i < pos[v])) nb[v].push_back(w);
        }
    }
}

// CFLStartVertex (inc/filter.h:437-478): the three vertices of smallest label frequency /
// degree among those with core > 1 (all when there is no 2-core), kept by the reference's
// own max-heap (its tie order is libstdc++'s), then the first in pop order of minimum
// NLF count / degree.
uint32_t cfl_start_vertex(const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls)
{
    int32_t core[SGM_MAXQ];
    const uint32_t klen = query_kcore(q, core);
    auto cmp = [](std::pair<uint32_t, double> l, std::pair<uint32_t, double> r) { return l.second < r.second; };
    std::priority_queue<std::pair<uint32_t, double>, std::vector<std::pair<uint32_t, double>>, decltype(cmp)> h(cmp);
    for (uint32_t u = 0; u < q.n; u++)
        if (klen == 0 || core[u] > 1) h.push(std::make_pair(u, ls.freq(q.label(u)) / (double)deg_of(q, u)));
    while (h.size() > 3) h.pop();
    uint32_t start = 0;
    double best = (double)(uint32_t)(ls.max_freq + 1u);
    while (!h.empty()) {
        const uint32_t u = h.top().first;
        const double s = nlf_cnt[u] / (double)deg_of(q, u);
        if (s < best) {
            start = u;
            best = s;
        }
        h.pop();
    }
    return start;
}

// CECIStartVertex (inc/filter.h:585-600) / DAFStartVertex (:922-940): minimum NLF count /
// degree, first on ties; DAF skips vertices of degree <= 1.
uint32_t ceci_daf_start_vertex(const QueryCSR& q, const uint32_t* nlf_cnt, uint32_t V, bool daf)
{
    double best = V;
    uint32_t start = 0;
    for (uint32_t u = 0; u < q.n; u++) {
        const uint32_t d = deg_of(q, u);
        if (daf && d <= 1) continue;
        const double s = nlf_cnt[u] / (double)d;
        if (s < best) {
            best = s;
            start = u;
        }
    }
    return start;
}

static uint32_t mask_of(const std::vector<uint32_t>& v)
{
    uint32_t m = 0;
    for (uint32_t x : v) m |= 1u << x;
    return m;
}

// The refinement program of one query (k_refine).  Every filter starts from the NLF words
// in vmask, so generating C(u) (generateCandidates) is a prune of u's still-untouched NLF bit
// by u's previous neighbours, whose sets are final for that step.
//   CFL (inc/filter.h:481-583): per BFS level, generate each vertex, then prune the level in
//     reverse by same-level following neighbours; then bottom-up, prune by deeper following
//     neighbours.
//   CECI (:602-788): per vertex u in BFS order, drop parent candidates with no NLF(u)
//     neighbour and generate u from its parent; then drop every candidate of x with no
//     NLF(u) neighbour for some following neighbour u (the non-tree edges' invalidation,
//     order-free because it tests NLF sets only).  The reverse-BFS refinement (:790-883)
//     changes no candidate set.
//   DAF (:942-1020): forward by previous neighbours, backward by following neighbours,
//     forward again.
// Returns the step count, or -1 for a disconnected query.
int refine_program(int kind, const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls, uint32_t* prog)
{
    const uint32_t n = q.n;
    uint32_t start;
    if (kind == SGM_TREE_CFL) start = cfl_start_vertex(q, nlf_cnt, ls);
    else start = ceci_daf_start_vertex(q, nlf_cnt, ls.V, kind == SGM_TREE_DAF);
    BfsTree t;
    bfs_tree(q, start, t);
    if (t.order.size() != n) return -1;
    std::vector<uint32_t> prev[SGM_MAXQ], fol[SGM_MAXQ];
    order_neighbours(q, t.order, false, prev);
    order_neighbours(q, t.order, true, fol);
    uint32_t k = 0;
    auto step = [&](uint32_t bit, uint32_t src, uint32_t piv) {
        if (!piv || k >= SGM_RPROG_MAXSTEPS) return;
        prog[1 + 2 * k] = bit | (src << 8);
        prog[2 + 2 * k] = piv;
        k++;
    };
    if (kind == SGM_TREE_DAF) {
        for (uint32_t pass = 0; pass < 3; pass++) {
            if (pass % 2 == 0)
                for (uint32_t i = 1; i < n; i++) step(t.order[i], 0, mask_of(prev[t.order[i]]));
            else
                for (int i = (int)n - 2; i >= 0; i--) step(t.order[i], 0, mask_of(fol[t.order[i]]));
        }
    } else if (kind == SGM_TREE_CFL) {
        uint32_t maxl = 0;
        for (uint32_t v = 0; v < n; v++) maxl = std::max(maxl, t.level[v]);
        std::vector<uint32_t> at[SGM_MAXQ];
        for (uint32_t v : t.order) at[t.level[v]].push_back(v);
        for (uint32_t l = 1; l <= maxl; l++) {
            for (uint32_t u : at[l]) step(u, 0, mask_of(prev[u]));
            for (size_t i = at[l].size(); i-- > 0;) {
                const uint32_t u = at[l][i];
                uint32_t same = 0;
                for (uint32_t w : fol[u])
                    if (t.level[w] == l) same |= 1u << w;
                step(u, 0, same);
            }
        }
        for (int l = (int)maxl - 1; l >= 0; l--)
            for (uint32_t u : at[l]) {
                uint32_t deeper = 0;
                for (uint32_t w : fol[u])
                    if (t.level[w] > (uint32_t)l) deeper |= 1u << w;
                step(u, 0, deeper);
            }
    } else {
        for (uint32_t i = 1; i < n; i++) {
            const uint32_t u = t.order[i], p = t.parent[u];
            step(p, 1, 1u << u);
            step(u, 0, 1u << p);
        }
        for (uint32_t x : t.order) step(x, 1, mask_of(fol[x]));
    }
    prog[0] = k;
    return (int)k;
}

// Order_QSI (inc/order.h:610-668, QSIStartEdge :146-179).  ecnt[a*n+b] = Edges::_e_cnt of the
// table (a, b), a u32 as in the reference.
void order_qsi(const QueryCSR& q, const uint32_t* ecnt, uint32_t* order)
{
    const uint32_t n = q.n;
    uint32_t vis = 0, minv = std::numeric_limits<uint32_t>::max(), mind = minv, s0 = 0, s1 = 1;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t p = q.off[i]; p < q.off[i + 1]; p++) {
            const uint32_t e = q.adj[p], cur = ecnt[i * n + e], ds = deg_of(q, i) + deg_of(q, e);
            if (cur < minv || (cur == minv && ds < mind)) {
                minv = cur;
                mind = ds;
                if (deg_of(q, i) < deg_of(q, e)) {
                    s0 = e;
                    s1 = i;
                } else {
                    s0 = i;
                    s1 = e;
                }
            }
        }
    order[0] = s0;
    order[1] = s1;
    vis = (1u << s0) | (1u << s1);
    for (uint32_t l = 2; l < n; l++) {
        uint32_t mv = std::numeric_limits<uint32_t>::max(), maxd = 0, maxadj = 0, sel = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!((vis >> i) & 1u)) continue;
            for (uint32_t p = q.off[i]; p < q.off[i + 1]; p++) {
                const uint32_t e = q.adj[p];
                if ((vis >> e) & 1u) continue;
                const uint32_t cur = ecnt[i * n + e], cd = deg_of(q, e);
                uint32_t adjs = 0;
                for (uint32_t r = q.off[e]; r < q.off[e + 1]; r++) adjs += (vis >> q.adj[r]) & 1u;
                if (cur < mv || (cur == mv && adjs < maxadj) || (cur == mv && adjs == maxadj && cd > maxd)) {
                    sel = e;
                    mv = cur;
                    maxd = cd;
                    maxadj = adjs;
                }
            }
        }
        order[l] = sel;
        vis |= 1u << sel;
    }
}

// Order_CFL (inc/order.h:781-976).
int order_cfl(const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls, const uint32_t* cnt,
              const PathEstimator& estimate, uint32_t* order)
{
    const uint32_t n = q.n;
    int32_t core[SGM_MAXQ];
    query_kcore(q, core);
    BfsTree t;
    bfs_tree(q, cfl_start_vertex(q, nlf_cnt, ls), t);
    if (t.order.size() != n) return -1;
    const uint32_t root = t.order[0];
    std::vector<std::vector<uint32_t>> cps;
    std::vector<std::vector<std::vector<uint32_t>>> forests;
    std::vector<uint32_t> leaves, cur;
    for (uint32_t u = 0; u < n; u++)
        if (deg_of(q, u) == 1) leaves.push_back(u);
    // generateCorePaths (:252-269) and generateTreePaths (:271-289)
    std::function<void(uint32_t, std::vector<std::vector<uint32_t>>&)> core_rec =
        [&](uint32_t u, std::vector<std::vector<uint32_t>>& out) {
            cur.push_back(u);
            bool leaf = true;
            for (uint32_t c : t.children[u])
                if (core[c] > 1) {
                    core_rec(c, out);
                    leaf = false;
                }
            if (leaf) out.push_back(cur);
            cur.pop_back();
        };
    std::function<void(uint32_t, std::vector<std::vector<uint32_t>>&)> tree_rec =
        [&](uint32_t u, std::vector<std::vector<uint32_t>>& out) {
            cur.push_back(u);
            bool leaf = true;
            for (uint32_t c : t.children[u])
                if (deg_of(q, c) > 1) {
                    tree_rec(c, out);
                    leaf = false;
                }
            if (leaf && cur.size() > 1) out.push_back(cur);
            cur.pop_back();
        };
    if (core[root] > 1) {
        core_rec(root, cps);
        for (uint32_t u = 0; u < n; u++) {
            if (core[u] <= 1) continue;
            std::vector<std::vector<uint32_t>> tp;
            cur.clear();
            tree_rec(u, tp);
            if (!tp.empty()) forests.push_back(tp);
        }
    } else {
        std::vector<std::vector<uint32_t>> tp;
        cur.clear();
        tree_rec(root, tp);
        if (!tp.empty()) forests.push_back(tp);
    }
    uint32_t sel_cnt = 0, vis = 0;
    auto take = [&](uint32_t v) {
        if ((vis >> v) & 1u) return;
        order[sel_cnt++] = v;
        vis |= 1u << v;
    };
    take(root);
    const double dmax = std::numeric_limits<double>::max();
    if (!cps.empty()) {
        std::vector<std::vector<uint64_t>> est(cps.size());
        std::vector<uint32_t> ntree(cps.size());
        for (size_t i = 0; i < cps.size(); i++) {
            // generateNoneTreeEdgesCount (:241-250), + 1
            uint32_t nt = deg_of(q, cps[i][0]) - (uint32_t)t.children[cps[i][0]].size();
            for (size_t j = 1; j < cps[i].size(); j++)
                nt += deg_of(q, cps[i][j]) - (uint32_t)t.children[cps[i][j]].size() - 1;
            ntree[i] = nt + 1;
            const int rc = estimate(cps[i], est[i]);
            if (rc) return rc;
        }
        double minv = dmax;
        size_t sel = 0;
        for (size_t i = 0; i < cps.size(); i++) {
            const double cv = est[i][0] / (double)ntree[i];
            if (cv < minv) {
                minv = cv;
                sel = i;
            }
        }
        for (;;) {
            for (size_t j = 1; j < cps[sel].size(); j++) take(cps[sel][j]);
            cps.erase(cps.begin() + sel);
            est.erase(est.begin() + sel);
            if (cps.empty()) break;
            minv = dmax;
            sel = 0;
            for (size_t i = 0; i < cps.size(); i++) {
                uint32_t r = 0;
                for (uint32_t j = 0; j < cps[i].size(); j++) {
                    if ((vis >> cps[i][j]) & 1u) continue;
                    r = j - 1;
                    break;
                }
                const double cv = est[i][r] / (double)cnt[cps[i][r]];
                if (cv < minv) {
                    minv = cv;
                    sel = i;
                }
            }
        }
    }
    for (auto& tp : forests) {
        std::vector<std::vector<uint64_t>> est(tp.size());
        for (size_t i = 0; i < tp.size(); i++) {
            const int rc = estimate(tp[i], est[i]);
            if (rc) return rc;
        }
        while (!tp.empty()) {
            double minv = dmax;
            size_t sel = 0;
            for (size_t i = 0; i < tp.size(); i++) {
                uint32_t r = 0;
                for (uint32_t j = 0; j < tp[i].size(); j++) {
                    if ((vis >> tp[i][j]) & 1u) continue;
                    r = j == 0 ? j : j - 1;
                    break;
                }
                const double cv = est[i][r] / (double)cnt[tp[i][r]];
                if (cv < minv) {
                    minv = cv;
                    sel = i;
                }
            }
            for (uint32_t v : tp[sel]) take(v);
            tp.erase(tp.begin() + sel);
            est.erase(est.begin() + sel);
        }
    }
    while (!leaves.empty()) {
        double minv = dmax;
        size_t sel = 0;
        for (size_t i = 0; i < leaves.size(); i++) {
            const double cv = cnt[leaves[i]];
            if (cv < minv) {
                minv = cv;
                sel = i;
            }
        }
        take(leaves[sel]);
        leaves.erase(leaves.begin() + sel);
    }
    return sel_cnt == n ? 0 : -1;
}

// _GuP_start_vertex (inc/order.h:481-535) over the NLF candidate counts.
static uint32_t gup_start(const QueryCSR& q, const int32_t* core, const uint32_t* nlf_cnt)
{
    uint32_t maxk = 0;
    std::vector<uint32_t> cand;
    for (uint32_t u = 0; u < q.n; u++) maxk = std::max(maxk, (uint32_t)core[u]);
    for (uint32_t u = 0; u < q.n; u++)
        if ((uint32_t)core[u] == maxk) cand.push_back(u);
    for (int k = (int)maxk; k >= 1; k--) {
        std::vector<std::pair<uint32_t, double>> ps;
        for (uint32_t u : cand) {
            int cd = 0;
            for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) cd += core[q.adj[p]] >= k;
            if (cd == 0) continue;
            ps.emplace_back(u, nlf_cnt[u] / (double)cd);
        }
        if (ps.empty()) continue;
        const double minp = std::min_element(ps.begin(), ps.end(), [](const auto& a, const auto& b) {
                                return a.second < b.second;
                            })->second;
        std::vector<uint32_t> mins;
        for (auto& [u, pen] : ps) {
            const double scale = std::max(1.0, std::max(std::fabs(pen), std::fabs(minp)));
            if (std::fabs(pen - minp) <= 1e-12 * scale) mins.push_back(u);
        }
        if (mins.size() == 1 || k == 1) return mins.front();
    }
    return cand.front();
}

// Order_GuP (inc/order.h:537-561): gupOrderCore (:299-384), then gupOrderTree (:412-450).
int order_gup(const QueryCSR& q, const uint32_t* nlf_cnt, const uint32_t* cnt, const uint32_t* ecnt,
              const CountEmbed& count_embed, uint32_t* order)
{
    const uint32_t n = q.n;
    bool tree_or_cycle = true;
    for (uint32_t u = 0; u < n; u++) tree_or_cycle &= deg_of(q, u) <= 2;
    if (tree_or_cycle) return -2;  // the reference asserts (:484-492)
    int32_t core[SGM_MAXQ];
    query_kcore(q, core);
    const uint32_t root = gup_start(q, core, nlf_cnt);
    BfsTree t;
    bfs_tree(q, root, t);
    if (t.order.size() != n) return -1;
    std::vector<uint32_t> out{root};
    std::vector<uint8_t> ordered(n, 0), queued(n, 0);
    ordered[root] = 1;
    std::vector<std::vector<double>> w(n, std::vector<double>(n, 0.0));
    for (uint32_t v = 0; v < n; v++) {
        if (core[v] < 2) continue;
        for (uint32_t p = q.off[v]; p < q.off[v + 1]; p++) {
            const uint32_t x = q.adj[p];
            // bfs_order[v] is the v-th vertex of the BFS order, compared by value (:321)
            if (core[x] >= 2 && t.order[v] < t.order[x]) {
                double wt = (double)ecnt[v * n + x];
                wt = wt / (double)cnt[v];
                wt = wt / (double)cnt[x];
                w[v][x] = wt;
                w[x][v] = wt;
            }
        }
    }
    std::vector<double> sel(n, 1.0);
    std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>, std::greater<std::pair<double, int>>>
        pq;
    uint32_t v = root;
    for (;;) {
        for (uint32_t p = q.off[v]; p < q.off[v + 1]; p++) {
            const uint32_t x = q.adj[p];
            if (core[x] >= 2 && !ordered[x]) {
                queued[x] = 1;
                sel[x] *= w[v][x];
            }
        }
        for (uint32_t p = q.off[v]; p < q.off[v + 1]; p++) {
            const uint32_t x = q.adj[p];
            if (core[x] >= 2 && !ordered[x]) {
                double c = sel[x] * static_cast<double>(cnt[x]);
                for (uint32_t r = q.off[x]; r < q.off[x + 1]; r++) {
                    const uint32_t y = q.adj[r];
                    if (queued[y] && !ordered[y]) c *= w[x][y] * sel[y];
                }
                pq.push(std::make_pair(c, (int)x));
            }
        }
        bool found = false;
        while (!pq.empty()) {
            v = (uint32_t)pq.top().second;
            pq.pop();
            if (!ordered[v]) {
                found = true;
                break;
            }
        }
        if (!found) break;
        out.push_back(v);
        ordered[v] = 1;
    }
    std::vector<double> nembeds(n, 0.0);
    using Item = std::tuple<int, double, uint32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (uint32_t u = 0; u < n; u++) {
        bool has = false;
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) has |= !ordered[q.adj[p]];
        if (ordered[u] && has) {
            const int rc = count_embed(u, ordered.data(), t, nembeds.data());
            if (rc) return rc;
            for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
                const uint32_t x = q.adj[p];
                if (!ordered[x]) heap.push(std::make_tuple(0, nembeds[x], x));
            }
        }
    }
    while (!heap.empty()) {
        const auto [d, e, u] = heap.top();
        (void)e;
        heap.pop();
        if (ordered[u]) return -2;  // the reference asserts (:434)
        out.push_back(u);
        ordered[u] = 1;
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
            const uint32_t x = q.adj[p];
            if (!ordered[x]) heap.push(std::make_tuple(d - 1, nembeds[x], x));
        }
    }
    if (out.size() != n) return -1;
    std::copy(out.begin(), out.end(), order);
    return 0;
}

}  // namespace sgm
