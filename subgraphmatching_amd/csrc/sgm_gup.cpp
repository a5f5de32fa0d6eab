// GuP's conflict detection (GUPConflictDetection, inc/prune.h:488-923): the reservation
// guards of every (depth, candidate position) of a query for an order, computed on the host
// from the query's aux tables (Init :828-905 over compute_rev :599-825), kept on the device
// for the aux-table search, whose waves prune a candidate when every vertex of its guard is
// matched (matchCheck :577-595).
//
// compute_rev for depth i reads the guards of deeper depths only, so a depth's candidates are
// computed in parallel on host threads, deepest depth first.  Each thread has its own
// frequency array (coverInfo) and scratch.  The element orders that decide ties are the
// reference's: its std::partition (libstdc++'s two-ended swap) of the reservation candidates,
// the last candidate of highest frequency, the guard of smallest size with ties to the
// deeper following neighbour, and a failed neighbour dropping the guard altogether.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "sgm_objects.h"

namespace {

struct GupShared {
    uint32_t n = 0, r = 3;
    std::vector<uint32_t> order, former, match;  // match: V words, bit i = v in C(order[i])
    std::vector<std::vector<uint32_t>> fdep;     // following neighbours' depths, ascending
    std::vector<std::vector<uint32_t>> cand;     // C(u) by query vertex
    // table (a, b) rows, by query vertex pair
    std::vector<std::vector<uint32_t>> toff, tidx;
    std::vector<uint32_t> goff;                  // first record of each depth
    std::vector<uint32_t> words;                 // 4 per record: size, vertices ascending
};

struct GupScratch {
    std::vector<uint32_t> freq;  // V
    std::vector<uint32_t> rev;
    std::vector<std::pair<uint32_t, uint32_t>> gr;
    uint32_t chosen[8];
    uint32_t nchosen = 0;

    bool has(uint32_t x) const
    {
        for (uint32_t i = 0; i < nchosen; i++)
            if (chosen[i] == x) return true;
        return false;
    }
    void add(uint32_t x)
    {
        if (!has(x)) chosen[nchosen++] = x;
    }
};

inline uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

void compute_rev(const GupShared& G, GupScratch& S, uint32_t ud, uint32_t vi, uint32_t* out)
{
    const uint32_t n = G.n, u = G.order[ud];
    const uint32_t v = G.cand[u][vi];
    uint32_t best_sz = 0, best[8];
    bool have = false;
    out[0] = 0;
    const std::vector<uint32_t>& fd = G.fdep[ud];
    for (size_t k = fd.size(); k-- > 0;) {  // deepest following neighbour first
        const uint32_t ufd = fd[k], uf = G.order[ufd];
        S.gr.clear();
        S.rev.clear();
        S.nchosen = 0;
        uint32_t dom = 0;
        const std::vector<uint32_t>& to = G.toff[(size_t)u * n + uf];
        const uint32_t* row = G.tidx[(size_t)u * n + uf].data() + to[vi];
        const uint32_t rn = to[vi + 1] - to[vi];
        for (uint32_t e = 0; e < rn; e++) {
            const uint32_t fi = row[e], vf = G.cand[uf][fi];
            const bool vf_ok = (G.match[vf] & G.former[ud]) != 0;
            const uint32_t* gg = G.words.data() + 4 * (size_t)(G.goff[ufd] + fi);
            const uint32_t wn = gg[0] ? gg[0] : 1u;  // no guard: the trivial one, {vf}
            for (uint32_t wi = 0; wi < wn; wi++) {
                const uint32_t w = gg[0] ? gg[1 + wi] : vf;
                if (w == v) continue;
                if (S.has(vf)) break;
                const bool w_ok = (G.match[w] & G.former[ud]) != 0;
                if (vf_ok && (vf == w || !w_ok)) {
                    S.add(vf);
                    dom |= G.match[vf];
                } else if (!vf_ok && w_ok) {
                    S.add(w);
                    dom |= G.match[w];
                }
                if ((!vf_ok && !w_ok) || S.nchosen > G.r || S.nchosen > popc(G.former[ud] & dom)) {
                    out[0] = 0;  // this neighbour cannot be reserved: no guard at all
                    return;
                }
            }
        }
        // pairs still to cover: a guard vertex of a following candidate not yet reserved
        for (uint32_t e = 0; e < rn; e++) {
            const uint32_t fi = row[e], vf = G.cand[uf][fi];
            const uint32_t* gg = G.words.data() + 4 * (size_t)(G.goff[ufd] + fi);
            if (!gg[0] || S.has(vf)) continue;
            for (uint32_t wi = 0; wi < gg[0]; wi++) {
                const uint32_t w = gg[1 + wi];
                if (w != v && !S.has(w)) S.gr.emplace_back(vf, w);
            }
        }
        for (const auto& p : S.gr) {
            if (S.freq[p.first] == 0) S.rev.push_back(p.first);
            if (S.freq[p.second] == 0) S.rev.push_back(p.second);
            S.freq[p.first]++;
            S.freq[p.second]++;
        }
        while (S.nchosen < G.r) {
            auto keep = [&](uint32_t x) {
                return S.freq[x] > 0 && S.nchosen + 1 <= popc((dom | G.match[x]) & G.former[ud]);
            };
            // libstdc++'s std::partition: two-ended swap (its element order decides ties)
            size_t first = 0, last = S.rev.size();
            while (true) {
                while (first != last && keep(S.rev[first])) first++;
                if (first == last) break;
                last--;
                while (first != last && !keep(S.rev[last])) last--;
                if (first == last) break;
                std::swap(S.rev[first], S.rev[last]);
                first++;
            }
            const size_t valid = first;
            if (valid == 0) break;
            size_t mi = 0;
            uint32_t mf = 0;
            for (size_t i = 0; i < valid; i++)
                if (S.freq[S.rev[i]] >= mf) {
                    mi = i;
                    mf = S.freq[S.rev[i]];
                }
            const uint32_t pick = S.rev[mi];
            S.add(pick);
            size_t w = 0;
            for (size_t i = 0; i < S.gr.size(); i++) {
                const auto p = S.gr[i];
                if (p.first == pick || p.second == pick) {
                    S.freq[p.first]--;
                    S.freq[p.second]--;
                } else {
                    S.gr[w++] = p;
                }
            }
            S.gr.resize(w);
        }
        for (uint32_t x : S.rev) S.freq[x] = 0;
        if (S.gr.empty() && S.nchosen && (!have || S.nchosen < best_sz)) {
            have = true;
            best_sz = S.nchosen;
            std::copy(S.chosen, S.chosen + S.nchosen, best);
        }
    }
    if (have) {
        std::sort(best, best + best_sz);
        out[0] = best_sz;
        std::copy(best, best + best_sz, out + 1);
    }
}

}  // namespace

// Guards of q for `order` (its tables built), uploaded to q->d_guard in depth order.
extern "C++" int gup_guards_build(sgm_query* q, const uint32_t* order)
{
    sgm_ctx* c = q->g->ctx;
    const uint32_t n = q->q.n, V = q->g->V;
    GupShared G;
    G.n = n;
    G.order.assign(order, order + n);
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; i++) pos[order[i]] = i;
    G.former.assign(n, 0);
    for (uint32_t i = 1; i < n; i++) G.former[i] = G.former[i - 1] | (1u << (i - 1));
    G.fdep.assign(n, {});
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t p = q->q.off[order[i]]; p < q->q.off[order[i] + 1]; p++)
            if (pos[q->q.adj[p]] > i) G.fdep[i].push_back(pos[q->q.adj[p]]);
        std::sort(G.fdep[i].begin(), G.fdep[i].end());
    }
    // candidates and the tables (u -> following neighbour) from the device
    G.cand.assign(n, {});
    for (uint32_t u = 0; u < n; u++) {
        G.cand[u].resize(q->counts[u]);
        if (q->counts[u])
            HIPCHK(c, hipMemcpy(G.cand[u].data(), q->d_list + q->cand_off[u], sizeof(uint32_t) * q->counts[u],
                                hipMemcpyDeviceToHost));
    }
    G.match.assign(V, 0);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t v : G.cand[order[i]]) G.match[v] |= 1u << i;
    G.toff.assign((size_t)n * n, {});
    G.tidx.assign((size_t)n * n, {});
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t fd : G.fdep[i]) {
            const uint32_t a = order[i], b = order[fd];
            const sgm_table& t = q->tables[(size_t)a * n + b];
            if (!t.built) return set_err(c, SGM_ESTATE, "conflict detection: table (%u,%u) missing", a, b);
            auto& to = G.toff[(size_t)a * n + b];
            auto& ti = G.tidx[(size_t)a * n + b];
            if (const int rc = table_fetch(q, a, b, to, ti)) return rc;
            if (ti.empty()) ti.resize(1);
        }
    G.goff.assign(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) G.goff[i + 1] = G.goff[i] + q->counts[order[i]];
    G.words.assign(4 * (size_t)std::max<uint32_t>(G.goff[n], 1), 0);
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<GupScratch> scratch(hw);
    for (auto& s : scratch) s.freq.assign(V, 0);
    for (int i = (int)n - 1; i >= 0; i--) {
        const uint32_t cnt = q->counts[order[i]];
        const uint32_t nt = cnt >= 4096 ? hw : 1u;
        auto work = [&](uint32_t t) {
            for (uint32_t j = t; j < cnt; j += nt)
                compute_rev(G, scratch[t], (uint32_t)i, j, G.words.data() + 4 * (size_t)(G.goff[i] + j));
        };
        if (nt == 1) {
            work(0);
            continue;
        }
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto& x : th) x.join();
    }
    const size_t need = G.words.size();
    if (q->guard_cap < need) {
        dfree(q->d_guard);
        q->d_guard = nullptr;
        q->guard_cap = 0;
        HIPCHK(c, dalloc(&q->d_guard, need));
        q->guard_cap = need;
    }
    HIPCHK(c, hipMemcpy(q->d_guard, G.words.data(), sizeof(uint32_t) * need, hipMemcpyHostToDevice));
    std::copy(G.goff.begin(), G.goff.end(), q->guard_off);
    std::copy(order, order + n, q->guard_order);
    q->guards_ok = true;
    return SGM_OK;
}

extern "C" {

int sgm_gup_guards(sgm_query* q, const uint32_t* order, uint32_t* goff, uint32_t* words, uint64_t words_cap)
{
    if (!q || !order) return SGM_EINVAL;
    sgm_ctx* c = q->g->ctx;
    if (const int rc = check_filtered(q, "GuP guards")) return rc;
    if (sgm::check_order(q->q, order)) return set_err(c, SGM_EORDER, "order is not a connected permutation");
    if (q->empty) return set_err(c, SGM_EEMPTY, "GuP guards: some candidate set is empty");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    bool built = true;
    for (uint32_t a = 0; a < q->q.n && built; a++)
        for (uint32_t p = q->q.off[a]; p < q->q.off[a + 1]; p++) built &= q->tables[(size_t)a * q->q.n + q->q.adj[p]].built;
    if (!built && (rc = build_tables_many(&q, 1))) return rc;
    if ((rc = gup_guards_build(q, order))) return rc;
    const uint32_t n = q->q.n;
    if (goff) std::copy(q->guard_off, q->guard_off + n + 1, goff);
    if (words) {
        const uint64_t total = 4ull * q->guard_off[n];
        if (words_cap < total) return set_err(c, SGM_EINVAL, "GuP guards: %llu words needed", (unsigned long long)total);
        if (total) HIPCHK(c, hipMemcpy(words, q->d_guard, sizeof(uint32_t) * total, hipMemcpyDeviceToHost));
    }
    return SGM_OK;
}

}  // extern "C"
