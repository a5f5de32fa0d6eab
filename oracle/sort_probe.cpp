// std::sort over records ordered like the reference's VertexSetIR::operator<
// (inc/code_gen.h:633-650): depth, label (always 0), size, then lexicographic edge
// indices.  Test infrastructure: it lets tests compare the oracle's restated sort with
// the library routine itself on the same inputs.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {
struct Rec {
    uint32_t depth, size, payload;
    std::vector<uint32_t> e;
    bool operator<(const Rec &r) const {
        if (this == &r) return false;
        if (depth != r.depth) return depth < r.depth;
        if (e.size() != r.e.size()) return e.size() < r.e.size();
        for (size_t i = 0; i < e.size(); i++)
            if (e[i] != r.e[i]) return e[i] < r.e[i];
        return false;
    }
};
}  // namespace

extern "C" void probe_std_sort(long n, const uint32_t *depth, const uint32_t *size,
                               const uint32_t *elems, uint32_t *perm_out) {
    std::vector<Rec> v(n);
    for (long i = 0; i < n; i++) {
        v[i].depth = depth[i];
        v[i].size = size[i];
        v[i].payload = (uint32_t)i;
        v[i].e.assign(elems + i * 32, elems + i * 32 + size[i]);
    }
    std::sort(v.begin(), v.end());
    for (long i = 0; i < n; i++) perm_out[i] = v[i].payload;
}

// std::priority_queue over (vertex, key) pairs with the comparator of the reference's
// CFLStartVertex (inc/filter.h:437-441: l.second < r.second, a max-heap on the key): push
// (v, keys[v]) for v < count, pop while more than `keep` remain, then pop the rest into
// out[] (the order in which CFLStartVertex visits them).
#include <queue>
#include <utility>

extern "C" void probe_std_heap(unsigned count, const double *keys, unsigned keep, uint32_t *out) {
    auto cmp = [](std::pair<uint32_t, double> l, std::pair<uint32_t, double> r) { return l.second < r.second; };
    std::priority_queue<std::pair<uint32_t, double>, std::vector<std::pair<uint32_t, double>>, decltype(cmp)> q(cmp);
    for (unsigned v = 0; v < count; v++) q.push(std::make_pair((uint32_t)v, keys[v]));
    while (q.size() > keep) q.pop();
    unsigned k = 0;
    while (!q.empty()) {
        out[k++] = q.top().first;
        q.pop();
    }
}

// The dynamic-order search (inc/explore.h:1674-2062) restated with the reference's own
// containers: a std::priority_queue per depth under extendable_vertex_compare, copied from the
// depth above (vector of queues, emplace_back(back())), candidate lists from the aux tables of
// the previous BFS neighbours.  It pins the C oracle's restated heap moves inside the whole
// search: counts, intersections and per-depth extensions must agree.  weight: per vertex
// pointer to cnt[u] u32 weights, or null (the VEQ variant ranks by candidate count).
#include <functional>

namespace {
using QEnt = std::pair<std::pair<uint32_t, uint32_t>, uint32_t>;
struct QCmp {
    bool operator()(const QEnt& l, const QEnt& r) const
    {
        if (l.first.second == 1 && r.first.second != 1) return true;
        if (l.first.second != 1 && r.first.second == 1) return false;
        return l.second > r.second;
    }
};
using DynPQ = std::priority_queue<QEnt, std::vector<QEnt>, QCmp>;
}  // namespace

extern "C" int probe_dyn_enum(uint32_t V, const uint32_t* off, const uint32_t* adj, uint32_t n, const uint32_t* qoff,
                              const uint32_t* qadj, const uint32_t* cand, const uint32_t* cnt, uint32_t stride,
                              const uint32_t* bfs, const uint32_t* const* weight, uint64_t cap, uint64_t* out_count,
                              uint64_t* out_inter, uint64_t* out_nodes)
{
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; i++) pos[bfs[i]] = i;
    std::vector<std::vector<uint32_t>> prev(n), fol(n);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t v = bfs[i];
        for (uint32_t j = 0; j < n; j++) {
            uint32_t w = bfs[j];
            bool a = false;
            for (uint32_t p = qoff[v]; p < qoff[v + 1]; p++) a |= qadj[p] == w;
            if (!a) continue;
            if (j < i) prev[v].push_back(w);
            if (j > i) fol[v].push_back(w);
        }
    }
    // tables[a][b][i] = ascending positions in C(b) adjacent to C(a)[i]
    std::vector<std::vector<std::vector<std::vector<uint32_t>>>> T(n, std::vector<std::vector<std::vector<uint32_t>>>(n));
    std::vector<int64_t> where(V, -1);
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = qoff[a]; p < qoff[a + 1]; p++) {
            uint32_t b = qadj[p];
            for (uint32_t k = 0; k < cnt[b]; k++) where[cand[(size_t)b * stride + k]] = k;
            T[a][b].assign(cnt[a], {});
            for (uint32_t i = 0; i < cnt[a]; i++) {
                uint32_t v = cand[(size_t)a * stride + i];
                for (uint32_t q = off[v]; q < off[v + 1]; q++)
                    if (where[adj[q]] >= 0) T[a][b][i].push_back((uint32_t)where[adj[q]]);
                std::sort(T[a][b][i].begin(), T[a][b][i].end());
            }
            for (uint32_t k = 0; k < cnt[b]; k++) where[cand[(size_t)b * stride + k]] = -1;
        }
    std::vector<std::vector<uint32_t>> valid(n);
    std::vector<uint32_t> ext(n), idx(n, 0), mo(n, 0);
    std::vector<char> vis(V, 0);
    for (uint32_t u = 0; u < n; u++) ext[u] = (uint32_t)prev[u].size();
    uint64_t ans = 0, inter = 0;
    for (uint32_t d = 0; d < n; d++) out_nodes[d] = 0;
    bool stop = false;
    std::vector<DynPQ> vq;
    auto cands = [&](uint32_t u) {
        valid[u] = T[prev[u][0]][u][idx[prev[u][0]]];
        for (size_t i = 1; i < prev[u].size(); i++) {
            const auto& row = T[prev[u][i]][u][idx[prev[u][i]]];
            std::vector<uint32_t> r;
            std::set_intersection(row.begin(), row.end(), valid[u].begin(), valid[u].end(), std::back_inserter(r));
            valid[u] = r;
            inter++;
        }
    };
    auto update = [&](uint32_t u) {
        for (uint32_t f : fol[u])
            if (--ext[f] == 0) {
                cands(f);
                uint32_t w = 0;
                if (weight)
                    for (uint32_t x : valid[f]) w += weight[f][x];
                else
                    w = (uint32_t)valid[f].size();
                vq.back().emplace(std::make_pair(std::make_pair(f, qoff[f + 1] - qoff[f]), w));
            }
    };
    auto restore = [&](uint32_t u) {
        for (uint32_t f : fol[u]) ext[f]++;
    };
    std::function<void(uint32_t)> rec = [&](uint32_t d) {
        uint32_t u = mo[d];
        if (valid[u].empty()) return;
        const std::vector<uint32_t> mine = valid[u];  // (valid[u] is not rewritten below this depth)
        for (uint32_t vi : mine) {
            uint32_t v = cand[(size_t)u * stride + vi];
            if (vis[v]) continue;
            out_nodes[d]++;
            if (d == n - 1) {
                if (++ans == cap) { stop = true; return; }
                continue;
            }
            idx[u] = vi;
            vis[v] = 1;
            vq.emplace_back(vq.back());
            update(u);
            mo[d + 1] = vq.back().top().first.first;
            vq.back().pop();
            rec(d + 1);
            if (stop) return;
            vis[v] = 0;
            vq.pop_back();
            restore(u);
        }
    };
    const uint32_t u0 = bfs[0];
    mo[0] = u0;
    for (uint32_t p = 0; p < cnt[u0] && !stop; p++) {
        out_nodes[0]++;
        idx[u0] = p;
        vis[cand[(size_t)u0 * stride + p]] = 1;
        vq.emplace_back(DynPQ(QCmp()));
        update(u0);
        mo[1] = vq.back().top().first.first;
        vq.back().pop();
        rec(1);
        if (stop) break;
        vis[cand[(size_t)u0 * stride + p]] = 0;
        vq.pop_back();
        restore(u0);
    }
    *out_count = ans;
    *out_inter = inter;
    return 0;
}

// ----------------------------------------------------------------------------------------
// GuP's reservation guards written with the standard containers the reference uses
// (std::unordered_set for the chosen vertices, std::partition over the reservation
// candidates, std::vector<pair>::erase, std::sort over the per-neighbour guards), to check
// the restatement of their element order in sgm_oracle_gup.c.  Tables come from the data
// graph directly; same output layout as orc_gup_guards.
#include <unordered_set>

extern "C" int probe_gup_guards(uint32_t V, const uint32_t* off, const uint32_t* adj, uint32_t n, const uint32_t* qoff,
                                const uint32_t* qadj, const uint32_t* cand, const uint32_t* cnt, uint32_t stride,
                                const uint32_t* order, uint32_t r, uint32_t* goff, uint32_t* out)
{
    std::vector<uint32_t> pos_of(n), former(n, 0), match(V, 0);
    for (uint32_t i = 0; i < n; i++) pos_of[order[i]] = i;
    for (uint32_t i = 1; i < n; i++) former[i] = former[i - 1] | (1u << (i - 1));
    std::vector<std::vector<uint32_t>> fdep(n);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t p = qoff[order[i]]; p < qoff[order[i] + 1]; p++)
            if (pos_of[qadj[p]] > i) fdep[i].push_back(pos_of[qadj[p]]);
        std::sort(fdep[i].begin(), fdep[i].end());
        for (uint32_t j = 0; j < cnt[order[i]]; j++) match[cand[(size_t)order[i] * stride + j]] |= 1u << i;
    }
    // row j of table (a, b): positions in C(b) of the neighbours of C(a)[j]
    auto row = [&](uint32_t a, uint32_t b, uint32_t j) {
        std::vector<uint32_t> rw;
        const uint32_t v = cand[(size_t)a * stride + j];
        const uint32_t* cb = cand + (size_t)b * stride;
        for (uint32_t q = off[v]; q < off[v + 1]; q++) {
            const uint32_t* it = std::lower_bound(cb, cb + cnt[b], adj[q]);
            if (it != cb + cnt[b] && *it == adj[q]) rw.push_back((uint32_t)(it - cb));
        }
        return rw;
    };
    goff[0] = 0;
    for (uint32_t i = 0; i < n; i++) goff[i + 1] = goff[i] + cnt[order[i]];
    std::fill(out, out + 4 * (size_t)goff[n], 0u);
    std::vector<uint32_t> freq(V, 0);
    auto pc = [](uint32_t x) { return (uint32_t)__builtin_popcount(x); };
    for (int i = (int)n - 1; i >= 0; i--) {
        const uint32_t u = order[i];
        for (uint32_t j = 0; j < cnt[u]; j++) {
            const uint32_t v = cand[(size_t)u * stride + j];
            std::vector<std::pair<uint32_t, std::vector<uint32_t>>> all;
            bool failed = false;
            for (auto it = fdep[i].rbegin(); it != fdep[i].rend() && !failed; ++it) {
                const uint32_t ufd = *it, uf = order[ufd];
                std::unordered_set<uint32_t> chosen;
                std::vector<uint32_t> rev;
                std::vector<std::pair<uint32_t, uint32_t>> gr;
                uint32_t dom = 0;
                const std::vector<uint32_t> rw = row(u, uf, j);
                for (uint32_t fi : rw) {
                    const uint32_t vf = cand[(size_t)uf * stride + fi];
                    const bool vf_ok = (match[vf] & former[i]) != 0;
                    const uint32_t* g = out + 4 * (size_t)(goff[ufd] + fi);
                    std::vector<uint32_t> ws = g[0] ? std::vector<uint32_t>(g + 1, g + 1 + g[0])
                                                    : std::vector<uint32_t>{vf};
                    for (uint32_t w : ws) {
                        if (w == v) continue;
                        if (chosen.count(vf)) break;
                        const bool w_ok = (match[w] & former[i]) != 0;
                        if (vf_ok && (vf == w || !w_ok)) {
                            chosen.insert(vf);
                            dom |= match[vf];
                        } else if (!vf_ok && w_ok) {
                            chosen.insert(w);
                            dom |= match[w];
                        }
                        if ((!vf_ok && !w_ok) || chosen.size() > r || chosen.size() > pc(former[i] & dom)) {
                            failed = true;
                            break;
                        }
                    }
                    if (failed) break;
                }
                if (failed) break;
                for (uint32_t fi : rw) {
                    const uint32_t vf = cand[(size_t)uf * stride + fi];
                    const uint32_t* g = out + 4 * (size_t)(goff[ufd] + fi);
                    if (!g[0] || chosen.count(vf)) continue;
                    for (uint32_t k = 0; k < g[0]; k++)
                        if (g[1 + k] != v && !chosen.count(g[1 + k])) gr.emplace_back(vf, g[1 + k]);
                }
                for (auto& e : gr) {
                    if (freq[e.first] == 0) rev.push_back(e.first);
                    if (freq[e.second] == 0) rev.push_back(e.second);
                    freq[e.first]++;
                    freq[e.second]++;
                }
                while (chosen.size() < r) {
                    auto mid = std::partition(rev.begin(), rev.end(), [&](uint32_t x) {
                        return freq[x] > 0 && chosen.size() + 1 <= pc((dom | match[x]) & former[i]);
                    });
                    const size_t valid = (size_t)(mid - rev.begin());
                    if (valid == 0) break;
                    size_t mi = 0;
                    uint32_t mf = 0;
                    for (size_t k = 0; k < valid; k++)
                        if (freq[rev[k]] >= mf) { mi = k; mf = freq[rev[k]]; }
                    const uint32_t pick = rev[mi];
                    chosen.insert(pick);
                    for (auto e = gr.begin(); e != gr.end();) {
                        if (e->first == pick || e->second == pick) {
                            freq[e->first]--;
                            freq[e->second]--;
                            e = gr.erase(e);
                        } else {
                            ++e;
                        }
                    }
                }
                for (uint32_t x : rev) freq[x] = 0;
                if (gr.empty() && !chosen.empty()) {
                    std::vector<uint32_t> gv(chosen.begin(), chosen.end());
                    std::sort(gv.begin(), gv.end());
                    all.emplace_back(ufd, gv);
                }
            }
            if (failed || all.empty()) continue;
            std::sort(all.begin(), all.end(), [](const auto& a, const auto& b) {
                if (a.second.size() != b.second.size()) return a.second.size() < b.second.size();
                return a.first > b.first;
            });
            uint32_t* o = out + 4 * (size_t)(goff[i] + j);
            o[0] = (uint32_t)all[0].second.size();
            for (size_t k = 0; k < all[0].second.size(); k++) o[1 + k] = all[0].second[k];
        }
    }
    return 0;
}
