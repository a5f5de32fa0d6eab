// Host-side query planning: RI and GQL matching orders, backward neighbours, DAF ancestor
// masks and the intersection-cache plan.  These are O(n^2..n^3) over queries of at most 32
// vertices and stay on the CPU, as in the reference.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "sgm_host.h"

namespace sgm {

static bool q_adj(const QueryCSR& q, uint32_t a, uint32_t b)
{
    for (uint32_t p = q.off[a]; p < q.off[a + 1]; p++)
        if (q.adj[p] == b) return true;
    return false;
}

static uint32_t q_deg(const QueryCSR& q, uint32_t a) { return q.off[a + 1] - q.off[a]; }

// RI order (inc/order.h:1138-1281): root of maximum degree (first on ties); then the vertex
// with most already-ordered neighbours; ties broken by (1) how many ordered vertices share
// a neighbour with the candidate's unordered neighbours, (2) how many unordered neighbours
// of the candidate touch no ordered vertex; final ties go to the lowest id.
void order_ri(const QueryCSR& q, uint32_t* order)
{
    const uint32_t n = q.n;
    std::vector<char> vis(n, 0);
    order[0] = 0;
    for (uint32_t i = 1; i < n; i++)
        if (q_deg(q, i) > q_deg(q, order[0])) order[0] = i;
    vis[order[0]] = 1;
    std::vector<uint32_t> ties, prev;
    for (uint32_t i = 1; i < n; i++) {
        ties.clear();
        uint32_t best = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (vis[u]) continue;
            uint32_t bn = 0;
            for (uint32_t j = 0; j < i; j++) bn += q_adj(q, u, order[j]);
            if (bn > best) { best = bn; ties.assign(1, u); }
            else if (bn == best) ties.push_back(u);
        }
        if (ties.size() != 1) {
            prev.swap(ties);
            ties.clear();
            best = 0;
            for (uint32_t u : prev) {
                std::vector<uint32_t> fn;
                for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++)
                    if (!vis[q.adj[p]]) fn.push_back(q.adj[p]);
                uint32_t cc = 0;
                for (uint32_t j = 0; j < i; j++) {
                    const uint32_t uu = order[j];
                    bool share = false;
                    for (uint32_t p = q.off[uu]; p < q.off[uu + 1] && !share; p++)
                        share = std::binary_search(fn.begin(), fn.end(), q.adj[p]);
                    cc += share;
                }
                if (cc > best) { best = cc; ties.assign(1, u); }
                else if (cc == best) ties.push_back(u);
            }
        }
        if (ties.size() != 1) {
            prev.swap(ties);
            ties.clear();
            best = 0;
            for (uint32_t u : prev) {
                uint32_t cc = 0;
                for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
                    const uint32_t uu = q.adj[p];
                    if (vis[uu]) continue;
                    bool touches = false;
                    for (uint32_t j = 0; j < i && !touches; j++) touches = q_adj(q, uu, order[j]);
                    cc += !touches;
                }
                if (cc > best) { best = cc; ties.assign(1, u); }
                else if (cc == best) ties.push_back(u);
            }
        }
        order[i] = ties[0];
        vis[order[i]] = 1;
    }
}

// GQL order (inc/order.h:563-597, start vertex :11-35): fewest candidates first (larger
// degree on ties), then repeatedly the unordered neighbour of the ordered set with the
// fewest candidates (larger degree on ties, else the lowest id).
void order_gql(const QueryCSR& q, const uint32_t* cand_cnt, uint32_t num_data_vertices, uint32_t* order)
{
    const uint32_t n = q.n;
    std::vector<char> vis(n, 0), near(n, 0);
    uint32_t s = 0;
    for (uint32_t i = 1; i < n; i++) {
        if (cand_cnt[i] < cand_cnt[s]) s = i;
        else if (cand_cnt[i] == cand_cnt[s] && q_deg(q, i) > q_deg(q, s)) s = i;
    }
    auto take = [&](uint32_t x) {
        vis[x] = 1;
        for (uint32_t p = q.off[x]; p < q.off[x + 1]; p++) near[q.adj[p]] = 1;
    };
    order[0] = s;
    take(s);
    for (uint32_t i = 1; i < n; i++) {
        uint32_t nx = 0;
        uint64_t mn = (uint64_t)num_data_vertices + 1;
        for (uint32_t j = 0; j < n; j++) {
            if (vis[j] || !near[j]) continue;
            if (cand_cnt[j] < mn) { mn = cand_cnt[j]; nx = j; }
            else if (cand_cnt[j] == mn && q_deg(q, j) > q_deg(q, nx)) nx = j;
        }
        order[i] = nx;
        take(nx);
    }
}

int check_order(const QueryCSR& q, const uint32_t* order)
{
    const uint32_t n = q.n;
    std::vector<char> seen(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        if (order[i] >= n || seen[order[i]]) return -1;
        seen[order[i]] = 1;
    }
    for (uint32_t i = 1; i < n; i++) {
        bool con = false;
        for (uint32_t j = 0; j < i && !con; j++) con = q_adj(q, order[i], order[j]);
        if (!con) return -2;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Intersection-cache plan (inc/code_gen.h:682-811).  For every depth `dep` and every later
// vertex, the set of its backward neighbours' depths that are <= dep is a candidate set op
// computed at depth `dep`.  Ops are ordered by (depth, label, size, lexicographic depths)
// with std::sort -- the same comparator as the reference, so that ops that tie keep the
// same relative order under the same standard library and the first copy of a duplicated
// op (which decides its v_depth, the failing-set source when it turns out empty) is the
// same one.  Duplicates (equal depth lists) are dropped, each op's parent is the op
// equal to its depth list minus the last element, and the iteration set of depth d is the
// op equal to all backward neighbours of order[d].
struct OpIR {
    uint32_t depth = SGM_INVALID_U32;
    uint32_t label = 0;
    uint32_t vdepth = SGM_INVALID_U32;
    std::vector<uint32_t> e;
    bool operator<(const OpIR& r) const
    {
        if (this == &r) return false;
        if (depth != r.depth) return depth < r.depth;
        if (label != r.label) return label < r.label;
        if (e.size() != r.e.size()) return e.size() < r.e.size();
        for (size_t i = 0; i < e.size(); i++)
            if (e[i] != r.e[i]) return e[i] < r.e[i];
        return false;
    }
    bool same(const OpIR& r) const { return label == r.label && e == r.e; }
};

int build_plan(const QueryCSR& q, const uint32_t* order, Plan& pl)
{
    const uint32_t n = q.n;
    if (n < 2 || n > 32) return -1;
    if (check_order(q, order)) return -2;
    std::vector<uint32_t> dep_of(n);
    for (uint32_t i = 0; i < n; i++) dep_of[order[i]] = i;
    std::vector<OpIR> vir(n);
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) vir[u].e.push_back(dep_of[q.adj[p]]);
        std::sort(vir[u].e.begin(), vir[u].e.end());
    }
    std::vector<OpIR> all, iter_tmp(n);
    for (uint32_t dep = 0; dep + 1 < n; dep++) {
        for (uint32_t fd = dep + 1; fd < n; fd++) {
            OpIR o;
            o.depth = dep;
            o.vdepth = fd;
            for (uint32_t x : vir[order[fd]].e)
                if (x <= dep) o.e.push_back(x);
            if (!o.e.empty()) all.push_back(o);
        }
        OpIR it;
        it.depth = dep;
        it.vdepth = dep + 1;
        for (uint32_t x : vir[order[dep + 1]].e)
            if (x <= dep) it.e.push_back(x);
        if (it.e.empty()) return -2;
        iter_tmp[dep + 1] = it;
    }
    std::sort(all.begin(), all.end());
    std::vector<OpIR> ops;
    for (size_t i = 0; i < all.size(); i++) {
        bool first = true;
        for (size_t j = 0; j < i && first; j++) first = !all[i].same(all[j]);
        if (first) ops.push_back(all[i]);
    }
    if (ops.size() > SGM_MAXOPS) return -3;
    pl.n = n;
    pl.order.assign(order, order + n);
    pl.ops.resize(ops.size());
    for (size_t i = 0; i < ops.size(); i++) {
        PlanOp& po = pl.ops[i];
        po.depth = ops[i].depth;
        po.vdepth = ops[i].vdepth;
        po.parent = SGM_INVALID_U32;
        po.edges = ops[i].e;
        if (i > 0) {
            OpIR pre = ops[i];
            pre.e.pop_back();
            for (size_t j = 0; j < i; j++)
                if (pre.same(ops[j])) { po.parent = (uint32_t)j; break; }
        }
    }
    pl.iter_set.assign(n, SGM_INVALID_U32);
    for (uint32_t d = 1; d < n; d++)
        for (size_t j = 0; j < ops.size(); j++)
            if (iter_tmp[d].same(ops[j])) { pl.iter_set[d] = (uint32_t)j; break; }
    // DAF ancestors over depths (inc/prune.h:100-104)
    pl.anc.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        uint64_t a = 1ull << i;
        const uint32_t u = order[i];
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) {
            const uint32_t j = dep_of[q.adj[p]];
            if (j < i) a |= pl.anc[j];
        }
        pl.anc[i] = a;
    }
    return 0;
}

void plan_to_device(const Plan& pl, sgm_plan_dev& d)
{
    std::memset(&d, 0, sizeof(d));
    d.n = pl.n;
    d.n_ops = (uint32_t)pl.ops.size();
    for (uint32_t i = 0; i < pl.n; i++) {
        d.order[i] = pl.order[i];
        d.iter_set[i] = pl.iter_set[i];
        d.anc[i] = pl.anc[i];
    }
    // ops are sorted by depth: op_first[d] = first op id whose depth >= d
    uint32_t o = 0;
    for (uint32_t dd = 0; dd <= pl.n; dd++) {
        while (o < d.n_ops && pl.ops[o].depth < dd) o++;
        d.op_first[dd] = o;
    }
    for (uint32_t i = 0; i < d.n_ops; i++) {
        d.op_parent[i] = pl.ops[i].parent == SGM_INVALID_U32 ? SGM_NOPARENT : (uint16_t)pl.ops[i].parent;
        d.op_vdepth[i] = (uint8_t)pl.ops[i].vdepth;
        d.op_depth[i] = (uint8_t)pl.ops[i].depth;
    }
}

// Backward neighbours in order position (inc/common_ops.h:259-313, GetPreviousNeb).
void previous_neighbours(const QueryCSR& q, const uint32_t* order, std::vector<std::vector<uint32_t>>& prev)
{
    const uint32_t n = q.n;
    std::vector<uint32_t> pos(n);
    for (uint32_t i = 0; i < n; i++) pos[order[i]] = i;
    prev.assign(n, {});
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++)
            if (pos[q.adj[p]] < pos[u]) prev[u].push_back(q.adj[p]);
        std::sort(prev[u].begin(), prev[u].end(), [&](uint32_t a, uint32_t b) { return pos[a] < pos[b]; });
    }
}

}  // namespace sgm
