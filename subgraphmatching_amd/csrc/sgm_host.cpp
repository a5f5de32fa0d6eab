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
    // queries have at most 32 vertices: neighbourhoods and the visited set are 32-bit masks
    const uint32_t n = q.n;
    uint32_t nbr[32] = {0};
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) nbr[u] |= 1u << q.adj[p];
    order[0] = 0;
    for (uint32_t i = 1; i < n; i++)
        if (q_deg(q, i) > q_deg(q, order[0])) order[0] = i;
    uint32_t vis = 1u << order[0];
    uint32_t ties[32], prev[32];
    for (uint32_t i = 1; i < n; i++) {
        // 1. most neighbours among the ordered vertices
        uint32_t nt = 0, best = 0;
        for (uint32_t u = 0; u < n; u++) {
            if ((vis >> u) & 1u) continue;
            const uint32_t bn = (uint32_t)__builtin_popcount(nbr[u] & vis);
            if (bn > best) { best = bn; nt = 0; }
            if (bn == best) ties[nt++] = u;
        }
        // 2. most ordered vertices sharing an unordered neighbour with u
        if (nt != 1) {
            const uint32_t np = nt;
            std::copy(ties, ties + nt, prev);
            nt = 0;
            best = 0;
            for (uint32_t k = 0; k < np; k++) {
                const uint32_t u = prev[k], fn = nbr[u] & ~vis;
                uint32_t cc = 0;
                for (uint32_t m = vis; m; m &= m - 1) cc += (nbr[__builtin_ctz(m)] & fn) != 0u;
                if (cc > best) { best = cc; nt = 0; }
                if (cc == best) ties[nt++] = u;
            }
        }
        // 3. most unordered neighbours not adjacent to any ordered vertex
        if (nt != 1) {
            const uint32_t np = nt;
            std::copy(ties, ties + nt, prev);
            nt = 0;
            best = 0;
            for (uint32_t k = 0; k < np; k++) {
                const uint32_t u = prev[k];
                uint32_t cc = 0;
                for (uint32_t m = nbr[u] & ~vis; m; m &= m - 1) cc += (nbr[__builtin_ctz(m)] & vis) == 0u;
                if (cc > best) { best = cc; nt = 0; }
                if (cc == best) ties[nt++] = u;
            }
        }
        order[i] = ties[0];
        vis |= 1u << order[i];
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
// computed at depth `dep`, labelled with the follow vertex's label (VertexSetIR::_v_label,
// :716).  Ops are ordered by (depth, label, size, lexicographic depths)
// with std::sort -- the same comparator as the reference, so that ops that tie keep the
// same relative order under the same standard library and the first copy of a duplicated
// op (which decides its v_depth, the failing-set source when it turns out empty) is the
// same one.  Ops are equal when label and depth list are (:652-654).  Duplicates are
// dropped, each op's parent is the op with its label and its depth list minus the last
// element (:777-787), and the iteration set of depth d is the op equal to order[d]'s label
// with all its backward neighbours (:800-808).
struct OpIR {
    uint32_t depth = SGM_INVALID_U32;
    uint32_t label = 0;
    uint32_t vdepth = SGM_INVALID_U32;
    uint32_t e = 0;  // the depth list as a set: depths are distinct and < 32
    // depth, label, list length, then the ascending lists lexicographically.  For two sets of
    // equal size the first differing position holds the smaller of the two lowest
    // non-shared elements, i.e. the lowest bit of e ^ r.e, and the list holding it is less.
    bool operator<(const OpIR& r) const
    {
        if (this == &r) return false;
        if (depth != r.depth) return depth < r.depth;
        if (label != r.label) return label < r.label;
        const int sa = __builtin_popcount(e), sb = __builtin_popcount(r.e);
        if (sa != sb) return sa < sb;
        const uint32_t x = e ^ r.e;
        return x != 0u && ((e >> __builtin_ctz(x)) & 1u) != 0u;
    }
};

int build_plan(const QueryCSR& q, const uint32_t* order, Plan& pl)
{
    const uint32_t n = q.n;
    if (n < 2 || n > 32) return -1;
    if (check_order(q, order)) return -2;
    uint32_t dep_of[32], back[32];  // back[u]: depths of u's neighbours, as a set
    for (uint32_t i = 0; i < n; i++) dep_of[order[i]] = i;
    for (uint32_t u = 0; u < n; u++) {
        back[u] = 0;
        for (uint32_t p = q.off[u]; p < q.off[u + 1]; p++) back[u] |= 1u << dep_of[q.adj[p]];
    }
    OpIR all[32 * 31 / 2];  // at most one op per (depth, later vertex)
    uint32_t na = 0;
    uint32_t iter_e[32] = {0};
    for (uint32_t dep = 0; dep + 1 < n; dep++) {
        const uint32_t upto = dep == 31 ? ~0u : ((2u << dep) - 1u);  // depths <= dep
        for (uint32_t fd = dep + 1; fd < n; fd++) {
            OpIR o;
            o.depth = dep;
            o.label = q.label(order[fd]);
            o.vdepth = fd;
            o.e = back[order[fd]] & upto;
            if (o.e) all[na++] = o;
        }
        iter_e[dep + 1] = back[order[dep + 1]] & upto;
        if (!iter_e[dep + 1]) return -2;
    }
    std::sort(all, all + na);
    // duplicates (equal label and set) keep the first in sorted order; parents and iteration
    // sets are then one probe each of an open-addressing table keyed by (label, edge set)
    // (edge sets are never empty, so key & 0xFFFFFFFF == 0 marks a free slot)
    OpIR ops[32 * 31 / 2];
    uint32_t nops = 0;
    constexpr uint32_t kSlots = 2048, kBits = 11;
    uint64_t tkey[kSlots];
    uint32_t tval[kSlots];
    std::fill(tkey, tkey + kSlots, 0ull);
    auto key = [](uint32_t label, uint32_t e) { return ((uint64_t)label << 32) | e; };
    auto slot = [&](uint64_t k) {
        uint32_t h = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - kBits));
        while ((uint32_t)tkey[h] != 0u && tkey[h] != k) h = (h + 1) & (kSlots - 1);
        return h;
    };
    auto find = [&](uint64_t k) -> uint32_t {
        if ((uint32_t)k == 0u) return SGM_INVALID_U32;
        const uint32_t h = slot(k);
        return tkey[h] == k ? tval[h] : SGM_INVALID_U32;
    };
    for (uint32_t a = 0; a < na; a++) {
        const uint64_t k = key(all[a].label, all[a].e);
        const uint32_t h = slot(k);
        if (tkey[h] == k) continue;
        tkey[h] = k;
        tval[h] = nops;
        ops[nops++] = all[a];
    }
    if (nops > SGM_MAXOPS) return -3;
    pl.n = n;
    pl.order.assign(order, order + n);
    pl.ops.resize(nops);
    for (uint32_t i = 0; i < nops; i++) {
        PlanOp& po = pl.ops[i];
        const uint32_t m = ops[i].e;
        po.depth = ops[i].depth;
        po.vdepth = ops[i].vdepth;
        po.emask = m;
        po.parent = SGM_INVALID_U32;
        // the parent computes the list without its deepest edge
        const uint32_t p = find(key(ops[i].label, m & ~(0x80000000u >> __builtin_clz(m))));
        if (p != SGM_INVALID_U32 && p < i) po.parent = p;
    }
    pl.iter_set.assign(n, SGM_INVALID_U32);
    for (uint32_t d = 1; d < n; d++) pl.iter_set[d] = find(key(q.label(order[d]), iter_e[d]));
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
