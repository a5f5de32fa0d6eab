// Host-side planning helpers (orders, intersection-cache plan).
#pragma once
#include <algorithm>
#include <functional>
#include <thread>
#include <stdint.h>

#include <unordered_map>
#include <vector>

#include "sgm_internal.h"

#define SGM_INVALID_U32 0xFFFFFFFFu

namespace sgm {

struct QueryCSR {
    uint32_t n = 0;
    std::vector<uint32_t> off, adj;
    std::vector<uint32_t> lab;  // vertex labels; empty = every label 0
    uint32_t label(uint32_t u) const { return lab.empty() ? 0u : lab[u]; }
};

struct PlanOp {
    uint32_t depth, vdepth, parent;
    uint32_t emask;  // the op's backward-neighbour depths, as a set
};

struct Plan {
    uint32_t n = 0;
    std::vector<uint32_t> order;
    std::vector<PlanOp> ops;
    std::vector<uint32_t> iter_set;
    std::vector<uint64_t> anc;
};

void order_ri(const QueryCSR& q, uint32_t* order);

// f(i) for i in [0, n) on up to 8 host threads (about `grain` items per thread; the calling
// thread takes residue 0, and every residue whose thread could not be started).  f must
// not throw.
template <class F>
void parallel_for(uint32_t n, uint32_t grain, F&& f)
{
    uint32_t T = grain ? n / grain : 1u;
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    T = std::max(1u, std::min({T, 8u, hw}));
    std::vector<std::thread> th;
    uint32_t started = 1;
    try {
        th.reserve(T - 1);
        for (uint32_t t = 1; t < T; t++, started++)
            th.emplace_back([&f, n, T, t] {
                for (uint32_t i = t; i < n; i += T) f(i);
            });
    } catch (...) {
        // out of threads: the residues from `started` on run here
    }
    for (uint32_t t = 0; t < T; t++)
        if (t == 0 || t >= started)
            for (uint32_t i = t; i < n; i += T) f(i);
    for (auto& x : th) x.join();
}
void order_gql(const QueryCSR& q, const uint32_t* cand_cnt, uint32_t num_data_vertices, uint32_t* order);
int check_order(const QueryCSR& q, const uint32_t* order);
int build_plan(const QueryCSR& q, const uint32_t* order, Plan& pl);
void plan_to_device(const Plan& pl, sgm_plan_dev& d);
void previous_neighbours(const QueryCSR& q, const uint32_t* order, std::vector<std::vector<uint32_t>>& prev);


// ---- BFS-tree filters and the QSI / CFL / CECI / GuP orders (sgm_tree.cpp)
// filter kinds of the refinement programs (the C ABI's SGM_FILTER_CFL / _CECI / _DAF)
enum { SGM_TREE_CFL = 3, SGM_TREE_CECI = 4, SGM_TREE_DAF = 5 };

struct BfsTree {
    std::vector<uint32_t> order;
    uint32_t parent[SGM_MAXQ] = {0};
    uint32_t level[SGM_MAXQ] = {0};
    std::vector<uint32_t> children[SGM_MAXQ];
};

// Data-graph label statistics (Graph::getLabelFreq / getMaxLabelFreq, inc/graph.h:339-345).
struct LabelStats {
    uint32_t V = 0;
    const std::unordered_map<uint32_t, uint32_t>* hist = nullptr;  // null: every label 0
    uint32_t max_freq = 0;
    uint32_t freq(uint32_t l) const
    {
        if (!hist) return l == 0 ? V : 0;
        auto it = hist->find(l);
        return it == hist->end() ? 0 : it->second;
    }
};

// Estimated embeddings along a root-to-leaf path of aux tables
// (estimatePathEmbeddsingsNum, inc/order.h:197-239): est[i] for i < path.size() - 1.
using PathEstimator = std::function<int(const std::vector<uint32_t>& path, std::vector<uint64_t>& est)>;
// count_embed (inc/order.h:386-410) from ordered vertex u: fills nembeds[c] of every
// unordered vertex c in u's BFS subtree reached through unordered tree children.
using CountEmbed = std::function<int(uint32_t u, const uint8_t* ordered, const BfsTree& t, double* nembeds)>;

uint32_t query_kcore(const QueryCSR& q, int32_t* core);
void bfs_tree(const QueryCSR& q, uint32_t start, BfsTree& t);
void order_neighbours(const QueryCSR& q, const std::vector<uint32_t>& order, bool following,
                      std::vector<uint32_t> (&nb)[SGM_MAXQ]);
uint32_t cfl_start_vertex(const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls);
uint32_t ceci_daf_start_vertex(const QueryCSR& q, const uint32_t* nlf_cnt, uint32_t V, bool daf);
int refine_program(int kind, const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls, uint32_t* prog);
void order_qsi(const QueryCSR& q, const uint32_t* ecnt, uint32_t* order);
int order_cfl(const QueryCSR& q, const uint32_t* nlf_cnt, const LabelStats& ls, const uint32_t* cnt,
              const PathEstimator& estimate, uint32_t* order);
// -2 where the reference asserts (a query whose vertices all have degree <= 2)
int order_gup(const QueryCSR& q, const uint32_t* nlf_cnt, const uint32_t* cnt, const uint32_t* ecnt,
              const CountEmbed& count_embed, uint32_t* order);

}  // namespace sgm
