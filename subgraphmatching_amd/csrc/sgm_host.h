// Host-side planning helpers (orders, intersection-cache plan).
#pragma once
#include <algorithm>
#include <thread>
#include <stdint.h>

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

}  // namespace sgm
