// Host-side planning helpers (orders, intersection-cache plan).
#pragma once
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
    std::vector<uint32_t> edges;
};

struct Plan {
    uint32_t n = 0;
    std::vector<uint32_t> order;
    std::vector<PlanOp> ops;
    std::vector<uint32_t> iter_set;
    std::vector<uint64_t> anc;
};

void order_ri(const QueryCSR& q, uint32_t* order);
void order_gql(const QueryCSR& q, const uint32_t* cand_cnt, uint32_t num_data_vertices, uint32_t* order);
int check_order(const QueryCSR& q, const uint32_t* order);
int build_plan(const QueryCSR& q, const uint32_t* order, Plan& pl);
void plan_to_device(const Plan& pl, sgm_plan_dev& d);
void previous_neighbours(const QueryCSR& q, const uint32_t* order, std::vector<std::vector<uint32_t>>& prev);

}  // namespace sgm
