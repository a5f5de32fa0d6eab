// Matching orders (inc/order.h:537-561 Order_GuP, :563-597 Order_GQL, :599-607 Order_CECI,
// :610-668 Order_QSI, :781-976 Order_CFL, :1138-1281 Order_RI) and the DAF weight array of the
// dynamic-order search (:978-1031 computeWeightArray), drop-in over the C ABI.  The
// rules are host code in the library; the orders over a candidate space read its NLF counts
// and aux tables on the device.  Order_GuP throws sgm::Error(SGM_EINVAL) where the reference
// asserts (a query whose vertices all have degree <= 2).
#pragma once

#include "api.h"

inline void Order_RI(const Graph& query, Order& order)
{
    const uint32_t n = query.getVertexCnt();
    order.assign(n, 0);
    sgm::check(sgm_order_ri(query.offsets(), query.neighbours(), n, order.data()), "Order_RI");
}

inline void Order_GQL(const Graph& data, const Query& query, const CandidateParam& can, Order& order)
{
    (void)data;
    order.assign(query.getVertexCnt(), 0);
    std::lock_guard<std::mutex> lk(sgm::device_mutex());
    sgm::check(sgm_order_gql(can.device(), order.data()), "Order_GQL");
}

namespace sgm {
inline void cand_order(int (*fn)(sgm_query*, uint32_t*), const Query& query, const CandidateParam& can, Order& order,
                       const char* what)
{
    order.assign(query.getVertexCnt(), 0);
    std::lock_guard<std::mutex> lk(device_mutex());
    check(fn(can.device(), order.data()), what);
}
}  // namespace sgm

inline void Order_QSI(const Graph& data, const Query& query, const CandidateParam& can, Order& order)
{
    (void)data;
    sgm::cand_order(sgm_order_qsi, query, can, order, "Order_QSI");
}

inline void Order_CFL(const Graph& data, const Query& query, const CandidateParam& can, Order& order)
{
    (void)data;
    sgm::cand_order(sgm_order_cfl, query, can, order, "Order_CFL");
}

inline void Order_GuP(const Graph& data, const Query& query, const CandidateParam& can, Order& order)
{
    (void)data;
    sgm::cand_order(sgm_order_gup, query, can, order, "Order_GuP");
}

// Order_CECI (inc/order.h:599-607) takes no candidate space: a device query object is made
// for the call (its NLF counts pick the start vertex).
inline void Order_CECI(const Graph& data, const Query& query, Order& order)
{
    const uint32_t n = query.getVertexCnt();
    order.assign(n, 0);
    sgm_graph* g = data.device();  // takes device_mutex itself: before the lock below
    std::lock_guard<std::mutex> lk(sgm::device_mutex());
    sgm_query* q = nullptr;
    sgm::check(sgm_query_create(g, query.offsets(), query.neighbours(), n, &q), "query create");
    int rc = query.unlabeled() ? SGM_OK : sgm_query_set_labels(q, query.getVertexLabels(), n);
    if (rc == SGM_OK) rc = sgm_order_ceci(q, order.data());
    sgm_query_destroy(q);
    sgm::check(rc, "Order_CECI");
}

// DAFStartVertex (inc/filter.h:922-940): the start of the BFS order of both dynamic searches.
// As for Order_CECI, a device query object is made for the call (its NLF counts decide).
inline VertexID DAFStartVertex(const Graph& data, const Query& query)
{
    const uint32_t n = query.getVertexCnt();
    sgm_graph* g = data.device();  // takes device_mutex itself: before the lock below
    std::lock_guard<std::mutex> lk(sgm::device_mutex());
    sgm_query* q = nullptr;
    sgm::check(sgm_query_create(g, query.offsets(), query.neighbours(), n, &q), "query create");
    int rc = query.unlabeled() ? SGM_OK : sgm_query_set_labels(q, query.getVertexLabels(), n);
    uint32_t s = 0;
    if (rc == SGM_OK) rc = sgm_daf_start_vertex(q, &s);
    sgm_query_destroy(q);
    sgm::check(rc, "DAFStartVertex");
    return s;
}

// computeWeightArray (inc/order.h:978-1031): bfs_order gets the BFS order from DAFStartVertex
// and weight_array[u] (caller-allocated, candidates_count[u] entries, as the reference's
// driver allocates it, src/order_ablations.cpp:133-138) the weights, computed on the device
// where they also stay for explore_Intersection_DynamicOrder_DAF.
inline void computeWeightArray(const Graph& data, const Query& query, const CandidateParam& canParam,
                               Order& bfs_order, uint32_t**& weight_array)
{
    (void)data;
    bfs_order.assign(query.getVertexCnt(), 0);
    std::lock_guard<std::mutex> lk(sgm::device_mutex());
    sgm::check(sgm_daf_weights(canParam.device(), bfs_order.data(), weight_array), "computeWeightArray");
}
