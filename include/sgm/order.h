// Matching orders (inc/order.h:563-597 Order_GQL, :1138-1281 Order_RI), drop-in over the
// C ABI (both are host code in the library).
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
