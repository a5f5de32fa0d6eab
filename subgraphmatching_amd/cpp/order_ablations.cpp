// Order-comparison driver over the drop-in C++ surface (include/sgm/*.h), with the reference
// harness's row semantics and CSV columns (src/order_ablations.cpp:43-185, 281-305, 343-357,
// 379): every row filters the query with GQL, then orders it and searches the candidate
// space with explore_Intersection_All<false x5> (static orders), or runs one of the
// dynamic-order searches ("DAF": computeWeightArray + explore_Intersection_DynamicOrder_DAF;
// "VEQ": DAFStartVertex + BFS + explore_Intersection_DynamicOrder_VEQ).  The order's host
// time is order_overhead; a timeout raises the stop flag.  One CSV line
//   data,query_name,explore,order,order_overhead,time_ms,ans,output_limit,intersection_cnt
// per row, on stdout.
//
// usage: sgm_orders DATA.txt QUERY.txt [output_limit=100000] [timeout_s=300] [orders...]
// orders: QSI  RI  GQL  CFL  CECI  DAF  VEQ   (default: all seven, the "filter" rows of the
// reference's list on the accelerated path; VF2PP, VF3, GraphPi, GraphZero and the "cache"
// rows are not)
#include <chrono>
#include <cstdio>
#include <future>
#include <iostream>
#include <sstream>

#include "../../include/sgm/explore.h"
#include "../../include/sgm/filter.h"
#include "../../include/sgm/order.h"

namespace {

std::string stem(const std::string& path)
{
    const size_t s = path.find_last_of('/');
    std::string base = s == std::string::npos ? path : path.substr(s + 1);
    const size_t d = base.rfind('.');
    return d == std::string::npos ? base : base.substr(0, d);
}

std::string fmt2(double x)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.2f", x);
    return buf;
}

double ms_since(std::chrono::high_resolution_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
}

// order_compare_filter (src/order_ablations.cpp:43-185)
std::string run_row(const Graph& data, const Query& query, const std::string& order_type,
                    std::atomic<bool>& stop_flag, uint64_t output_limit)
{
    double time = 0, order_overhead = 0;
    size_t ans = 0;
    Profiler::getInst().reset();
    CandidateParam can(data, query);
    if (Filter_GQL(data, query, can)) {
        Config cfg;
        cfg.useFilter = true;
        cfg.can = &can;
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (order_type == "DAF") {
            const uint32_t n = query.getVertexCnt();
            uint32_t** weight_array = new uint32_t*[n];
            for (uint32_t i = 0; i < n; ++i) {
                weight_array[i] = new uint32_t[can.candidates_count[i] ? can.candidates_count[i] : 1];
                std::fill(weight_array[i], weight_array[i] + can.candidates_count[i], 0xFFFFFFFFu);
            }
            Order bfs_order, match_order;
            computeWeightArray(data, query, can, bfs_order, weight_array);
            order_overhead = ms_since(t0);
            explore_Intersection_DynamicOrder_DAF(data, query, bfs_order, match_order, cfg, weight_array, time, ans,
                                                  stop_flag, output_limit);
            for (uint32_t i = 0; i < n; ++i) delete[] weight_array[i];
            delete[] weight_array;
        } else if (order_type == "VEQ") {
            Order bfs_order, match_order;
            const VertexID start_vertex = DAFStartVertex(data, query);
            TreeNode* tree = nullptr;
            BFS(query, start_vertex, tree, bfs_order);
            order_overhead = ms_since(t0);
            explore_Intersection_DynamicOrder_VEQ(data, query, bfs_order, match_order, cfg, time, ans, stop_flag,
                                                  output_limit);
            delete[] tree;
        } else {
            Order order;
            if (order_type == "GQL") Order_GQL(data, query, can, order);
            else if (order_type == "CECI") Order_CECI(data, query, order);
            else if (order_type == "CFL") Order_CFL(data, query, can, order);
            else if (order_type == "QSI") Order_QSI(data, query, can, order);
            else if (order_type == "RI") Order_RI(query, order);
            else throw sgm::Error(SGM_EINVAL, "order '" + order_type + "' is not on the accelerated path");
            order_overhead = ms_since(t0);
            explore_Intersection_All<false, false, false, false, false>(data, query, order, cfg, time, ans, stop_flag,
                                                                        output_limit);
        }
    }
    std::ostringstream os;
    os << "filter," << order_type << ',' << fmt2(order_overhead) << ',' << fmt2(time) << ',' << ans << ','
       << output_limit << ',' << Profiler::getInst().total_intersection;
    return os.str();
}

// run_order_with_timeout (src/order_ablations.cpp:281-305)
std::string run_with_timeout(const Graph& data, const Query& query, const std::string& order_type,
                             uint64_t output_limit, uint32_t timeout_s)
{
    std::atomic<bool> stop_flag(false);
    auto fut = std::async(std::launch::async, [&]() { return run_row(data, query, order_type, stop_flag, output_limit); });
    if (fut.wait_for(std::chrono::seconds(timeout_s)) == std::future_status::timeout) stop_flag.store(true);
    return fut.get();
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::cerr << "usage: " << argv[0] << " DATA.txt QUERY.txt [output_limit] [timeout_s] [orders...]\n";
        return 2;
    }
    const std::string data_path = argv[1], query_path = argv[2];
    const uint64_t output_limit = argc > 3 ? std::stoull(argv[3]) : 100000ull;
    const uint32_t timeout_s = argc > 4 ? (uint32_t)std::stoul(argv[4]) : 300u;
    std::vector<std::string> orders;
    for (int i = 5; i < argc; i++) orders.push_back(argv[i]);
    if (orders.empty()) orders = {"QSI", "RI", "GQL", "CFL", "CECI", "DAF", "VEQ"};
    try {
        Graph data;
        data.Load(data_path);
        Query query;
        query.Load(query_path);
        std::cout << "data,query_name,explore,order,order_overhead,time_ms,ans,output_limit,intersection_cnt\n";
        for (const auto& o : orders)
            std::cout << stem(data_path) << ',' << stem(query_path) << ','
                      << run_with_timeout(data, query, o, output_limit, timeout_s) << '\n'
                      << std::flush;
    } catch (const sgm::Error& e) {
        std::cerr << "error " << e.code << ": " << e.what() << '\n';
        return 1;
    }
    return 0;
}
