// Algorithm-row driver over the drop-in C++ surface (include/sgm/*.h), with the reference
// harness's row semantics and CSV columns (src/overall_performance.cpp:52-97, 244-262,
// 281-312, 355): each row filters, orders and enumerates one query on a std::async thread,
// a timeout raises the stop flag, and one CSV line
//   data,query_name,algorithm,time_ms,ans,output_limit,intersection_cnt
// is printed per row.
//
// usage: sgm_overall DATA.txt QUERY.txt [output_limit=100000] [timeout_s=300] [rows...]
// rows: RI  QSI  GQL  CFL  CECI  DAF  RI+cache+filter+fs   (default: all seven)
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

// One algorithm row (src/overall_performance.cpp:52-127 for "RI", "QSI", "GQL", "CFL" and
// "CECI", :158-172 for "DAF", :243-262 for "RI+cache+filter+fs").  The promise is set once the search is about to start, so the
// timeout covers the search only, as in the reference.
std::string run_row(const Graph& data, const Query& query, const std::string& alg, std::atomic<bool>& stop_flag,
                    std::promise<void> started, uint64_t output_limit)
{
    double time = 0;
    size_t ans = 0;
    Profiler::getInst().reset();
    bool signalled = false;
    auto start = [&] { started.set_value(); signalled = true; };
    try {
        if (alg == "RI") {
            CandidateParam can(data, query);
            if (Filter_NLF(data, query, can)) {
                Order order;
                Order_RI(query, order);
                start();
                Config cfg;
                cfg.useFilter = true;
                cfg.can = &can;
                explore_Intersection_All<false, false, false, false, false>(data, query, order, cfg, time, ans,
                                                                            stop_flag, output_limit);
            }
        } else if (alg == "GQL") {
            CandidateParam can(data, query);
            if (Filter_GQL(data, query, can)) {
                Order order;
                Order_GQL(data, query, can, order);
                start();
                Config cfg;
                cfg.useFilter = true;
                cfg.can = &can;
                explore_Intersection_All<false, false, false, false, false>(data, query, order, cfg, time, ans,
                                                                            stop_flag, output_limit);
            }
        } else if (alg == "QSI" || alg == "CFL" || alg == "CECI" || alg == "DAF") {
            // the candidate-space rows of the other filters and orders; DAF adds failing sets
            CandidateParam can(data, query);
            const bool ok = alg == "QSI"    ? Filter_NLF(data, query, can)
                            : alg == "CFL"  ? Filter_CFL(data, query, can)
                            : alg == "CECI" ? Filter_CECI(data, query, can)
                                            : Filter_DAF(data, query, can);
            if (ok) {
                Order order;
                if (alg == "QSI") Order_QSI(data, query, can, order);
                else if (alg == "CFL") Order_CFL(data, query, can, order);
                else if (alg == "CECI") Order_CECI(data, query, order);
                else Order_RI(query, order);
                start();
                Config cfg;
                cfg.useFilter = true;
                cfg.can = &can;
                if (alg == "DAF")
                    explore_Intersection_All<true, false, false, false, false>(data, query, order, cfg, time, ans,
                                                                               stop_flag, output_limit);
                else
                    explore_Intersection_All<false, false, false, false, false>(data, query, order, cfg, time, ans,
                                                                                stop_flag, output_limit);
            }
        } else if (alg == "RI+cache+filter+fs") {
            Order order;
            Order_RI(query, order);
            CandidateParam can(data, query);
            if (Filter_GQL(data, query, can)) {
                start();
                Config cfg;
                cfg.useFilter = true;
                cfg.can = &can;
                ReuseParam reuse(query, order);
                cfg.useCache = true;
                cfg.reuse = &reuse;
                explore_Intersection_Cache_DAF_Filter(data, query, order, cfg, time, ans, stop_flag, output_limit);
            }
        } else {
            std::cerr << "row '" << alg << "' is not on the accelerated path\n";
        }
    } catch (...) {
        if (!signalled) start();
        throw;
    }
    if (!signalled) start();  // the reference's "goto Ret" path: empty candidates, ans 0
    std::ostringstream os;
    os << alg << ',' << fmt2(time) << ',' << ans << ',' << output_limit << ','
       << Profiler::getInst().total_intersection;
    return os.str();
}

std::string run_with_timeout(const Graph& data, const Query& query, const std::string& alg, uint64_t output_limit,
                             uint32_t timeout_s)
{
    std::atomic<bool> stop_flag(false);
    std::promise<void> started;
    std::future<void> started_f = started.get_future();
    auto fut = std::async(std::launch::async, [&, p = std::move(started)]() mutable {
        return run_row(data, query, alg, stop_flag, std::move(p), output_limit);
    });
    started_f.get();
    if (fut.wait_for(std::chrono::seconds(timeout_s)) == std::future_status::timeout) stop_flag.store(true);
    return fut.get();
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::cerr << "usage: " << argv[0] << " DATA.txt QUERY.txt [output_limit] [timeout_s] [rows...]\n";
        return 2;
    }
    const std::string data_path = argv[1], query_path = argv[2];
    const uint64_t output_limit = argc > 3 ? std::stoull(argv[3]) : 100000ull;
    const uint32_t timeout_s = argc > 4 ? (uint32_t)std::stoul(argv[4]) : 300u;
    std::vector<std::string> rows;
    for (int i = 5; i < argc; i++) rows.push_back(argv[i]);
    if (rows.empty()) rows = {"RI", "QSI", "GQL", "CFL", "CECI", "DAF", "RI+cache+filter+fs"};
    try {
        Graph data;
        data.Load(data_path);
        Query query;
        query.Load(query_path);
        std::cout << "data,query_name,algorithm,time_ms,ans,output_limit,intersection_cnt\n";
        for (const auto& alg : rows)
            std::cout << stem(data_path) << ',' << stem(query_path) << ','
                      << run_with_timeout(data, query, alg, output_limit, timeout_s) << '\n'
                      << std::flush;
    } catch (const sgm::Error& e) {
        std::cerr << "error " << e.code << ": " << e.what() << '\n';
        return 1;
    }
    return 0;
}
