// Enumeration entry points (inc/explore.h:1073-1129, :1849-2062 and :2757-2848), drop-in over
// the C ABI.
//
//   explore_Intersection_Cache_DAF_Filter  intersection cache + candidate filter + DAF
//                                          failing sets (the "RI+cache+filter+fs" row)
//   explore_Intersection_All<false x5>     plain candidate-space search over the aux-edge
//                                          tables (the "RI" / "GQL" rows)
//   explore_Intersection_All<true, false x4>  the same with DAF failing sets
//   explore_Intersection_All<f, false, true, false, false>  with GuP's conflict detection
//                                          (reservation guards computed on the host)
//   explore_Intersection_DynamicOrder_DAF  the next vertex chosen per search node, by DAF
//   explore_Intersection_DynamicOrder_VEQ  weights / by candidate counts (the "DAF" and
//                                          "VEQ" rows of the order ablation)
//
// Semantics kept from the reference: `ans` is incremented by min(#embeddings,
// output_limit); `time` is the search time in ms (here the enumeration kernel's, measured
// with HIP events, where the reference reads the thread CPU clock); Profiler's
// total_intersection and total_iter_count are incremented; stop_flag is watched while the
// search runs and a raised flag ends it early with the partial count, as the reference's
// polling does (inc/explore.h:1057).  The other explore_Intersection_All pruning variants
// (GuP, conflict detection, VEQ) are not on the accelerated path and fail to compile.
#pragma once

#include "api.h"

namespace sgm {

inline void run_enumeration(const Query& query, const Order& order, const Config& cfg, double& time, size_t& ans,
                            std::atomic<bool>& stop_flag, uint64_t output_limit, int method,
                            uint32_t* const* weights = nullptr)
{
    cfg.Check();
    if (cfg.can == nullptr)
        throw Error(SGM_EINVAL, "the accelerated search needs filtered candidates (cfg.useFilter and cfg.can)");
    if (order.size() != query.getVertexCnt()) throw Error(SGM_EORDER, "order length differs from the query size");
    sgm_enum_opts opts;
    std::memset(&opts, 0, sizeof opts);
    opts.cap = output_limit;
    opts.method = method;
    sgm_stats st;
    std::memset(&st, 0, sizeof st);
    std::atomic<bool> done(false);
    // the watcher repeats the stop request while the flag stays raised, so a request that
    // lands before the launch has reset the device stop word is not lost
    std::thread watcher([&] {
        while (!done.load(std::memory_order_acquire)) {
            if (stop_flag.load()) sgm_stop(context());
            std::this_thread::sleep_for(std::chrono::milliseconds(stop_flag.load() ? 5 : 1));
        }
    });
    int rc;
    {
        std::lock_guard<std::mutex> lk(device_mutex());
        rc = weights ? sgm_set_daf_weights(cfg.can->device(), order.data(), weights) : SGM_OK;
        if (rc == SGM_OK) rc = sgm_enumerate(cfg.can->device(), order.data(), &opts, &st);
    }
    done.store(true, std::memory_order_release);
    watcher.join();
    if (rc != SGM_OK && rc != SGM_ESTOPPED) check(rc, "enumerate");
    ans += (size_t)st.count;
    time = st.kernel_ms;
    Profiler& prof = Profiler::getInst();
    prof.total_intersection += (size_t)st.intersections;
    // the dynamic searches count no iterations (their vertex per depth varies)
    if (method != SGM_ENUM_DYN_DAF && method != SGM_ENUM_DYN_VEQ)
        for (uint32_t d = 0; d < query.getVertexCnt() && d < 32; d++) prof.total_iter_count[d] += (size_t)st.nodes[d];
    // the prunes of the method's pruning (DAF cut-offs and / or GuP guard prunes)
    if (method == SGM_ENUM_AUX_TABLES_GUPCD) prof.gup_cd.cnt += (uint32_t)st.fs_prunes;
    else prof.daf_fp.cnt += (uint32_t)st.fs_prunes;
}

}  // namespace sgm

inline void explore_Intersection_Cache_DAF_Filter(const Graph& data, const Query& query, const Order& order,
                                                  const Config& cfg, double& time, size_t& ans,
                                                  std::atomic<bool>& stop_flag, uint64_t output_limit = -1)
{
    (void)data;
    sgm::run_enumeration(query, order, cfg, time, ans, stop_flag, output_limit, SGM_ENUM_CACHE_FILTER_FS);
}

template <bool useDAFFailingPrune, bool useGuPFailingPrune, bool useGUPConflictDetection, bool useBICEConflictDetection,
          bool useVEQAutomorphismPrune>
inline void explore_Intersection_All(const Graph& data, const Query& query, const Order& order, const Config& cfg,
                                     double& time, size_t& ans, std::atomic<bool>& stop_flag,
                                     uint64_t output_limit = -1)
{
    static_assert(!useGuPFailingPrune && !useBICEConflictDetection && !useVEQAutomorphismPrune,
                  "only explore_Intersection_All<false|true,false,false|true,false,false> (plain, DAF failing "
                  "sets, GuP conflict detection) is on the accelerated path");
    (void)data;
    sgm::run_enumeration(query, order, cfg, time, ans, stop_flag, output_limit,
                         useGUPConflictDetection
                             ? (useDAFFailingPrune ? SGM_ENUM_AUX_TABLES_FS_GUPCD : SGM_ENUM_AUX_TABLES_GUPCD)
                             : (useDAFFailingPrune ? SGM_ENUM_AUX_TABLES_FS : SGM_ENUM_AUX_TABLES));
}

// explore_Intersection_DynamicOrder_DAF (inc/explore.h:1849-1931): bfs_order is the start
// order (computeWeightArray's), weight_array the weights ranked by (the caller's are used).
// The reference writes match_order past its cleared size, so it comes back empty.
inline void explore_Intersection_DynamicOrder_DAF(const Graph& data, const Query& query, const Order& bfs_order,
                                                  Order& match_order, const Config& cfg, uint32_t** weight_array,
                                                  double& time, size_t& ans, std::atomic<bool>& stop_flag,
                                                  uint64_t output_limit = -1)
{
    (void)data;
    if (weight_array == nullptr) throw sgm::Error(SGM_EINVAL, "weight_array is null");
    match_order.clear();
    sgm::run_enumeration(query, bfs_order, cfg, time, ans, stop_flag, output_limit, SGM_ENUM_DYN_DAF, weight_array);
}

// explore_Intersection_DynamicOrder_VEQ (inc/explore.h:1993-2062): ranked by candidate counts.
inline void explore_Intersection_DynamicOrder_VEQ(const Graph& data, const Query& query, const Order& bfs_order,
                                                  Order& match_order, const Config& cfg, double& time, size_t& ans,
                                                  std::atomic<bool>& stop_flag, uint64_t output_limit = -1)
{
    (void)data;
    match_order.clear();
    sgm::run_enumeration(query, bfs_order, cfg, time, ans, stop_flag, output_limit, SGM_ENUM_DYN_VEQ);
}
