// Drop-in C++ surface over the C ABI (sgm_c.h): shared types, the process-wide device
// context and the per-thread Profiler.
//
// The reference defines these in inc/common_type.h (VertexID, Order :38-48, Profiler
// :150-226).  Everything here is header-only and links against libsgm_hip.so.
#pragma once

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../sgm_c.h"

typedef uint32_t VertexID;
typedef uint32_t LabelID;
typedef std::vector<VertexID> Order;
typedef std::pair<VertexID, VertexID> PartialOrder;  // inc/common_type.h:41-42
typedef std::vector<PartialOrder> Restriction;

namespace sgm {

// Errors the reference reports with exit(-1) (inc/api.h:211-225, inc/graph.h:194-197,
// inc/code_gen.h:743-746) are thrown instead.
class Error : public std::runtime_error {
public:
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// One device context per process (device from SGM_DEVICE, default 0), created on first
// use.  Device calls are serialised by device_mutex(); sgm_stop() needs no lock.
inline sgm_ctx* context()
{
    static sgm_ctx* ctx = [] {
        const char* env = std::getenv("SGM_DEVICE");
        sgm_ctx* c = nullptr;
        const int rc = sgm_ctx_create(env ? std::atoi(env) : 0, &c);
        if (rc != SGM_OK) throw Error(rc, "sgm: cannot create a device context (no usable GPU?)");
        return c;
    }();
    return ctx;
}

inline std::mutex& device_mutex()
{
    static std::mutex m;
    return m;
}

inline void check(int rc, const char* what)
{
    if (rc != SGM_OK) throw Error(rc, std::string(what) + ": " + sgm_ctx_last_error(context()));
}

}  // namespace sgm

// Per-thread search counters (inc/common_type.h:150-226): the fields the accelerated path
// fills are total_intersection and total_iter_count (extensions per depth); daf_fp.cnt
// counts failing-set cut-offs (and the guard prunes too when both prunings run), gup_cd.cnt
// GuP's conflict-detection prunes.
class Profiler {
public:
    Profiler(const Profiler&) = delete;
    Profiler& operator=(const Profiler&) = delete;
    static bool constexpr useProfiling = true;

    static Profiler& getInst()
    {
        static thread_local Profiler inst;
        return inst;
    }

    struct PruneStatistics {
        uint32_t cnt{0};
        size_t mem_overhead{0};
        double mem_overhead_KB{0};
        void reset() { cnt = 0; mem_overhead = 0; mem_overhead_KB = 0.0; }
    };
    PruneStatistics daf_fp;
    PruneStatistics gup_cd;

    size_t total_intersection{0};
    size_t total_iter_count[300];

    void reset()
    {
        total_intersection = 0;
        daf_fp.reset();
        gup_cd.reset();
        std::memset(total_iter_count, 0, sizeof total_iter_count);
    }

private:
    Profiler() { reset(); }
};
