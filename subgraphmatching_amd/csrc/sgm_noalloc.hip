// sgm_enumerate_noalloc: the cache + filter (+ failing-set) search of one query with every
// per-call device buffer taken from one static device array, so the call allocates no
// device or pinned memory (for callers that cannot: code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The array (`scratch`, kScratchBytes per GPU, held until the process ends) is laid out per
// call as: plan | query descriptor | statistics | stop / item / chunk words | depth-1 split
// block | chunk records | bitmap of C(u_L) | set-op stack (the rest).  In host code the
// array's name is not a device address: its address on each device comes from
// hipGetSymbolAddress, looked up once per device when a context is created there.  Every
// call on one GPU shares that GPU's array: calls that overlap on one GPU race; each call
// clears what must start at zero (with fill kernels, not memsets).
#include <hip/hip_runtime.h>

#include <mutex>

#include "sgm_objects.h"

namespace {

constexpr size_t kScratchBytes = 256ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};  // the array's address on each device (null: not looked up)

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Looks up the array's address on `device` (once per device; the device must be current).
int noalloc_setup(sgm_ctx* c)
{
    if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, SGM_EINVAL, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return SGM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(c, SGM_EHIP, "scratch array address: %s", hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<uint8_t*>(p);
    return SGM_OK;
}

extern "C" int sgm_enumerate_noalloc(sgm_query* q, const uint32_t* order, const sgm_enum_opts* opts, sgm_stats* out)
{
    if (!q || !order || !out) return SGM_EINVAL;
    sgm_graph* g = q->g;
    sgm_ctx* c = g->ctx;
    std::memset(out, 0, sizeof *out);
    const sgm_enum_opts o = normalize_opts(opts);
    if (o.method != SGM_ENUM_CACHE_FILTER_FS && o.method != SGM_ENUM_CACHE_FILTER)
        return set_err(c, SGM_EINVAL, "the no-allocation entry point runs the cache search methods (got %d)", o.method);
    if (o.hub_min_degree > kHubMinDegree)
        return set_err(c, SGM_EINVAL, "the no-allocation entry point uses the graph's own hub bitmaps (hub_min_degree "
                                      "<= %u)", kHubMinDegree);
    if (const int frc = check_filtered(q, "enumerate")) return frc;
    if (q->q.n < 2) return set_err(c, SGM_EINVAL, "enumeration needs at least 2 query vertices");
    uint8_t* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
    if (!base) return set_err(c, SGM_ESTATE, "scratch array not set up on device %d", c->device);
    sgm::Plan pl;
    const int prc = sgm::build_plan(q->q, order, pl);
    if (prc == -2) return set_err(c, SGM_EORDER, "order is not a connected permutation");
    if (prc) return set_err(c, SGM_EINVAL, "plan construction failed (%d)", prc);
    if (q->empty) return SGM_OK;  // the reference's driver reports 0 without enumerating
    sgm_plan_dev plan;
    sgm::plan_to_device(pl, plan);

    // layout of this call in the array
    const uint32_t V = g->V;
    const size_t lb_words = ((size_t)V + 31) / 32;
    size_t at = 0;
    auto take = [&](size_t bytes) { const size_t p = at; at += up256(bytes); return base + p; };
    sgm_plan_dev* d_plan = reinterpret_cast<sgm_plan_dev*>(take(sizeof(sgm_plan_dev)));
    sgm_equery* d_eq = reinterpret_cast<sgm_equery*>(take(sizeof(sgm_equery)));
    sgm_stats_dev* d_st = reinterpret_cast<sgm_stats_dev*>(take(sizeof(sgm_stats_dev)));
    uint32_t* d_words = reinterpret_cast<uint32_t*>(take(3 * sizeof(uint32_t)));  // stop, next, snext
    uint32_t* d_split = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * SGM_SPLIT_WORDS));
    unsigned long long* d_rec =
        reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS));
    uint32_t* d_lastbits = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * lb_words));
    if (at >= kScratchBytes)
        return set_err(c, SGM_EINVAL, "a graph of %u vertices needs %zu bytes of fixed scratch; the array holds %zu", V,
                       at, kScratchBytes);
    uint32_t* d_stack = reinterpret_cast<uint32_t*>(base + at);
    const size_t stack_bytes = kScratchBytes - at;

    // the set-op stack: as many waves (4 per block) as fit with the default per-wave depth
    // (a few hub lists); a stack overflow re-runs with 4x the depth on a quarter of the waves
    uint64_t words = o.stack_words ? o.stack_words : std::min<uint64_t>((uint64_t)8 * g->max_deg + 1024, 1u << 20);
    if ((uint64_t)4 * words * sizeof(uint32_t) > stack_bytes)
        return set_err(c, SGM_EINVAL, "a set-op stack of %llu words per wave needs %llu bytes per block; the array "
                                      "holds %zu after the fixed part", (unsigned long long)words,
                       (unsigned long long)(4 * words * sizeof(uint32_t)), stack_bytes);
    uint32_t blocks = default_blocks(c, o, 1);
    const uint32_t* hub_slot = (o.flags & SGM_ENUM_F_NO_BITMAPS) || !g->n_hubs ? nullptr : g->d_hub_slot;
    const uint32_t u0 = order[0];
    const uint32_t roots = q->counts[u0];
    const bool split = !(o.flags & SGM_ENUM_F_NO_SPLIT) &&
                       ((o.flags & SGM_ENUM_F_SPLIT_ALL) || roots <= SGM_SPLIT_MAX_ROOTS || narrow_roots(o));
    sgm_equery eq;
    std::memset(&eq, 0, sizeof eq);
    eq.plan = d_plan;
    eq.vmask = q->d_vmask;
    eq.lastbits = d_lastbits;
    eq.stop = d_words;
    eq.next = d_words + 1;
    eq.snext = d_words + 2;
    eq.split = split ? d_split : nullptr;
    eq.rec = split ? d_rec : nullptr;
    eq.st = d_st;
    eq.cap = (unsigned long long)o.cap;
    eq.budget = budget_ticks(c, o);
    root_range(o, &eq.root_lo, &eq.root_hi);
    eq.n_items = g->ge_cnt[std::min<uint32_t>(q->q.off[u0 + 1] - q->q.off[u0], 32)];
    eq.root_u = u0;
    const int use_fs = o.method == SGM_ENUM_CACHE_FILTER_FS ? 1 : 0;
    const int serial = (o.flags & SGM_ENUM_F_SERIAL_LAST) ? 1 : 0;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(d_plan, &plan, sizeof plan, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_eq, &eq, sizeof eq, hipMemcpyHostToDevice, c->stream));
    sgm_stats_dev got;
    for (uint32_t attempt = 0;; attempt++) {
        const uint64_t fit = stack_bytes / ((uint64_t)4 * words * sizeof(uint32_t));
        blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks, fit));
        // the array keeps the last call's contents: clear what the search reads as zero
        sgm::launch_fill_u32(c->stream, reinterpret_cast<uint32_t*>(d_st), sizeof(sgm_stats_dev) / sizeof(uint32_t), 0);
        sgm::launch_fill_u32(c->stream, d_words, 3, 0);
        sgm::launch_fill_u32(c->stream, c->d_stop, 1, 0);
        clear_host_stop(c);
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        sgm::launch_last_bits(c->stream, d_eq, 1, V);
        if (split) sgm::launch_split_roots(c->stream, g->d_off, g->d_bydeg, d_eq, 1, V, o.split_chunks);
        sgm::launch_enum_cache(c->stream, blocks, g->d_off, g->d_adj, g->d_bydeg, d_eq, 1, plan.n_ops, d_stack,
                               (uint32_t)words, c->d_stop, c->d_hstop, use_fs, hub_slot, g->d_hub_bits, g->hub_words,
                               serial, split);
        if (split) sgm::launch_split_reduce(c->stream, d_eq, 1);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipMemcpyAsync(&got, d_st, sizeof got, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        out->kernel_ms += ms;
        if (!got.overflow) break;
        // a wave's set-op stack overflowed (the result is void): deeper stacks, fewer waves
        out->retries++;
        if (attempt >= 3 || (uint64_t)4 * words * 4 * sizeof(uint32_t) > stack_bytes)
            return set_err(c, SGM_EDEVICE, "set-op stack overflow within the %zu-byte scratch array", kScratchBytes);
        words *= 4;
    }
    const double kms = out->kernel_ms;
    const uint32_t retries = out->retries;
    stats_out(got, o.cap, out, c->wall_ns_per_tick);
    out->kernel_ms = kms;
    out->retries = retries;
    if (c->stop_requested && !out->capped) {
        out->stopped = 1;
        return set_err(c, SGM_ESTOPPED, "stopped by sgm_stop");
    }
    return SGM_OK;
}
