// Wave64 primitives shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgm {

// ------------------------------------------------------------------ global pointers
// A pointer read from device memory (a descriptor or plan field) is accessed with flat
// instructions unless its type names the global address space; flat accesses count against
// both the vector-memory and the LDS/scalar wait counters, so every LDS or scalar-load wait
// then also waits for them.  gcast() types such a pointer as global.
template <class T>
using glb = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ const glb<T>* gcast(const T* p)
{
    return (const glb<T>*)p;
}
template <class T>
__device__ __forceinline__ glb<T>* gcast(T* p)
{
    return (glb<T>*)p;
}

// ------------------------------------------------------------------ wave primitives
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t rdlane(uint32_t x, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)l);
}

// a 64-bit value (a pointer) held by lane l, as a wave-uniform value
__device__ __forceinline__ uint64_t rdlane64(uint64_t x, uint32_t l)
{
    return ((uint64_t)rdlane((uint32_t)(x >> 32), l) << 32) | rdlane((uint32_t)x, l);
}

__device__ __forceinline__ uint32_t uni(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

__device__ __forceinline__ uint32_t wave_or32(uint32_t x)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) x |= (uint32_t)__shfl_xor((int)x, s);
    return x;
}

// inclusive prefix sum over the 64 lanes (every lane must be active)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, o);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// Segmented walks: lanes hold the inclusive prefix sums `incl` of per-lane segment lengths;
// returns the lane owning element t of the concatenation (the number of lanes whose
// segment ends at or before t; ends are sorted).  Every lane must be active.
__device__ __forceinline__ uint32_t wave_seg_owner(uint32_t incl, uint32_t t)
{
    uint32_t j = 0;
#pragma unroll
    for (uint32_t s = 32; s > 0; s >>= 1) {
        const uint32_t e = (uint32_t)__shfl((int)incl, (int)(j + s - 1));
        if (e <= t) j += s;
    }
    return j;
}

// OR of x over the lanes in `lanes` (a ballot), one readlane pair per set bit: cheaper than a
// butterfly when few lanes contribute, and it keeps no shuffle indices live.
__device__ __forceinline__ uint64_t lanes_or64(uint64_t lanes, uint64_t x)
{
    uint64_t r = 0;
    while (lanes) {
        const uint32_t l = (uint32_t)__ffsll((long long)lanes) - 1u;
        lanes &= lanes - 1;
        r |= ((uint64_t)rdlane((uint32_t)(x >> 32), l) << 32) | rdlane((uint32_t)x, l);
    }
    return r;
}

// later reads of LDS words other lanes wrote see the writes
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t x)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) x += (uint32_t)__shfl_xor((int)x, s);
    return x;
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t x)
{
    uint32_t lo = wave_or32((uint32_t)x), hi = wave_or32((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// One 64-entry step of a segmented walk over the concatenated lists of the wave's lanes
// (lane l's list of len_l entries; incl = the inclusive prefix sums of the lengths, total =
// their sum): the entries of one list are a contiguous lane range of a step.
struct TabStep {
    uint32_t own;    // lane whose row holds this entry (64 when none)
    uint32_t pos;    // entry index within that row
    bool act;        // the lane holds an entry
    bool last;       // the last entry of its row in this step
    uint64_t below;  // lanes of the same row before this one in this step
};

__device__ __forceinline__ TabStep tab_step(uint32_t incl, uint32_t len, uint32_t total, uint32_t b0, uint32_t lane)
{
    TabStep r;
    const uint32_t t = b0 + lane;
    r.own = wave_seg_owner(incl, t);
    const uint32_t sj = (uint32_t)__shfl((int)(incl - len), (int)(r.own & 63u));
    const uint32_t ej = (uint32_t)__shfl((int)incl, (int)(r.own & 63u));
    r.act = t < total;
    r.pos = t - sj;
    r.last = r.act && (t + 1 == ej || lane == 63);
    const uint32_t segstart = r.act ? (sj > b0 ? sj - b0 : 0u) : lane;
    r.below = ((1ull << lane) - 1ull) & ~((1ull << segstart) - 1ull);
    return r;
}

__device__ __forceinline__ bool lower_found(const uint32_t* __restrict__ a, uint32_t n, uint32_t key)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == key;
}

}  // namespace sgm
