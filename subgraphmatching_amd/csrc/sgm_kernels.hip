// gfx950 kernels for the filter -> order -> enumerate path (wave64, CDNA4).
//
// All matching arithmetic is integer.  Every kernel keeps control flow wave-uniform
// where lanes cooperate (ballots, readlane), so per-wave state lives in LDS or in
// uniform registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgm_internal.h"
#include "sgm_kernels.h"

namespace sgm {

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

__device__ __forceinline__ uint64_t wave_or64(uint64_t x)
{
    uint32_t lo = wave_or32((uint32_t)x), hi = wave_or32((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
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

// ================================================================== NLF
// Unlabeled NLF (inc/filter.h:24-58): v may take u iff deg(v) >= deg_q(u).  One thread
// per data vertex writes its membership word for every query vertex at once.
__global__ void __launch_bounds__(256) k_nlf(const uint32_t* __restrict__ off, uint32_t V,
                                             const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    uint32_t* __restrict__ vm = d->vmask;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint32_t dv = off[v + 1] - off[v], m = 0;
        for (uint32_t u = 0; u < n; u++) m |= (dv >= d->qdeg[u] ? 1u : 0u) << u;
        vm[v] = m;
    }
}

// ================================================================== GQL
// Local semi-perfect matching (inc/filter.h:168-204): left = N_q(u) as a bitmask L of query
// vertices, right = N_G(v); right vertex w is adjacent to left vertex u' iff bit u' of
// vmask[w].  Exact maximum-matching test, wave-uniform:
//   * right vertices stream in adjacency order, 64 per chunk (one coalesced load of N(v),
//     one load of vmask per neighbour);
//   * each new right vertex is matched greedily to a free left vertex, else by a BFS for
//     an augmenting path over left vertices (lane i keeps the mask of the right vertex
//     matched to left vertex i);
//   * "closed" accumulates the left vertices reached by failed searches: they are all
//     matched and every path from them stays inside, so right vertices whose mask lies in
//     closed are skipped (Kuhn's algorithm never re-tries a failed vertex);
//   * the test stops as soon as |L| vertices are matched or too few right vertices remain.
// Only the boolean result matters, and any exact maximum matching gives the same one.
__device__ bool lsm_wave(const uint32_t* __restrict__ nb, uint32_t deg, const uint32_t* __restrict__ vm,
                         uint32_t L, uint32_t k)
{
    const uint32_t lane = lane_id();
    uint32_t freeL = L, closed = 0, matched = 0;
    uint32_t Ml = 0;
    for (uint32_t base = 0; base < deg; base += 64) {
        const uint32_t j = base + lane;
        const uint32_t m = (j < deg) ? (vm[nb[j]] & L) : 0u;
        uint64_t pend = __ballot((m & ~closed) != 0u);
        while (pend) {
            const uint32_t t = (uint32_t)__ffsll((long long)pend) - 1u;
            pend &= pend - 1;
            const uint32_t mf = rdlane(m, t);
            const uint32_t mm = mf & ~closed;
            if (!mm) continue;
            const uint32_t fr = mm & freeL;
            if (fr) {
                const uint32_t i = (uint32_t)__ffs(fr) - 1u;
                if (lane == i) Ml = mf;
                freeL &= ~(1u << i);
                if (++matched == k) return true;
                continue;
            }
            uint32_t visited = mm, frontier = mm;
            uint32_t via = ((mm >> lane) & 1u) ? 64u : 255u;  // 64: reached from the new vertex
            bool found = false;
            uint32_t tgt = 0;
            while (true) {
                const uint32_t x = ((frontier >> lane) & 1u) ? (Ml & ~visited & ~closed) : 0u;
                const uint32_t nxt = wave_or32(x);
                if (!nxt) break;
                uint32_t fr2 = frontier;
                while (fr2) {
                    const uint32_t i = (uint32_t)__ffs(fr2) - 1u;
                    fr2 &= fr2 - 1u;
                    const uint32_t xi = rdlane(x, i);
                    if (((xi >> lane) & 1u) && via == 255u) via = i;
                }
                visited |= nxt;
                const uint32_t hit = nxt & freeL;
                if (hit) { tgt = (uint32_t)__ffs(hit) - 1u; found = true; break; }
                frontier = nxt;
            }
            if (found) {
                uint32_t cur = tgt;
                while (true) {
                    const uint32_t p = rdlane(via, cur);
                    const uint32_t val = (p == 64u) ? mf : rdlane(Ml, p);
                    if (lane == cur) Ml = val;
                    if (p == 64u) break;
                    cur = p;
                }
                freeL &= ~(1u << tgt);
                if (++matched == k) return true;
            } else {
                closed |= visited;
            }
        }
        const uint32_t left = deg > base + 64 ? deg - (base + 64) : 0u;
        if (left < k - matched) return false;
    }
    return false;
}

// One refinement phase (round r, query vertex u) of the reference's GQL schedule
// (inc/filter.h:273-290) for every query of a batch (blockIdx.y = query).  Candidates of u
// are walked hub-first through the degree-sorted vertex list: its prefix of length
// ge_cnt[deg_q(u)] is exactly the NLF set, and vmask bit u says whether v is still valid.
// Only bit u changes during the phase and u is not its own neighbour, so the checks of
// one phase are independent (the reference's sequential order inside a phase does not
// matter).  A round-2 phase whose left-side sets did not change since round 1 would
// repeat round 1's answers exactly and is skipped.
//
// A wave takes 64 candidates at a time: lanes fetch their ids, validity bits and
// offsets together.  A candidate of degree <= GQL_LANE_DEG is first decided by its own lane
// from its right-side masks: a greedy matching that saturates the left side proves the
// answer true; a left vertex no right vertex covers, or fewer useful right vertices than
// left vertices, proves it false.  Only the remaining candidates (and all higher-degree
// ones) go through the exact wave matcher, one after another.
#define GQL_LANE_DEG 32u

__device__ __forceinline__ int lsm_lane(const uint32_t* __restrict__ adj, uint32_t o, uint32_t deg,
                                        const uint32_t* __restrict__ vm, uint32_t L, uint32_t k)
{
    // returns 1 = passes, 0 = fails, -1 = undecided
    uint32_t orm = 0, useful = 0, freeL = L, matched = 0;
    for (uint32_t j0 = 0; j0 < deg; j0 += 8) {
        uint32_t w[8], m[8];
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = (j0 + t < deg) ? adj[o + j0 + t] : 0u;
#pragma unroll
        for (int t = 0; t < 8; t++) m[t] = (j0 + t < deg) ? (vm[w[t]] & L) : 0u;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            orm |= m[t];
            useful += m[t] != 0u;
            const uint32_t f = m[t] & freeL;
            if (f) { freeL &= ~(f & (0u - f)); matched++; }
        }
    }
    if (matched == k) return 1;
    if (orm != L || useful < k) return 0;
    return -1;
}

__global__ void __launch_bounds__(256) k_gql_phase(const uint32_t* __restrict__ off,
                                                   const uint32_t* __restrict__ adj,
                                                   const uint32_t* __restrict__ by_deg,
                                                   const uint32_t* __restrict__ ge_cnt,
                                                   const sgm_qdesc* __restrict__ qs, uint32_t phase)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (phase >= 2 * n) return;
    const uint32_t r = phase / n, u = phase % n;
    const uint32_t L = d->nbr[u];
    if (L == 0) return;  // empty left side: the matching trivially saturates it
    uint32_t* __restrict__ removed = d->removed;
    if (r == 1) {
        uint32_t touched = 0;
        for (uint32_t p = u + 1; p < n; p++) if (removed[p]) touched |= 1u << p;
        for (uint32_t p = 0; p < u; p++) if (removed[n + p]) touched |= 1u << p;
        if (!(touched & L)) return;
    }
    const uint32_t k = __popc(L);
    const uint32_t cnt = ge_cnt[d->qdeg[u]];
    uint32_t* __restrict__ vm = d->vmask;
    const uint32_t lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t w0 = blockIdx.x * waves + (threadIdx.x >> 6);
    const uint32_t stride = gridDim.x * waves;
    uint32_t local_removed = 0;
    for (uint32_t base = w0 * 64; base < cnt; base += stride * 64) {
        const uint32_t i = base + lane;
        const uint32_t v = i < cnt ? by_deg[i] : 0u;
        const bool valid = i < cnt && ((vm[v] >> u) & 1u);
        uint32_t o = 0, dv = 0;
        if (valid) { o = off[v]; dv = off[v + 1] - o; }
        int verdict = -1;
        if (valid && dv <= GQL_LANE_DEG) verdict = lsm_lane(adj, o, dv, vm, L, k);
        if (valid && verdict == 0) atomicAnd(&vm[v], ~(1u << u));
        local_removed += (uint32_t)__popcll(__ballot(valid && verdict == 0));
        uint64_t pend = __ballot(valid && verdict < 0);
        while (pend) {
            const uint32_t t = (uint32_t)__ffsll((long long)pend) - 1u;
            pend &= pend - 1;
            const uint32_t vt = rdlane(v, t), ot = rdlane(o, t), dt = rdlane(dv, t);
            if (!lsm_wave(adj + ot, dt, vm, L, k)) {
                if (lane == 0) atomicAnd(&vm[vt], ~(1u << u));
                local_removed++;
            }
        }
    }
    if (lane == 0 && local_removed) atomicAdd(&removed[phase], local_removed);
}

// ================================================================== hub adjacency bitmaps
// One V-bit bitmap per high-degree vertex (slot order = hubs[]), so intersections against
// a hub's adjacency list test membership with one load instead of a binary search.
// The bitmaps start zeroed; lanes of a block set the bits of one hub's neighbours.
__global__ void __launch_bounds__(256) k_hub_bits(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                  const uint32_t* __restrict__ hubs, uint32_t n_hubs,
                                                  uint32_t* __restrict__ bits, uint32_t words)
{
    for (uint32_t h = blockIdx.x; h < n_hubs; h += gridDim.x) {
        const uint32_t v = hubs[h];
        uint32_t* __restrict__ b = bits + (size_t)h * words;
        for (uint32_t p = off[v] + threadIdx.x; p < off[v + 1]; p += blockDim.x) {
            const uint32_t w = adj[p];
            atomicOr(&b[w >> 5], 1u << (w & 31u));
        }
    }
}

// ================================================================== candidate bitmaps
// Per query vertex u: bm[u][w] = 64-vertex bitmap of C(u) (from vmask bit u), then
// rank[u][w] = exclusive popcount prefix and cnt[u] = |C(u)|.
__global__ void __launch_bounds__(256) k_bitmaps(const sgm_cdesc* __restrict__ cs, uint32_t V)
{
    const sgm_cdesc* c = cs + blockIdx.y;
    const uint32_t n = c->n, W = (V + 63) / 64;
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t lane = lane_id();
    for (uint32_t w = blockIdx.x * waves + (threadIdx.x >> 6); w < W; w += gridDim.x * waves) {
        const uint32_t v = w * 64 + lane;
        const uint32_t m = v < V ? c->vmask[v] : 0u;
        for (uint32_t u = 0; u < n; u++) {
            const uint64_t b = __ballot((m >> u) & 1u);
            if (lane == 0) c->bm[(size_t)u * W + w] = b;
        }
    }
}

// One block per (query, u): exclusive scan of popcounts over the W words.
__global__ void __launch_bounds__(256) k_rank_scan(const sgm_cdesc* __restrict__ cs, uint32_t V)
{
    const sgm_cdesc* c = cs + blockIdx.y;
    const uint32_t u = blockIdx.x;
    if (u >= c->n) return;
    const uint32_t W = (V + 63) / 64;
    const uint64_t* __restrict__ bm = c->bm + (size_t)u * W;
    uint32_t* __restrict__ rank = c->rank + (size_t)u * W;
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < W; base += 256) {
        const uint32_t w = base + threadIdx.x;
        const uint32_t x = w < W ? (uint32_t)__popcll(bm[w]) : 0u;
        // inclusive wave scan
        uint32_t s = x;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)s, o);
            if (lane >= (uint32_t)o) s += y;
        }
        if (lane == 63) wsum[wv] = s;
        __syncthreads();
        uint32_t pre = carry;
        for (uint32_t i = 0; i < wv; i++) pre += wsum[i];
        if (w < W) rank[w] = pre + s - x;
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) c->cnt[u] = carry;
}

// Compact C(u) in ascending vertex order into list[cand_off[u] ...].
__global__ void __launch_bounds__(256) k_compact(const sgm_cdesc* __restrict__ cs, uint32_t V)
{
    const sgm_cdesc* c = cs + blockIdx.y;
    const uint32_t n = c->n, W = (V + 63) / 64;
    const uint32_t waves = blockDim.x >> 6, lane = lane_id();
    for (uint32_t w = blockIdx.x * waves + (threadIdx.x >> 6); w < W; w += gridDim.x * waves) {
        for (uint32_t u = 0; u < n; u++) {
            const uint64_t b = c->bm[(size_t)u * W + w];
            if ((b >> lane) & 1ull)
                c->list[c->cand_off[u] + c->rank[(size_t)u * W + w] + mbcnt64(b)] = w * 64 + lane;
        }
    }
}

// ================================================================== aux tables
// Table of directed query edge (a,b) (inc/table.h:7-137): for the i-th candidate of a, the
// ascending positions k in C(b) of its data neighbours that are candidates of b.
// Pass 1 counts per row, pass 2 fills (positions are ranks in b's bitmap, so rows come out
// sorted because adjacency lists are).  One wave per row, coalesced over N(v).
__device__ __forceinline__ uint32_t cand_rank(const uint64_t* __restrict__ bm, const uint32_t* __restrict__ rank,
                                              uint32_t w)
{
    const uint32_t word = w >> 6, bit = w & 63u;
    const uint64_t low = bit ? (bm[word] & ((1ull << bit) - 1ull)) : 0ull;
    return rank[word] + (uint32_t)__popcll(low);
}

__global__ void __launch_bounds__(256) k_table_count(const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ adj,
                                                     const uint32_t* __restrict__ vmask,
                                                     const uint32_t* __restrict__ rows, uint32_t nrows,
                                                     uint32_t b, uint32_t* __restrict__ rowcnt)
{
    const uint32_t waves = blockDim.x >> 6, lane = lane_id();
    for (uint32_t i = blockIdx.x * waves + (threadIdx.x >> 6); i < nrows; i += gridDim.x * waves) {
        const uint32_t v = rows[i], o = off[v], dv = off[v + 1] - o;
        uint32_t c = 0;
        for (uint32_t j = lane; j < dv; j += 64) c += (vmask[adj[o + j]] >> b) & 1u;
        for (int s = 32; s > 0; s >>= 1) c += (uint32_t)__shfl_xor((int)c, s);
        if (lane == 0) rowcnt[i] = c;
    }
}

__global__ void __launch_bounds__(256) k_table_fill(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ adj,
                                                    const uint32_t* __restrict__ vmask,
                                                    const uint32_t* __restrict__ rows, uint32_t nrows,
                                                    uint32_t b, const uint64_t* __restrict__ bm_b,
                                                    const uint32_t* __restrict__ rank_b,
                                                    const uint32_t* __restrict__ rowoff,
                                                    uint32_t* __restrict__ eidx, uint32_t* __restrict__ ev)
{
    const uint32_t waves = blockDim.x >> 6, lane = lane_id();
    for (uint32_t i = blockIdx.x * waves + (threadIdx.x >> 6); i < nrows; i += gridDim.x * waves) {
        const uint32_t v = rows[i], o = off[v], dv = off[v + 1] - o;
        uint32_t outp = rowoff[i];
        for (uint32_t base = 0; base < dv; base += 64) {
            const uint32_t j = base + lane;
            const uint32_t w = j < dv ? adj[o + j] : 0u;
            const bool hit = j < dv && ((vmask[w] >> b) & 1u);
            const uint64_t bal = __ballot(hit);
            if (hit) {
                const uint32_t p = outp + mbcnt64(bal);
                eidx[p] = cand_rank(bm_b, rank_b, w);
                ev[p] = w;
            }
            outp += (uint32_t)__popcll(bal);
        }
    }
}

// Exclusive scan of n u32 counts into n+1 u64-safe u32 offsets (single block, looped).
__global__ void __launch_bounds__(256) k_scan_u32(const uint32_t* __restrict__ in, uint32_t n,
                                                  uint32_t* __restrict__ out, uint32_t* __restrict__ maxv)
{
    __shared__ uint32_t wsum[4], wmax[4];
    __shared__ uint32_t carry, cmax;
    if (threadIdx.x == 0) { carry = 0; cmax = 0; }
    __syncthreads();
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < n; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < n ? in[i] : 0u;
        uint32_t s = x, mx = x;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)s, o);
            if (lane >= (uint32_t)o) s += y;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) { uint32_t y = (uint32_t)__shfl_xor((int)mx, o); mx = y > mx ? y : mx; }
        if (lane == 63) wsum[wv] = s;
        if (lane == 0) wmax[wv] = mx;
        __syncthreads();
        uint32_t pre = carry;
        for (uint32_t k = 0; k < wv; k++) pre += wsum[k];
        if (i < n) out[i] = pre + s - x;
        __syncthreads();
        if (threadIdx.x == 0) {
            carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
            for (int k = 0; k < 4; k++) cmax = wmax[k] > cmax ? wmax[k] : cmax;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[n] = carry; *maxv = cmax; }
}

// ================================================================== enumeration
// Cache + filter + DAF failing-set search (inc/explore.h:968-1129, inc/prune.h:82-171),
// one wave per depth-0 candidate.  The reference never cuts at depth 0 and resets every
// failing-set slot before reading it, so the subtrees of different depth-0 candidates are
// independent: each wave reproduces the reference's search of its subtree exactly
// (same set-op evaluations, same failing-set cut-offs), and per-depth node and
// intersection counts add up to the reference's.
//
// Per wave: the partial embedding, per-depth iteration cursors and failing masks, and the
// location/size of every cached set op live in LDS; materialised ops are kept on a per-wave
// stack in HBM (ops computed at depth d start at stack_top[d] and are overwritten when the
// vertex at depth d changes).  Each level is scanned 64 candidates at a time: lanes test
// the candidate filter (vmask bit) and injectivity, ballots give the ordered list of
// extensions, and the conflicts of a chunk are OR-ed into the parent's failing mask at once
// (a later cut-off overwrites that mask anyway, so this equals the one-by-one order).
// The last level is never materialised: its set op is intersected and counted in one pass.
struct WaveDFS {
    uint64_t rem[SGM_MAXQ];
    uint64_t fail[SGM_MAXQ];
    unsigned long long nodes[SGM_MAXQ];
    uint32_t emb[SGM_MAXQ];
    uint32_t pos[SGM_MAXQ];
    uint32_t tot[SGM_MAXQ];
    uint32_t lst[SGM_MAXQ];
    uint32_t chunk_base[SGM_MAXQ];
    uint32_t flags[SGM_MAXQ];
    uint32_t stack_top[SGM_MAXQ + 1];
};

#define F_ANY 1u
#define F_PRUNED 2u
#define STACK_FLAG 0x80000000u

__device__ __forceinline__ const uint32_t* op_list(uint32_t o_off, const uint32_t* adj, const uint32_t* ws)
{
    return (o_off & STACK_FLAG) ? ws + (o_off & ~STACK_FLAG) : adj + o_off;
}

// Membership of key in the sorted list b: one bit test when b is the adjacency list of a
// vertex that has a bitmap (bits != nullptr), else a binary search.
__device__ __forceinline__ bool member(const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bits,
                                       uint32_t key)
{
    return bits ? ((bits[key >> 5] >> (key & 31u)) & 1u) != 0u : lower_found(b, nb, key);
}

// Each lane handles UNR elements per step (256 per wave), so a wave keeps UNR independent
// memory chains in flight instead of one: intersections and scans are latency-bound, and
// the time to reach the first embeddings is a chain of them.
#define UNR 4

// Membership of UNR keys in the sorted list b (nb > 0) at once: a bit test each when b has
// a bitmap, else UNR branch-free binary searches in lockstep (the step count depends only
// on nb, so the loads of the UNR searches overlap).
__device__ __forceinline__ void member4(const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bits,
                                        const uint32_t (&key)[UNR], const bool (&valid)[UNR], bool (&hit)[UNR])
{
    if (bits) {
#pragma unroll
        for (int u = 0; u < UNR; u++) hit[u] = valid[u] && ((bits[key[u] >> 5] >> (key[u] & 31u)) & 1u);
        return;
    }
    uint32_t p[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) p[u] = 0;
    uint32_t n = nb;
    while (n > 1) {
        const uint32_t half = n >> 1;
#pragma unroll
        for (int u = 0; u < UNR; u++) p[u] = (b[p[u] + half] < key[u]) ? p[u] + half : p[u];
        n -= half;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        const uint32_t lb = p[u] + (b[p[u]] < key[u] ? 1u : 0u);
        hit[u] = valid[u] && lb < nb && b[lb] == key[u];
    }
}

// Intersect sorted lists a and b into out (ascending); returns the size.  Lanes take
// elements of the shorter list and test them against the longer one (bitmap or binary
// search).  abits/bbits: bitmaps of a / b when they are hub adjacency lists, else null.
// Stops early (result unused) once the stop word is raised.
__device__ uint32_t isect_write(const uint32_t* __restrict__ a, uint32_t na, const uint32_t* __restrict__ abits,
                                const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bbits,
                                uint32_t* __restrict__ out, unsigned long long& elems, unsigned long long& outs,
                                const uint32_t* stop)
{
    if (na > nb) {
        const uint32_t* t = a; a = b; b = t;
        uint32_t tn = na; na = nb; nb = tn;
        bbits = abits;
    }
    const uint32_t lane = lane_id();
    uint32_t cnt = 0;
    elems += na + nb;
    if (nb == 0) return 0;
    for (uint32_t base = 0; base < na; base += 64 * UNR) {
        uint32_t key[UNR];
        bool valid[UNR], hit[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t i = base + 64 * u + lane;
            valid[u] = i < na;
            key[u] = valid[u] ? a[i] : 0u;
        }
        member4(b, nb, bbits, key, valid, hit);
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint64_t bal = __ballot(hit[u]);
            if (hit[u]) out[cnt + mbcnt64(bal)] = key[u];
            cnt += (uint32_t)__popcll(bal);
        }
        if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    }
    outs += cnt;
    return cnt;
}

// A wave publishes its embedding count at least every FLUSH_AT embeddings (and at the end
// of each depth-0 item), so the cap can stop the whole grid while waves are still deep
// inside huge subtrees; the stop word is raised once the global count reaches the cap,
// so the final total is >= cap whenever the cap stopped the search.
#define FLUSH_AT 512ull

__device__ __forceinline__ void flush_count(unsigned long long& pending, unsigned long long* gcount, uint32_t* stop,
                                            unsigned long long cap)
{
    if (!pending) return;
    unsigned long long prev = 0;
    if (lane_id() == 0) prev = atomicAdd(gcount, pending);
    prev = __shfl(prev, 0);
    if (prev + pending >= cap && lane_id() == 0) atomicExch(stop, 1u);
    pending = 0;
}

struct LastRes {
    uint32_t size;      // iteration-set size
    uint32_t ok;        // extensions (embeddings)
    uint64_t conf;      // OR of conflict failing sets
    uint32_t any;       // some element passed the filter
};

// Classify UNR elements per lane at the last depth L: filter bit of u_L, then injectivity
// against the L mapped vertices.
__device__ __forceinline__ void classify_last4(const bool (&valid)[UNR], const uint32_t (&w)[UNR],
                                               const uint32_t* __restrict__ vmask, uint32_t uL, const WaveDFS& S,
                                               const uint64_t* anc, uint32_t L, LastRes& r)
{
    bool f[UNR];
    uint32_t c[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        f[u] = valid[u] && ((vmask[w[u]] >> uL) & 1u);
        c[u] = SGM_INVALID;
    }
    for (uint32_t j = 0; j < L; j++) {
        const uint32_t e = S.emb[j];
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (e == w[u]) c[u] = j;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        const bool ok = f[u] && c[u] == SGM_INVALID, cf = f[u] && c[u] != SGM_INVALID;
        const uint64_t okb = __ballot(ok);
        const uint64_t cfb = __ballot(cf);
        r.ok += (uint32_t)__popcll(okb);
        if (okb | cfb) r.any = 1;
        if (cfb) r.conf |= wave_or64(cf ? (anc[L] | anc[c[u]]) : 0ull);
    }
}

// Last level, scanning an existing list.
__device__ void last_scan(const uint32_t* __restrict__ list, uint32_t size, const uint32_t* __restrict__ vmask,
                          uint32_t uL, const WaveDFS& S, const uint64_t* anc, uint32_t L, LastRes& r,
                          const uint32_t* stop)
{
    const uint32_t lane = lane_id();
    r.size = size;
    for (uint32_t base = 0; base < size; base += 64 * UNR) {
        uint32_t w[UNR];
        bool valid[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t i = base + 64 * u + lane;
            valid[u] = i < size;
            w[u] = valid[u] ? list[i] : 0u;
        }
        classify_last4(valid, w, vmask, uL, S, anc, L, r);
        if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    }
}

// Last level, intersecting (parent op) with N(v) and classifying in the same pass.
__device__ void last_isect(const uint32_t* __restrict__ a, uint32_t na, const uint32_t* __restrict__ abits,
                           const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bbits,
                           const uint32_t* __restrict__ vmask, uint32_t uL, const WaveDFS& S,
                           const uint64_t* anc, uint32_t L, LastRes& r, unsigned long long& elems,
                           const uint32_t* stop)
{
    if (na > nb) {
        const uint32_t* t = a; a = b; b = t;
        uint32_t tn = na; na = nb; nb = tn;
        bbits = abits;
    }
    const uint32_t lane = lane_id();
    elems += na + nb;
    uint32_t size = 0;
    if (nb == 0) { r.size = 0; return; }
    for (uint32_t base = 0; base < na; base += 64 * UNR) {
        uint32_t key[UNR];
        bool valid[UNR], hit[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t i = base + 64 * u + lane;
            valid[u] = i < na;
            key[u] = valid[u] ? a[i] : 0u;
        }
        member4(b, nb, bbits, key, valid, hit);
#pragma unroll
        for (int u = 0; u < UNR; u++) size += (uint32_t)__popcll(__ballot(hit[u]));
        classify_last4(hit, key, vmask, uL, S, anc, L, r);
        if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    }
    r.size = size;
}

__global__ void __launch_bounds__(256) k_enum_cache(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ adj,
                                                    const uint32_t* __restrict__ vmask,
                                                    const uint32_t* __restrict__ items, uint32_t n_items,
                                                    uint32_t root_u,
                                                    const sgm_plan_dev* __restrict__ gplan,
                                                    uint32_t* __restrict__ stack, uint32_t stack_words,
                                                    uint32_t* __restrict__ work, sgm_stats_dev* __restrict__ gst,
                                                    uint32_t* __restrict__ stop, unsigned long long cap,
                                                    int use_fs, const uint32_t* __restrict__ hub_slot,
                                                    const uint32_t* __restrict__ hub_bits, uint32_t hub_words)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* smem = reinterpret_cast<uint8_t*>(smem_u64);
    sgm_plan_dev& P = *reinterpret_cast<sgm_plan_dev*>(smem);
    // bitmap of a vertex's adjacency list, or null when it has none
    auto vbits = [&](uint32_t v) -> const uint32_t* {
        if (!hub_slot) return nullptr;
        const uint32_t s = hub_slot[v];
        return s == SGM_INVALID ? nullptr : hub_bits + (size_t)s * hub_words;
    };
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(gplan);
        uint32_t* dst = reinterpret_cast<uint32_t*>(smem);
        for (uint32_t i = threadIdx.x; i < sizeof(sgm_plan_dev) / 4; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const uint32_t n = P.n, n_ops = P.n_ops;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const size_t plan_bytes = (sizeof(sgm_plan_dev) + 15) & ~(size_t)15;
    const size_t wave_bytes = ((sizeof(WaveDFS) + 12u * n_ops) + 15) & ~(size_t)15;
    WaveDFS& S = *reinterpret_cast<WaveDFS*>(smem + plan_bytes + wave * wave_bytes);
    uint32_t* op_off = reinterpret_cast<uint32_t*>(smem + plan_bytes + wave * wave_bytes + sizeof(WaveDFS));
    uint32_t* op_size = op_off + n_ops;
    uint32_t* op_vtx = op_size + n_ops;  // vertex of an adjacency (alias) op, else SGM_INVALID
    uint32_t* ws = stack + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * stack_words;
    const uint32_t L = n - 1;
    const uint32_t uL = P.order[L];

    for (uint32_t d = 0; d < SGM_MAXQ; d++) S.nodes[d] = 0;
    unsigned long long inter = 0, fsp = 0, elems = 0, items_done = 0, outs = 0, scanned = 0;
    unsigned long long cnt_pending = 0;
    bool overflow = false;
    uint32_t poll = 0;

    while (true) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(work, 1u);
        it = uni(__shfl((int)it, 0));
        if (it >= n_items) break;
        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) break;
        const uint32_t v0 = items[it];
        if (!((vmask[v0] >> root_u) & 1u)) continue;  // not (or no longer) a candidate of u0
        items_done++;
        S.emb[0] = v0;
        S.nodes[0] += 1;
        S.stack_top[0] = 0;
        bool aborted = false;

        // ---- ops of depth 0 (or the whole last level when n == 2)
        if (n == 2) {
            LastRes r = {0, 0, 0, 0};
            const uint32_t o = P.iter_set[1];
            op_off[o] = off[v0];
            op_size[o] = off[v0 + 1] - off[v0];
            last_scan(adj + off[v0], op_size[o], vmask, uL, S, P.anc, L, r, stop);
            scanned += r.size;
            cnt_pending += r.ok;
            S.nodes[L] += r.ok;
        } else {
            bool ok = true;
            uint32_t top = 0;
            const uint32_t nbo = off[v0], nbs = off[v0 + 1] - nbo;
            for (uint32_t o = P.op_first[0]; o < P.op_first[1]; o++) {
                const uint32_t p = P.op_parent[o];
                if (p == SGM_NOPARENT) { op_off[o] = nbo; op_size[o] = nbs; op_vtx[o] = v0; }
                else {
                    const uint32_t need = op_size[p] < nbs ? op_size[p] : nbs;
                    if (top + need > stack_words) { overflow = true; aborted = true; ok = false; break; }
                    inter++;
                    const uint32_t sz = isect_write(op_list(op_off[p], adj, ws), op_size[p],
                                                    op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr,
                                                    adj + nbo, nbs, vbits(v0), ws + top, elems, outs, stop);
                    op_off[o] = top | STACK_FLAG;
                    op_size[o] = sz;
                    op_vtx[o] = SGM_INVALID;
                    top += sz;
                }
                if (op_size[o] == 0) { ok = false; break; }
            }
            if (ok) {
                S.stack_top[1] = top;
                // enter depth 1
                S.pos[1] = 0; S.rem[1] = 0; S.flags[1] = 0;
                S.tot[1] = op_size[P.iter_set[1]];
                S.lst[1] = op_off[P.iter_set[1]];
                S.fail[0] = 0;
                uint32_t d = 1;
                while (true) {
                    if ((++poll & 3u) == 0 && __atomic_load_n(stop, __ATOMIC_RELAXED)) { aborted = true; break; }
                    if (S.rem[d] == 0 && ((S.flags[d] & F_PRUNED) || S.pos[d] >= S.tot[d])) {
                        // end of the loop at depth d (inc/explore.h:1068-1070)
                        if (use_fs && !(S.flags[d] & (F_PRUNED | F_ANY))) S.fail[d - 1] = P.anc[d];
                        d--;
                        if (d == 0) break;
                        // PruneCheck(d) after the child returned (inc/prune.h:119-137)
                        if (use_fs) {
                            if (!((S.fail[d] >> d) & 1ull)) {
                                S.fail[d - 1] = S.fail[d];
                                S.flags[d] |= F_PRUNED;
                                S.rem[d] = 0;
                                fsp++;
                            } else S.fail[d - 1] |= S.fail[d];
                        }
                        continue;
                    }
                    const uint32_t* list = op_list(S.lst[d], adj, ws);
                    if (S.rem[d] == 0) {
                        // classify the next 64 elements of depth d's iteration set
                        const uint32_t u = P.order[d];
                        const uint32_t base = S.pos[d];
                        S.pos[d] = base + 64;
                        scanned += (S.tot[d] - base < 64u) ? (S.tot[d] - base) : 64u;
                        S.chunk_base[d] = base;
                        const uint32_t i = base + lane;
                        const bool valid = i < S.tot[d];
                        const uint32_t w = valid ? list[i] : 0u;
                        const bool f = valid && ((vmask[w] >> u) & 1u);
                        uint32_t c = SGM_INVALID;
                        if (f)
                            for (uint32_t j = 0; j < d; j++)
                                if (S.emb[j] == w) c = j;
                        const uint64_t okb = __ballot(f && c == SGM_INVALID);
                        const uint64_t cfb = __ballot(f && c != SGM_INVALID);
                        if (okb | cfb) S.flags[d] |= F_ANY;
                        if (use_fs && cfb) {
                            const uint64_t m = (f && c != SGM_INVALID) ? (P.anc[d] | P.anc[c]) : 0ull;
                            S.fail[d - 1] |= wave_or64(m);
                        }
                        S.rem[d] = okb;
                        continue;
                    }
                    // extend with the next candidate of this chunk
                    const uint64_t rm = S.rem[d];
                    const uint32_t j = (uint32_t)__ffsll((long long)rm) - 1u;
                    S.rem[d] = rm & (rm - 1);
                    const uint32_t v = list[S.chunk_base[d] + j];
                    S.emb[d] = v;
                    S.nodes[d] += 1;
                    const uint32_t nbo = off[v], nbs = off[v + 1] - nbo;
                    if (d == n - 2) {
                        // last level: child at depth L evaluated in one pass
                        LastRes r = {0, 0, 0, 0};
                        bool empty = false;
                        uint32_t conf_v = 0;
                        if (P.op_first[d] < P.op_first[d + 1]) {
                            const uint32_t o = P.op_first[d];
                            const uint32_t p = P.op_parent[o];
                            if (p == SGM_NOPARENT) {
                                last_scan(adj + nbo, nbs, vmask, uL, S, P.anc, L, r, stop);
                            } else {
                                inter++;
                                last_isect(op_list(op_off[p], adj, ws), op_size[p],
                                           op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr, adj + nbo, nbs,
                                           vbits(v), vmask, uL, S, P.anc, L, r, elems, stop);
                            }
                            if (r.size == 0) { empty = true; conf_v = P.op_vdepth[o]; }
                        } else {
                            const uint32_t o = P.iter_set[L];
                            last_scan(op_list(op_off[o], adj, ws), op_size[o], vmask, uL, S, P.anc, L, r, stop);
                        }
                        cnt_pending += r.ok;
                        if (cnt_pending >= FLUSH_AT) flush_count(cnt_pending, &gst->count, stop, cap);
                        S.nodes[L] += r.ok;
                        scanned += r.size;
                        if (P.op_first[d] < P.op_first[d + 1] && P.op_parent[P.op_first[d]] != SGM_NOPARENT)
                            outs += r.size;
                        if (use_fs) {
                            uint64_t f;
                            if (empty) f = P.anc[conf_v];
                            else {
                                f = r.conf;
                                if (r.ok) f = ~0ull;
                                if (!r.any) f = P.anc[L];
                            }
                            S.fail[d] = f;
                            if (!((f >> d) & 1ull)) {
                                S.fail[d - 1] = f;
                                S.flags[d] |= F_PRUNED;
                                S.rem[d] = 0;
                                fsp++;
                            } else S.fail[d - 1] |= f;
                        }
                        continue;
                    }
                    // compute the set ops of depth d (CacheBuffer::getNextCandidatesImpl)
                    bool ok2 = true;
                    uint32_t conf = 0;
                    uint32_t top = S.stack_top[d];
                    for (uint32_t o = P.op_first[d]; o < P.op_first[d + 1]; o++) {
                        const uint32_t p = P.op_parent[o];
                        if (p == SGM_NOPARENT) { op_off[o] = nbo; op_size[o] = nbs; op_vtx[o] = v; }
                        else {
                            const uint32_t need = op_size[p] < nbs ? op_size[p] : nbs;
                            if (top + need > stack_words) { overflow = true; aborted = true; ok2 = false; break; }
                            inter++;
                            const uint32_t sz = isect_write(op_list(op_off[p], adj, ws), op_size[p],
                                                            op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr,
                                                            adj + nbo, nbs, vbits(v), ws + top, elems, outs, stop);
                            op_off[o] = top | STACK_FLAG;
                            op_size[o] = sz;
                            op_vtx[o] = SGM_INVALID;
                            top += sz;
                        }
                        if (op_size[o] == 0) { ok2 = false; conf = P.op_vdepth[o]; break; }
                    }
                    if (aborted) break;
                    if (!ok2) {
                        if (use_fs) {
                            const uint64_t f = P.anc[conf];
                            S.fail[d] = f;
                            if (!((f >> d) & 1ull)) {
                                S.fail[d - 1] = f;
                                S.flags[d] |= F_PRUNED;
                                S.rem[d] = 0;
                                fsp++;
                            } else S.fail[d - 1] |= f;
                        }
                        continue;
                    }
                    S.stack_top[d + 1] = top;
                    const uint32_t nd = d + 1;
                    S.pos[nd] = 0; S.rem[nd] = 0; S.flags[nd] = 0;
                    S.tot[nd] = op_size[P.iter_set[nd]];
                    S.lst[nd] = op_off[P.iter_set[nd]];
                    S.fail[d] = 0;
                    d = nd;
                }
            }
        }
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (aborted || overflow) break;
    }
    flush_count(cnt_pending, &gst->count, stop, cap);
    if (lane == 0) {
        atomicAdd(&gst->intersections, inter);
        atomicAdd(&gst->fs_prunes, fsp);
        atomicAdd(&gst->elems_read, elems);
        atomicAdd(&gst->elems_out, outs);
        atomicAdd(&gst->elems_scanned, scanned);
        atomicAdd(&gst->items_done, items_done);
        if (overflow) { atomicAdd(&gst->overflow, 1ull); atomicExch(stop, 1u); }
        for (uint32_t d = 0; d < n; d++)
            if (S.nodes[d]) atomicAdd(&gst->nodes[d], S.nodes[d]);
    }
}


// ================================================================== aux-table enumeration
// Plain candidate-space search over the auxiliary tables (inc/explore.h:2610-2848 with every
// pruning option off; CandidatesBuffer::getNextCandidatesImpl :284-354): the candidates of
// depth d are the intersection, over the backward neighbours p of order[d] in order
// position, of the table rows edge_matrix[p][order[d]].getNeb(idx(p)) -- positions in
// C(order[d]), ascending; each extra backward neighbour is one intersection (counted even
// when an earlier one came out empty, as the reference does).  Injectivity is the only
// check.  One wave per depth-0 candidate; the last level is intersected and counted in one
// pass.  Per-wave intermediate lists live on the same HBM stack as the cache search.
struct WaveAux {
    uint64_t rem[SGM_MAXQ];
    unsigned long long nodes[SGM_MAXQ];
    const uint32_t* lst[SGM_MAXQ];
    uint32_t emb[SGM_MAXQ];
    uint32_t iemb[SGM_MAXQ];
    uint32_t pos[SGM_MAXQ];
    uint32_t tot[SGM_MAXQ];
    uint32_t chunk_base[SGM_MAXQ];
    uint32_t stack_top[SGM_MAXQ + 1];
};

// Build the candidate list of depth d into (*list, *size); all but the last intersection
// are materialised on the stack from *top; when `leave_last` is set the last backward
// neighbour's row is returned separately (for a fused count) in (*last, *nlast).
__device__ bool aux_list(const sgm_auxplan_dev& A, const WaveAux& S, uint32_t d, uint32_t* ws, uint32_t stack_words,
                         uint32_t& top, const uint32_t*& list, uint32_t& size, bool leave_last,
                         const uint32_t*& last, uint32_t& nlast, unsigned long long& elems, unsigned long long& outs,
                         const uint32_t* stop)
{
    const uint32_t f = A.prev_first[d], e_end = A.prev_first[d + 1];
    const uint32_t np = e_end - f;
    {
        const uint32_t* o = A.tab_off[f];
        const uint32_t r = S.iemb[A.prev_depth[f]];
        const uint32_t b = o[r];
        list = A.tab_idx[f] + b;
        size = o[r + 1] - b;
    }
    const uint32_t stop_at = leave_last ? e_end - 1 : e_end;
    for (uint32_t e = f + 1; e < stop_at; e++) {
        const uint32_t* o = A.tab_off[e];
        const uint32_t r = S.iemb[A.prev_depth[e]];
        const uint32_t b = o[r], nb = o[r + 1] - b;
        const uint32_t need = size < nb ? size : nb;
        if (top + need > stack_words) return false;
        const uint32_t sz = isect_write(list, size, nullptr, A.tab_idx[e] + b, nb, nullptr, ws + top, elems, outs, stop);
        list = ws + top;
        size = sz;
        top += sz;
    }
    if (leave_last && np > 1) {
        const uint32_t e = e_end - 1;
        const uint32_t* o = A.tab_off[e];
        const uint32_t r = S.iemb[A.prev_depth[e]];
        const uint32_t b = o[r];
        last = A.tab_idx[e] + b;
        nlast = o[r + 1] - b;
    } else {
        last = nullptr;
        nlast = 0;
    }
    return true;
}

// count positions whose vertex is not used by depths < L
__device__ __forceinline__ uint32_t aux_count_chunk(bool valid, uint32_t vidx, const uint32_t* __restrict__ cand,
                                                    const WaveAux& S, uint32_t L)
{
    bool ok = false;
    if (valid) {
        const uint32_t v = cand[vidx];
        ok = true;
        for (uint32_t j = 0; j < L; j++)
            if (S.emb[j] == v) ok = false;
    }
    return (uint32_t)__popcll(__ballot(ok));
}

__global__ void __launch_bounds__(256) k_enum_aux(const sgm_auxplan_dev* __restrict__ gA, uint32_t n_items,
                                                  uint32_t* __restrict__ stack, uint32_t stack_words,
                                                  uint32_t* __restrict__ work, sgm_stats_dev* __restrict__ gst,
                                                  uint32_t* __restrict__ stop, unsigned long long cap)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* smem = reinterpret_cast<uint8_t*>(smem_u64);
    const sgm_auxplan_dev& A = *gA;
    const uint32_t n = A.n;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    WaveAux& S = *reinterpret_cast<WaveAux*>(smem + wave * ((sizeof(WaveAux) + 15) & ~(size_t)15));
    uint32_t* ws = stack + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * stack_words;
    const uint32_t L = n - 1;
    for (uint32_t d = 0; d < SGM_MAXQ; d++) S.nodes[d] = 0;
    unsigned long long inter = 0, elems = 0, outs = 0, scanned = 0, cnt_pending = 0, items_done = 0;
    bool overflow = false;
    uint32_t poll = 0;
    while (true) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(work, 1u);
        it = uni(__shfl((int)it, 0));
        if (it >= n_items) break;
        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) break;
        items_done++;
        bool aborted = false;
        S.emb[0] = A.cand[0][it];
        S.iemb[0] = it;
        S.nodes[0] += 1;
        S.stack_top[1] = 0;
        uint32_t d = 1;
        // enter depth d (d <= L): build the list; at d == L count directly
        bool entering = true;
        while (true) {
            if ((++poll & 63u) == 0 && __atomic_load_n(stop, __ATOMIC_RELAXED)) { aborted = true; break; }
            if (entering) {
                entering = false;
                const uint32_t np = A.prev_first[d + 1] - A.prev_first[d];
                inter += np - 1;
                uint32_t top = S.stack_top[d];
                const uint32_t* list;
                uint32_t size;
                const uint32_t* last;
                uint32_t nlast;
                const bool is_last = d == L;
                if (!aux_list(A, S, d, ws, stack_words, top, list, size, is_last, last, nlast, elems, outs, stop)) {
                    overflow = true; aborted = true; break;
                }
                if (is_last) {
                    const uint32_t* cand = A.cand[d];
                    uint32_t ok = 0;
                    if (!last) {
                        scanned += size;
                        for (uint32_t base = 0; base < size; base += 64) {
                            const uint32_t i = base + lane;
                            const bool valid = i < size;
                            ok += aux_count_chunk(valid, valid ? list[i] : 0u, cand, S, L);
                        }
                    } else {
                        const uint32_t* a = list; uint32_t na = size;
                        const uint32_t* b = last; uint32_t nb = nlast;
                        if (na > nb) { const uint32_t* t = a; a = b; b = t; uint32_t tn = na; na = nb; nb = tn; }
                        elems += na + nb;
                        for (uint32_t base = 0; base < na; base += 64) {
                            const uint32_t i = base + lane;
                            const uint32_t key = i < na ? a[i] : 0u;
                            const bool hit = i < na && lower_found(b, nb, key);
                            const uint32_t h = (uint32_t)__popcll(__ballot(hit));
                            outs += h;
                            scanned += h;
                            ok += aux_count_chunk(hit, key, cand, S, L);
                        }
                    }
                    cnt_pending += ok;
                    if (cnt_pending >= FLUSH_AT) flush_count(cnt_pending, &gst->count, stop, cap);
                    S.nodes[L] += ok;
                    d--;
                    if (d == 0) break;
                    continue;
                }
                S.stack_top[d + 1] = top;
                S.lst[d] = list;
                S.tot[d] = size;
                S.pos[d] = 0;
                S.rem[d] = 0;
                continue;
            }
            if (S.rem[d] == 0 && S.pos[d] >= S.tot[d]) {
                d--;
                if (d == 0) break;
                continue;
            }
            const uint32_t* list = S.lst[d];
            if (S.rem[d] == 0) {
                const uint32_t base = S.pos[d];
                S.pos[d] = base + 64;
                S.chunk_base[d] = base;
                const uint32_t i = base + lane;
                const bool valid = i < S.tot[d];
                scanned += (S.tot[d] - base < 64u) ? (S.tot[d] - base) : 64u;
                bool ok = false;
                if (valid) {
                    const uint32_t v = A.cand[d][list[i]];
                    ok = true;
                    for (uint32_t j = 0; j < d; j++)
                        if (S.emb[j] == v) ok = false;
                }
                S.rem[d] = __ballot(ok);
                continue;
            }
            const uint64_t rm = S.rem[d];
            const uint32_t j = (uint32_t)__ffsll((long long)rm) - 1u;
            S.rem[d] = rm & (rm - 1);
            const uint32_t vidx = list[S.chunk_base[d] + j];
            S.emb[d] = A.cand[d][vidx];
            S.iemb[d] = vidx;
            S.nodes[d] += 1;
            d++;
            entering = true;
        }
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (aborted || overflow) break;
    }
    flush_count(cnt_pending, &gst->count, stop, cap);
    if (lane == 0) {
        atomicAdd(&gst->intersections, inter);
        atomicAdd(&gst->elems_read, elems);
        atomicAdd(&gst->elems_out, outs);
        atomicAdd(&gst->elems_scanned, scanned);
        atomicAdd(&gst->items_done, items_done);
        if (overflow) { atomicAdd(&gst->overflow, 1ull); atomicExch(stop, 1u); }
        for (uint32_t dd = 0; dd < n; dd++)
            if (S.nodes[dd]) atomicAdd(&gst->nodes[dd], S.nodes[dd]);
    }
}

}  // namespace sgm

// ================================================================== launch wrappers
namespace sgm {

static inline uint32_t clampg(uint64_t x, uint32_t lo, uint32_t hi)
{
    return (uint32_t)(x < lo ? lo : (x > hi ? hi : x));
}

void launch_nlf(hipStream_t s, const uint32_t* off, uint32_t V, const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_nlf, grid, dim3(256), 0, s, off, V, qs);
}

void launch_gql_phase(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* by_deg,
                      const uint32_t* ge_cnt, const sgm_qdesc* qs, uint32_t nq, uint32_t phase, uint32_t max_cand)
{
    dim3 grid(clampg(((uint64_t)max_cand + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_gql_phase, grid, dim3(256), 0, s, off, adj, by_deg, ge_cnt, qs, phase);
}

void launch_bitmaps(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V)
{
    const uint32_t W = (V + 63) / 64;
    dim3 grid(clampg((W + 3) / 4, 1, 4096), nq);
    hipLaunchKernelGGL(k_bitmaps, grid, dim3(256), 0, s, cs, V);
}

void launch_rank_scan(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V, uint32_t max_n)
{
    dim3 grid(max_n, nq);
    hipLaunchKernelGGL(k_rank_scan, grid, dim3(256), 0, s, cs, V);
}

void launch_compact(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V)
{
    const uint32_t W = (V + 63) / 64;
    dim3 grid(clampg((W + 3) / 4, 1, 4096), nq);
    hipLaunchKernelGGL(k_compact, grid, dim3(256), 0, s, cs, V);
}

void launch_table(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* vmask,
                  const uint32_t* rows, uint32_t nrows, uint32_t b, const uint64_t* bm_b, const uint32_t* rank_b,
                  uint32_t* rowcnt, uint32_t* rowoff, uint32_t* maxv, uint32_t* eidx, uint32_t* ev, int pass)
{
    dim3 grid(clampg(((uint64_t)nrows + 3) / 4, 1, 8192));
    if (pass == 0) {
        hipLaunchKernelGGL(k_table_count, grid, dim3(256), 0, s, off, adj, vmask, rows, nrows, b, rowcnt);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(256), 0, s, rowcnt, nrows, rowoff, maxv);
    } else {
        hipLaunchKernelGGL(k_table_fill, grid, dim3(256), 0, s, off, adj, vmask, rows, nrows, b, bm_b, rank_b,
                           rowoff, eidx, ev);
    }
}

size_t enum_smem_bytes(uint32_t n_ops, uint32_t waves)
{
    const size_t plan_bytes = (sizeof(sgm_plan_dev) + 15) & ~(size_t)15;
    const size_t wave_bytes = ((sizeof(WaveDFS) + 12u * n_ops) + 15) & ~(size_t)15;
    return plan_bytes + waves * wave_bytes;
}

void launch_hub_bits(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* hubs, uint32_t n_hubs,
                     uint32_t* bits, uint32_t words)
{
    hipLaunchKernelGGL(k_hub_bits, dim3(clampg(n_hubs, 1, 65535)), dim3(256), 0, s, off, adj, hubs, n_hubs, bits, words);
}

void launch_enum_aux(hipStream_t s, uint32_t blocks, const sgm_auxplan_dev* plan, uint32_t n_items, uint32_t* stack,
                     uint32_t stack_words, uint32_t* work, sgm_stats_dev* st, uint32_t* stop, unsigned long long cap)
{
    const size_t smem = 4 * ((sizeof(WaveAux) + 15) & ~(size_t)15);
    hipLaunchKernelGGL(k_enum_aux, dim3(blocks), dim3(256), smem, s, plan, n_items, stack, stack_words, work, st, stop,
                       cap);
}

void launch_enum_cache(hipStream_t s, uint32_t blocks, const uint32_t* off, const uint32_t* adj,
                       const uint32_t* vmask, const uint32_t* items, uint32_t n_items, uint32_t root_u,
                       const sgm_plan_dev* plan, uint32_t n_ops, uint32_t* stack, uint32_t stack_words,
                       uint32_t* work, sgm_stats_dev* st, uint32_t* stop, unsigned long long cap, int use_fs,
                       const uint32_t* hub_slot, const uint32_t* hub_bits, uint32_t hub_words)
{
    hipLaunchKernelGGL(k_enum_cache, dim3(blocks), dim3(256), enum_smem_bytes(n_ops, 4), s, off, adj, vmask, items,
                       n_items, root_u, plan, stack, stack_words, work, st, stop, cap, use_fs, hub_slot, hub_bits,
                       hub_words);
}

}  // namespace sgm
