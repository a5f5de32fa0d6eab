// gfx950 enumeration kernels: cache + filter + failing-set search (batched persistent
// launch) and the plain aux-table search.  All matching arithmetic is integer.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgm_internal.h"
#include "sgm_kernels.h"
#include "sgm_wave.h"

namespace sgm {

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
    // the current item when it is a split chunk (kept here, not in registers, across the
    // search): depth-1 start sp_lo (SGM_INVALID for a whole root), chunk index, root index
    // and its cutmin word
    uint32_t sp_lo, sp_it, sp_ri, sp_pad;
    uint32_t* sp_cut;
    // probe_words of the wave's current query not yet published (kept in LDS: the search is
    // at its register limit)
    unsigned long long probes;
};

#define F_ANY 1u
#define F_PRUNED 2u
#define STACK_FLAG 0x80000000u

// LDS slot of one wave in the batched search: its copy of the current query's plan, its
// search state, and per-op list location / size / alias vertex (max_ops entries each).
__host__ __device__ inline size_t enum_plan_bytes() { return (sizeof(sgm_plan_dev) + 15) & ~(size_t)15; }

__host__ __device__ inline size_t enum_wave_bytes(uint32_t max_ops)
{
    return enum_plan_bytes() + (((sizeof(WaveDFS) + 12u * max_ops) + 15) & ~(size_t)15);
}

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
    // last element <= key: every step issues its UNR loads before the first compare (into
    // separate registers), and one final load per key decides membership
    uint32_t p[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) p[u] = 0;
    uint32_t n = nb;
    while (n > 1) {
        const uint32_t half = n >> 1;
        uint32_t x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) x[u] = b[p[u] + half];
#pragma unroll
        for (int u = 0; u < UNR; u++) p[u] = x[u] <= key[u] ? p[u] + half : p[u];
        n -= half;
    }
    uint32_t x[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) x[u] = b[p[u]];
#pragma unroll
    for (int u = 0; u < UNR; u++) hit[u] = valid[u] && x[u] == key[u];
}

// Words a membership test of one key in a list of n reads under the strategy used: one
// bitmap word, or the ceil(log2 n) steps of the branch-free binary search plus its final
// load.  An intersection's algorithmic words are its keys plus their probes (probe_words).
__device__ __forceinline__ uint32_t probe_cost(uint32_t n, bool bitmap)
{
    return bitmap ? 1u : (n > 1 ? 33u - (uint32_t)__clz((int)(n - 1)) : 1u);
}

// Intersect sorted lists a and b into out (ascending); returns the size.  Lanes take
// elements of the shorter list and test them against the longer one (bitmap or binary
// search).  abits/bbits: bitmaps of a / b when they are hub adjacency lists, else null.
// Stops early (result unused) once the stop word is raised.
__device__ uint32_t isect_write(const uint32_t* __restrict__ a, uint32_t na, const uint32_t* __restrict__ abits,
                                const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bbits,
                                uint32_t* __restrict__ out, unsigned long long& elems, unsigned long long& outs,
                                unsigned long long* probes, const uint32_t* stop)
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
    if (lane == 0) *probes += (unsigned long long)na * (1u + probe_cost(nb, bbits != nullptr));
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

// Intersect two sorted lists of positions in C(u) (cand: C(u)'s vertices) into out.  A list
// that is the table row of a hub data vertex (its slot, else SGM_INVALID) is probed through
// the vertex's adjacency bitmap when the other list is walked: position x is in the row iff
// cand[x] is adjacent to the vertex -- two loads per key instead of a binary search.
__device__ uint32_t isect_pos(const uint32_t* __restrict__ a, uint32_t na, uint32_t aslot,
                              const uint32_t* __restrict__ b, uint32_t nb, uint32_t bslot,
                              const uint32_t* __restrict__ cand, const uint32_t* __restrict__ hub_bits,
                              uint32_t hub_words, uint32_t* __restrict__ out, unsigned long long& elems,
                              unsigned long long& outs, unsigned long long* probes, const uint32_t* stop)
{
    if (na > nb) {
        const uint32_t* t = a; a = b; b = t;
        uint32_t tn = na; na = nb; nb = tn;
        bslot = aslot;
    }
    if (bslot == SGM_INVALID || !hub_bits)
        return isect_write(a, na, nullptr, b, nb, nullptr, out, elems, outs, probes, stop);
    const uint32_t lane = lane_id();
    const uint32_t* __restrict__ bits = hub_bits + (size_t)bslot * hub_words;
    uint32_t cnt = 0;
    elems += na + nb;
    if (nb == 0) return 0;
    if (lane == 0) *probes += (unsigned long long)na * 3u;  // key, its vertex, the bitmap word
    for (uint32_t base = 0; base < na; base += 64 * UNR) {
        uint32_t key[UNR], w[UNR];
        bool valid[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t i = base + 64 * u + lane;
            valid[u] = i < na;
            key[u] = valid[u] ? a[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) w[u] = valid[u] ? cand[key[u]] : 0u;
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const bool hit = valid[u] && ((bits[w[u] >> 5] >> (w[u] & 31u)) & 1u);
            const uint64_t bal = __ballot(hit);
            if (hit) out[cnt + mbcnt64(bal)] = key[u];
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

// Copies the host's stop request (a word of pinned host memory) into the device stop word
// that every wave polls, at most once per SGM_RELAY_TICKS of this wave's clock.
__device__ __forceinline__ void relay_stop(uint64_t now, uint64_t& relay_at, const uint32_t* host_stop,
                                           uint32_t* dev_stop)
{
    if (now < relay_at) return;
    relay_at = now + SGM_RELAY_TICKS;
    const uint32_t h = uni(__hip_atomic_load(host_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
    if (h && lane_id() == 0) atomicExch(dev_stop, 1u);
}

// Adds a wave's unpublished time on a query (busy, ticks) to the query's total; once the total
// passes the query's budget, the query is marked timed out and its stop word raised (its
// waves leave it at their next poll).
__device__ __forceinline__ void charge_time(sgm_stats_dev* st, uint32_t* stop, unsigned long long budget,
                                            unsigned long long& busy)
{
    unsigned long long prev = 0;
    if (lane_id() == 0) prev = atomicAdd(&st->busy_ticks, busy);
    prev = __shfl(prev, 0);
    if (prev + busy > budget && lane_id() == 0) {
        atomicExch(&st->timed_out, 1ull);
        atomicExch(stop, 1u);
    }
    busy = 0;
}

// LDS state is per wave, so its values are wave-uniform; reading them through readfirstlane
// keeps them in scalar registers (a plain LDS load lands in a vector register).
__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | uni((uint32_t)x);
}

// A wave's walk over the queries of a launch whose items are the depth-0 candidates of each
// query q, [0, item_first[q+1] - item_first[q]), handed out by the query's own counter
// next[q].  Workgroups are dispatched round-robin over the 8 XCDs, so the queries are split
// into 8 classes (q mod 8): a wave serves its own class first, starting at a wave-dependent
// query, so the waves spread over the batch instead of all piling onto one query, and each
// L2 mostly holds the rows of one eighth of it.  A stopped query (cap, budget, overflow) is
// passed over without touching its items.  The wave is done once a pass over every class in
// a row found no item (counters only grow and stops are never cleared).
struct QueryWalk {
    uint32_t nq, ncls, cls, k, seen, classes, gw;
};

__device__ __forceinline__ uint32_t qw_csize(const QueryWalk& w, uint32_t c)
{
    return (w.nq - c + w.ncls - 1) / w.ncls;
}

__device__ __forceinline__ void qw_init(QueryWalk& w, uint32_t nq, uint32_t gwave)
{
    w.nq = nq;
    w.ncls = nq < 8u ? nq : 8u;
    w.cls = blockIdx.x % w.ncls;
    w.gw = gwave;
    w.k = (gwave / w.ncls) % qw_csize(w, w.cls);
    w.seen = 0;
    w.classes = 0;
}

// The next (query, item), or false when the wave is done (or the launch stopped).
__device__ bool qw_next(QueryWalk& w, const uint32_t* __restrict__ item_first, uint32_t* __restrict__ next,
                        const uint32_t* __restrict__ stops, const uint32_t* __restrict__ dev_stop, uint32_t& q,
                        uint32_t& it)
{
    while (w.classes < w.ncls) {
        if (__atomic_load_n(dev_stop, __ATOMIC_RELAXED)) return false;
        const uint32_t qq = uni(w.cls + w.ncls * w.k);
        const uint32_t nit = item_first[qq + 1] - item_first[qq];
        if (!__atomic_load_n(stops + qq, __ATOMIC_RELAXED) && __atomic_load_n(next + qq, __ATOMIC_RELAXED) < nit) {
            uint32_t t = 0;
            if (lane_id() == 0) t = atomicAdd(next + qq, 1u);
            t = uni((uint32_t)__shfl((int)t, 0));
            if (t < nit) {
                w.seen = 0;
                w.classes = 0;
                q = qq;
                it = t;
                return true;
            }
        }
        const uint32_t cs = qw_csize(w, w.cls);
        w.k = w.k + 1 == cs ? 0 : w.k + 1;
        if (++w.seen == cs) {
            w.cls = w.cls + 1 == w.ncls ? 0 : w.cls + 1;
            w.k = (w.gw / w.ncls) % qw_csize(w, w.cls);
            w.seen = 0;
            w.classes++;
        }
    }
    return false;
}

// bit v of a per-query bitmap (the last query vertex's candidates)
__device__ __forceinline__ bool last_bit(const glb<uint32_t>* __restrict__ lb, uint32_t v)
{
    return ((lb[v >> 5] >> (v & 31u)) & 1u) != 0u;
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
                                               const glb<uint32_t>* __restrict__ lbits, const WaveDFS& S,
                                               const uint64_t* anc, uint32_t L, LastRes& r)
{
    bool f[UNR];
    uint32_t c[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        f[u] = valid[u] && last_bit(lbits, w[u]);
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
        if (cfb) r.conf |= lanes_or64(cfb, cf ? (anc[L] | anc[c[u]]) : 0ull);
    }
}

// Last level, scanning an existing list.
__device__ void last_scan(const uint32_t* __restrict__ list, uint32_t size, const glb<uint32_t>* __restrict__ lbits,
                          const WaveDFS& S, const uint64_t* anc, uint32_t L, LastRes& r,
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
        classify_last4(valid, w, lbits, S, anc, L, r);
        if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    }
}

// Last level, intersecting (parent op) with N(v) and classifying in the same pass.
__device__ void last_isect(const uint32_t* __restrict__ a, uint32_t na, const uint32_t* __restrict__ abits,
                           const uint32_t* __restrict__ b, uint32_t nb, const uint32_t* __restrict__ bbits,
                           const glb<uint32_t>* __restrict__ lbits, const WaveDFS& S,
                           const uint64_t* anc, uint32_t L, LastRes& r, unsigned long long& elems,
                           unsigned long long* probes, const uint32_t* stop)
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
    if (lane == 0) *probes += (unsigned long long)na * (1u + probe_cost(nb, bbits != nullptr));
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
        classify_last4(hit, key, lbits, S, anc, L, r);
        if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    }
    r.size = size;
}

// Batched persistent launch: every query keeps its own counter over its depth-0 candidates
// (its roots in degree order, hubs first).  Wave w starts on query w mod nq and takes roots
// from that query's counter; when the query is exhausted or stopped (cap reached), the wave
// moves on to the next query, and it exits after a full lap over the batch finds no work.
// So all queries progress at once -- one query's slow start or drain overlaps the others'
// work -- and waves flow to the queries that still have roots.  A wave keeps a copy of its
// current query's plan in its own LDS slot and publishes counts and statistics to that
// query's slot when it moves on.  A batch of one query is the single-query search.
// The last level tests only bit u_L of the candidate words: a bitmap of C(u_L) (V/8 bytes
// per query instead of 4·V) keeps those probes in L2.  Wave per 64 vertices, one ballot.
// Split table of each query (layout at SGM_SPLIT_ROOTS): one wave per query walks the
// degree-ordered items until SGM_SPLIT_ROOTS candidate roots are found (or SGM_SPLIT_SCAN
// positions are seen), then cuts root r's depth-1 loop -- N(v_r), deg(v_r) elements -- into
// min(max_chunks, ceil(deg/64)) chunks of a whole number of 64-element batches
// (max_chunks <= SGM_SPLIT_CHUNKS).  A
// root range of at most SGM_SPLIT_ROOTS ids is read directly instead: all its candidates
// are split and no root is left to the item counter (root_first = n_items).
__global__ void __launch_bounds__(64) k_split_roots(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ items,
                                                    const sgm_equery* __restrict__ qs, uint32_t V,
                                                    uint32_t max_chunks)
{
    const sgm_equery& Q = qs[blockIdx.x];
    uint32_t* __restrict__ sp = Q.split;
    if (!sp) return;
    const uint32_t lane = lane_id();
    const uint32_t Rn = SGM_SPLIT_ROOTS;
    uint32_t* __restrict__ rvtx = sp + 2;
    uint32_t* __restrict__ cutmin = rvtx + Rn;
    uint32_t* __restrict__ done = cutmin + Rn;
    uint32_t* __restrict__ it3 = done + SGM_SPLIT_ITEMS;
    uint32_t found = 0, root_first = 0;
    const uint32_t lim = Q.n_items < SGM_SPLIT_SCAN ? Q.n_items : SGM_SPLIT_SCAN;
    if (Q.plan->n > 2 && Q.root_hi - Q.root_lo <= Rn) {
        // the candidates of a narrow root range, in id order (Rn <= 64: one ballot)
        const uint32_t v = Q.root_lo + lane;
        const bool c = lane < Q.root_hi - Q.root_lo && v < V && ((Q.vmask[v] >> Q.root_u) & 1u);
        const uint64_t b = __ballot(c);
        if (c) rvtx[mbcnt64(b)] = v;
        found = (uint32_t)__popcll(b);
        root_first = Q.n_items;
    } else if (Q.plan->n > 2) {  // two-vertex queries have no depth-1 loop to split
        for (uint32_t base = 0; base < lim && found < Rn; base += 64) {
            const uint32_t p = base + lane;
            const uint32_t v = p < lim ? items[p] : 0u;
            const bool c = p < lim && ((Q.vmask[v] >> Q.root_u) & 1u) && v >= Q.root_lo && v < Q.root_hi;
            const uint64_t b = __ballot(c);
            const uint32_t r = found + mbcnt64(b);
            if (c && r < Rn) rvtx[r] = v;
            const uint32_t cnt = (uint32_t)__popcll(b);
            if (found + cnt >= Rn) {
                uint64_t bb = b;  // the (Rn - found)-th set bit is the last split root
                for (uint32_t k = 0; k + 1 < Rn - found; k++) bb &= bb - 1;
                root_first = base + (uint32_t)__ffsll((long long)bb);
                found = Rn;
            } else {
                found += cnt;
            }
        }
        if (found < Rn) root_first = lim;  // every candidate below lim is split
    }
    // chunks per root, then their (root, lo, hi) triples in root order
    uint32_t deg = 0, nch = 0, W = 64;
    if (lane < found) {
        const uint32_t v = rvtx[lane];
        deg = off[v + 1] - off[v];
        const uint32_t b64 = (deg + 63) / 64;
        const uint32_t ch = b64 < max_chunks ? b64 : max_chunks;
        W = ch ? ((b64 + ch - 1) / ch) * 64 : 64;
        nch = deg ? (deg + W - 1) / W : 1u;  // a root without neighbours still counts its node
        cutmin[lane] = SGM_INVALID;
    }
    const uint32_t incl = wave_incl_scan(nch);
    const uint32_t total = (uint32_t)__shfl((int)incl, 63);
    for (uint32_t c = 0; c < nch; c++) {
        const uint32_t s = incl - nch + c;
        it3[3 * s] = lane;
        it3[3 * s + 1] = c * W;
        it3[3 * s + 2] = (c + 1) * W < deg ? (c + 1) * W : deg;
    }
    for (uint32_t s = lane; s < total; s += 64) done[s] = 0;
    if (lane == 0) {
        sp[0] = total;
        sp[1] = root_first;
    }
}

// Adds the split chunks' records to their queries' statistics: chunk s of root r counts iff
// it ran to its end and starts at or before r's first cut-off (cutmin[r]).  Grid
// (SGM_SPLIT_ITEMS / 256, queries).
__global__ void __launch_bounds__(256) k_split_reduce(const sgm_equery* __restrict__ qs)
{
    const sgm_equery& Q = qs[blockIdx.y];
    const uint32_t* __restrict__ sp = Q.split;
    if (!sp) return;
    const uint32_t ns = sp[0];
    if (blockIdx.x * blockDim.x >= ns) return;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t* __restrict__ cutmin = sp + 2 + SGM_SPLIT_ROOTS;
    const uint32_t* __restrict__ done = cutmin + SGM_SPLIT_ROOTS;
    const uint32_t* __restrict__ it3 = done + SGM_SPLIT_ITEMS;
    bool inc = false;
    if (s < ns) inc = done[s] && it3[3 * s + 1] <= cutmin[it3[3 * s]];
    __shared__ unsigned long long red[4];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    for (uint32_t f = 0; f < SGM_SPLIT_FIELDS; f++) {
        unsigned long long v = inc ? Q.rec[(size_t)f * SGM_SPLIT_ITEMS + s] : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
            v += ((unsigned long long)hi << 32) | lo;
        }
        if (lane == 0) red[wave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t = red[0] + red[1] + red[2] + red[3];
            unsigned long long* dst = f < SGM_MAXQ ? &Q.st->nodes[f]
                                    : f == 32u ? &Q.st->intersections
                                    : f == 33u ? &Q.st->fs_prunes
                                    : f == 34u ? &Q.st->elems_read
                                    : f == 35u ? &Q.st->elems_out
                                    : f == 36u ? &Q.st->elems_scanned
                                               : &Q.st->probe_words;
            if (t) atomicAdd(dst, t);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_last_bits(const sgm_equery* __restrict__ qs, uint32_t V)
{
    const sgm_equery& Q = qs[blockIdx.y];
    const uint32_t uL = Q.plan->order[Q.plan->n - 1];
    const uint32_t W = (V + 31) / 32, lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t g = blockIdx.x * waves + (threadIdx.x >> 6); g * 64 < V; g += gridDim.x * waves) {
        const uint32_t v = g * 64 + lane;
        const uint64_t b = __ballot(v < V && ((Q.vmask[v] >> uL) & 1u));
        if (lane == 0) Q.lastbits[2 * g] = (uint32_t)b;
        if (lane == 1 && 2 * g + 1 < W) Q.lastbits[2 * g + 1] = (uint32_t)(b >> 32);
    }
}

struct WaveAcc {
    unsigned long long inter, fsp, elems, outs, scanned, items_done, cnt_pending, busy;
};

// Per-depth node counters: lane d publishes (when out != null) and clears depth d, one LDS
// address per lane (a loop over the depths would keep one address register per depth).
__device__ __forceinline__ void clear_nodes(WaveDFS& S, unsigned long long* out, uint32_t n)
{
    const uint32_t lane = lane_id();
    if (lane < SGM_MAXQ) {
        if (out && lane < n && S.nodes[lane]) atomicAdd(out + lane, S.nodes[lane]);
        S.nodes[lane] = 0;
    }
    // later uniform reads of any depth's counter see the clears
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ void flush_stats(const sgm_equery& Q, WaveDFS& S, WaveAcc& a, uint32_t n, bool overflow)
{
    flush_count(a.cnt_pending, &Q.st->count, Q.stop, Q.cap);
    if (lane_id() == 0) {
        if (a.inter) atomicAdd(&Q.st->intersections, a.inter);
        if (a.fsp) atomicAdd(&Q.st->fs_prunes, a.fsp);
        if (a.elems) atomicAdd(&Q.st->elems_read, a.elems);
        if (a.outs) atomicAdd(&Q.st->elems_out, a.outs);
        if (a.scanned) atomicAdd(&Q.st->elems_scanned, a.scanned);
        if (a.items_done) atomicAdd(&Q.st->items_done, a.items_done);
        if (a.busy) atomicAdd(&Q.st->busy_ticks, a.busy);
        if (overflow) { atomicAdd(&Q.st->overflow, 1ull); atomicExch(Q.stop, 1u); }
    }
    a = WaveAcc{0, 0, 0, 0, 0, 0, 0, 0};
    const unsigned long long pw = uni64(S.probes);
    if (lane_id() == 0) {
        if (pw) atomicAdd(&Q.st->probe_words, pw);
        S.probes = 0;
    }
    clear_nodes(S, &Q.st->nodes[0], n);
}

// ---- last level for all remaining siblings of a chunk at once
//
// At depth d = n-2 every extension v of the current chunk (the set bits of S.rem[d]) ends
// in one last-level evaluation over the candidates w of u_L: (parent op) ∩ N(v) when the
// last-level op has a parent, N(v) when it has none, or a list cached at an earlier depth
// when depth d computes no op.  One sibling at a time, each evaluation is a chain of
// dependent loads (keys, membership, filter word) over a few dozen elements.  Here the
// siblings' (v, element) pairs are concatenated -- wave prefix sum of the per-sibling
// lengths, each the shorter list of its intersection -- and 64 * UNR pairs are evaluated
// per step; per-sibling sums accumulate in the sibling's own lane (lane i = i-th sibling).
// A cached list is evaluated once for the chunk: sibling v then differs only at w == v
// (an extension becomes a conflict at depth d).  Results are applied in sibling order with
// the same failing-set cut-off as one at a time, so the search, its counts and its
// statistics are unchanged; siblings after a cut-off were evaluated for nothing.

// Membership of K keys, each in its own sorted list b[u] (length nb[u] > 0) or bitmap.
// Branch-free search for the last element <= key: every step issues the K loads together
// (lanes with nothing to search read safe[0]); the wave steps until its longest list is done.
template <int K>
__device__ __forceinline__ void member4v(const uint32_t* const (&b)[K], const uint32_t (&nb)[K],
                                         const uint32_t* const (&bits)[K], const uint32_t (&key)[K],
                                         const bool (&valid)[K], bool (&hit)[K], const uint32_t* safe)
{
    uint32_t p[K], n[K];
    const uint32_t* bb[K];
    bool more = false;
#pragma unroll
    for (int u = 0; u < K; u++) {
        const bool srch = valid[u] && !bits[u];
        p[u] = 0;
        n[u] = srch ? nb[u] : 1u;
        bb[u] = srch ? b[u] : safe;
        more |= n[u] > 1;
    }
    while (__ballot(more)) {
        more = false;
        uint32_t x[K], half[K];
#pragma unroll
        for (int u = 0; u < K; u++) {
            half[u] = n[u] >> 1;
            x[u] = bb[u][p[u] + half[u]];
        }
#pragma unroll
        for (int u = 0; u < K; u++) {
            p[u] = x[u] <= key[u] ? p[u] + half[u] : p[u];
            n[u] -= half[u];
            more |= n[u] > 1;
        }
    }
    // one load per key, bitmap word or list element, all issued before the first use
    uint32_t wd[K];
#pragma unroll
    for (int u = 0; u < K; u++) wd[u] = bits[u] ? bits[u][key[u] >> 5] : bb[u][p[u]];
#pragma unroll
    for (int u = 0; u < K; u++) {
        const bool e = bits[u] ? ((wd[u] >> (key[u] & 31u)) & 1u) != 0u : wd[u] == key[u];
        hit[u] = valid[u] && e;
    }
}

#define LAST_PU 2  // 4 gives more loads in flight per wave but costs a wave per SIMD

struct SibAcc {
    uint32_t ok, size, any;
    uint64_t conf;
};

// Adds one 64-pair group's results (pairs gbase .. gbase+63 of the concatenation) to the
// owner lanes: owner lane j's pairs are [excl_j, incl_j), so its share of the group is a
// contiguous lane range, counted from the wave-wide ballots with mask arithmetic alone.
// Conflicts (rare) are added by their own lanes' owners.
__device__ __forceinline__ void sib_accumulate(SibAcc& r, uint32_t excl, uint32_t incl, uint32_t gbase,
                                               uint32_t own, bool hit, bool filt, uint32_t c,
                                               const uint64_t* anc, uint32_t L)
{
    const uint32_t lane = lane_id();
    const bool ok = filt && c == SGM_INVALID, cf = filt && c != SGM_INVALID;
    const uint64_t H = __ballot(hit), O = __ballot(ok), F = __ballot(filt);
    uint64_t CF = __ballot(cf);
    if (H) {
        const uint32_t lo = excl > gbase ? (excl - gbase < 64u ? excl - gbase : 64u) : 0u;
        const uint32_t hi = incl > gbase ? (incl - gbase < 64u ? incl - gbase : 64u) : 0u;
        const uint64_t below_hi = hi >= 64u ? ~0ull : ((1ull << hi) - 1ull);
        const uint64_t below_lo = lo >= 64u ? ~0ull : ((1ull << lo) - 1ull);
        const uint64_t seg = below_hi & ~below_lo;
        r.size += (uint32_t)__popcll(H & seg);
        r.ok += (uint32_t)__popcll(O & seg);
        r.any |= (F & seg) ? 1u : 0u;
    }
    while (CF) {
        const uint32_t l = (uint32_t)__ffsll((long long)CF) - 1u;
        CF &= CF - 1;
        const uint32_t j = rdlane(own, l), cc = rdlane(c, l);
        if (lane == j) r.conf |= anc[L] | anc[cc];
    }
}

__device__ void last_chunk(const sgm_plan_dev& P, WaveDFS& S, uint32_t d, const uint32_t* __restrict__ list,
                           const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                           const uint32_t* __restrict__ ws, const uint32_t* op_off, const uint32_t* op_size,
                           const uint32_t* op_vtx, const uint32_t* __restrict__ hub_slot,
                           const uint32_t* __restrict__ hub_bits, uint32_t hub_words,
                           const glb<uint32_t>* __restrict__ lbits, uint32_t embv, int use_fs, WaveAcc& acc,
                           const uint32_t* stop)
{
    const uint32_t lane = lane_id();
    const uint32_t L = d + 1;
    const uint64_t rm = uni64(S.rem[d]);
    const uint32_t m = (uint32_t)__popcll(rm);
    const bool mine = lane < m;
    // chunk position of the lane-th set bit of rm: the largest p with at most `lane` set
    // bits below it
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t s = 32; s > 0; s >>= 1) {
        const uint32_t q = pos + s;
        const uint64_t below = q >= 64u ? ~0ull : ((1ull << q) - 1ull);
        if ((uint32_t)__popcll(rm & below) <= lane) pos = q;
    }
    pos = pos > 63u ? 63u : pos;
    const uint32_t v = mine ? list[uni(S.chunk_base[d]) + pos] : 0u;
    auto bits_of = [&](uint32_t x) -> const uint32_t* {
        if (!hub_slot) return nullptr;
        const uint32_t s = hub_slot[x];
        return s == SGM_INVALID ? nullptr : hub_bits + (size_t)s * hub_words;
    };
    const uint32_t o = uni(P.op_first[d]);
    const bool has_op = o < uni(P.op_first[d + 1]);
    const uint32_t par = has_op ? uni(P.op_parent[o]) : SGM_NOPARENT;
    const int mode = !has_op ? 2 : (par == SGM_NOPARENT ? 1 : 0);  // 0 isect, 1 N(v), 2 cached list
    SibAcc r = {0, 0, 0, 0};
    uint32_t nbs = 0, pn = 0;
    uint32_t pcost = 0;  // this sibling's intersection: keys + probe words (probe_cost)
    constexpr int PU = LAST_PU;  // pairs per lane per step
    if (mode == 2) {
        const uint32_t ci = uni(P.iter_set[L]);
        const uint32_t* __restrict__ cl = op_list(uni(op_off[ci]), adj, ws);
        const uint32_t cn = uni(op_size[ci]);
        uint32_t b_ok = 0, b_any = 0;
        uint64_t b_conf = 0;
        for (uint32_t base = 0; base < cn; base += 64 * UNR) {
            uint32_t w[UNR];
            bool f[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const uint32_t i = base + 64 * u + lane;
                f[u] = i < cn;
                w[u] = f[u] ? cl[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) f[u] = f[u] && last_bit(lbits, w[u]);
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                uint32_t c = SGM_INVALID;
                for (uint32_t j = 0; j < d; j++)
                    if (rdlane(embv, j) == w[u]) c = j;
                const bool ok = f[u] && c == SGM_INVALID, cf = f[u] && c != SGM_INVALID;
                const uint64_t okb = __ballot(ok), cfb = __ballot(cf);
                b_ok += (uint32_t)__popcll(okb);
                if (okb | cfb) b_any = 1;
                if (cfb) b_conf |= lanes_or64(cfb, cf ? (P.anc[L] | P.anc[c]) : 0ull);
            }
            if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
        }
        // sibling v: an extension w == v of the cached list becomes a conflict at depth d
        const bool self = mine && lower_found(cl, cn, v) && last_bit(lbits, v);
        r.ok = b_ok - (self ? 1u : 0u);
        r.size = cn;
        r.any = b_any;
        r.conf = b_conf | (self ? (P.anc[L] | P.anc[d]) : 0ull);
    } else {
        uint32_t nbo = 0;
        if (mine) { nbo = off[v]; nbs = off[v + 1] - nbo; }
        const uint32_t* __restrict__ pl = nullptr;
        const uint32_t* pbits = nullptr;
        if (mode == 0) {
            pl = op_list(uni(op_off[par]), adj, ws);
            pn = uni(op_size[par]);
            const uint32_t pv = uni(op_vtx[par]);
            pbits = pv != SGM_INVALID ? bits_of(pv) : nullptr;
        }
        // keys of sibling i: N(v) when it is the shorter list (always without a parent op),
        // else the parent op, tested against N(v) (its bitmap when v is a hub)
        const uint32_t kv = (mode == 1 || pn > nbs) ? 1u : 0u;
        const uint32_t len = mine ? (mode == 1 ? nbs : (pn < nbs ? pn : nbs)) : 0u;
        const uint32_t slot = (mine && !kv && hub_slot) ? hub_slot[v] : SGM_INVALID;
        const uint32_t incl = wave_incl_scan(len);
        const uint32_t total = rdlane(incl, 63);
        const uint32_t excl = incl - len;
        // pair t -> owner lane and key; the keys of pass i + 1 are loaded while pass i's
        // membership and filter probes are in flight (software pipelining over passes)
        auto load_keys = [&](uint32_t b0, uint32_t (&k)[PU], uint32_t (&ow)[PU], bool (&va)[PU]) {
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const uint32_t t = b0 + 64 * u + lane;
                ow[u] = wave_seg_owner(incl, t);
                va[u] = t < total;
                const uint32_t q = t - (uint32_t)__shfl((int)excl, (int)ow[u]);
                const uint32_t onbo = (uint32_t)__shfl((int)nbo, (int)ow[u]);
                const uint32_t okv = (uint32_t)__shfl((int)kv, (int)ow[u]);
                k[u] = va[u] ? (okv ? adj[onbo + q] : pl[q]) : 0u;
            }
        };
        uint32_t key[PU], own[PU];
        bool valid[PU];
        load_keys(0, key, own, valid);
        for (uint32_t base = 0; base < total; base += 64 * PU) {
            uint32_t ov[PU], tn[PU];
            const uint32_t* tb[PU];
            const uint32_t* tbits[PU];
            bool hit[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const uint32_t onbo = (uint32_t)__shfl((int)nbo, (int)own[u]);
                const uint32_t onbs = (uint32_t)__shfl((int)nbs, (int)own[u]);
                const uint32_t okv = (uint32_t)__shfl((int)kv, (int)own[u]);
                const uint32_t oslot = (uint32_t)__shfl((int)slot, (int)own[u]);
                ov[u] = (uint32_t)__shfl((int)v, (int)own[u]);
                tb[u] = okv ? pl : adj + onbo;
                tn[u] = okv ? pn : onbs;
                tbits[u] = okv ? pbits
                               : (oslot == SGM_INVALID ? nullptr : hub_bits + (size_t)oslot * hub_words);
            }
            // the filter words, loaded with the keys: they do not depend on the membership test
            uint32_t lw[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) lw[u] = valid[u] ? lbits[key[u] >> 5] : 0u;
            uint32_t nkey[PU], nown[PU];
            bool nvalid[PU];
            if (base + 64 * PU < total) {
                load_keys(base + 64 * PU, nkey, nown, nvalid);
            } else {
#pragma unroll
                for (int u = 0; u < PU; u++) { nkey[u] = 0; nown[u] = 0; nvalid[u] = false; }
            }
            if (mode == 1) {
#pragma unroll
                for (int u = 0; u < PU; u++) hit[u] = valid[u];
            } else {
                member4v(tb, tn, tbits, key, valid, hit, adj);
            }
            bool filt[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) filt[u] = hit[u] && ((lw[u] >> (key[u] & 31u)) & 1u) != 0u;
#pragma unroll
            for (int u = 0; u < PU; u++) {
                uint32_t c = SGM_INVALID;
                if (filt[u]) {
                    for (uint32_t j = 0; j < d; j++)
                        if (rdlane(embv, j) == key[u]) c = j;
                    if (key[u] == ov[u]) c = d;
                }
                sib_accumulate(r, excl, incl, base + 64 * u, own[u], hit[u], filt[u], c, P.anc, L);
            }
            if (((base >> 8) & 7u) == 7u && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
#pragma unroll
            for (int u = 0; u < PU; u++) { key[u] = nkey[u]; own[u] = nown[u]; valid[u] = nvalid[u]; }
        }
        // (after the pair loop, so that it holds no register during it)
        if (mode == 0 && mine && len)
            pcost = len * (1u + probe_cost(kv ? pn : nbs, kv ? pbits != nullptr : slot != SGM_INVALID));
    }
    // apply the siblings' results in order (inc/explore.h:1036-1066 per extension), all lanes
    // at once: sibling i's failing mask f_i; the first sibling whose mask omits depth d cuts
    // the rest (it and those before it count); the parent's mask is the OR of the counted
    // masks, or the cutting sibling's mask alone
    uint64_t f = 0;
    if (use_fs && mine) {
        if (mode != 2 && r.size == 0) f = P.anc[P.op_vdepth[o]];
        else if (!r.any) f = P.anc[L];
        else if (r.ok) f = ~0ull;
        else f = r.conf;
    }
    const uint64_t cut = __ballot(use_fs && mine && !((f >> d) & 1ull));
    const uint32_t last = cut ? (uint32_t)__ffsll((long long)cut) - 1u : m - 1u;  // last counted sibling
    const bool counted = mine && lane <= last;
    const uint32_t ok_sum = rdlane(wave_incl_scan(counted ? r.ok : 0u), 63);
    const uint32_t size_sum = rdlane(wave_incl_scan(counted ? r.size : 0u), 63);
    const uint32_t n_counted = last + 1u;
    S.emb[d] = rdlane(v, last);
    S.nodes[d] += n_counted;
    S.nodes[L] += ok_sum;
    acc.cnt_pending += ok_sum;
    acc.scanned += size_sum;
    if (mode == 0) {
        acc.inter += n_counted;
        acc.elems += (unsigned long long)pn * n_counted + rdlane(wave_incl_scan(counted ? nbs : 0u), 63);
        const uint32_t pw = rdlane(wave_incl_scan(counted ? pcost : 0u), 63);
        if (lane == 0) S.probes += pw;
        acc.outs += size_sum;
    }
    if (use_fs) {
        const uint64_t fl = ((uint64_t)rdlane((uint32_t)(f >> 32), last) << 32) | rdlane((uint32_t)f, last);
        S.fail[d] = fl;
        if (cut) {
            S.fail[d - 1] = fl;
            S.flags[d] |= F_PRUNED;
            acc.fsp++;
        } else {
            // an extension with embeddings has the full mask; otherwise OR the few others
            const uint64_t all = __ballot(counted && f == ~0ull) ? ~0ull : lanes_or64(__ballot(counted), f);
            S.fail[d - 1] |= all;
        }
    }
    S.rem[d] = 0;
}

// SPLIT: the launch holds queries with split chunks (the variant without them keeps the
// register budget of the plain search).
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_enum_cache(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ adj,
                                                    const uint32_t* __restrict__ items,
                                                    const sgm_equery* __restrict__ qs, uint32_t nq,
                                                    uint32_t max_ops,
                                                    uint32_t* __restrict__ stack, uint32_t stack_words,
                                                    uint32_t* __restrict__ dev_stop,
                                                    const uint32_t* __restrict__ host_stop, int use_fs,
                                                    const uint32_t* __restrict__ hub_slot,
                                                    const uint32_t* __restrict__ hub_bits, uint32_t hub_words,
                                                    int serial_last)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* smem = reinterpret_cast<uint8_t*>(smem_u64);
    const uint32_t wave = uni(threadIdx.x >> 6), lane = lane_id();  // uniform: keeps per-wave values scalar
    const size_t wave_bytes = enum_wave_bytes(max_ops);
    uint8_t* wbase = smem + wave * wave_bytes;
    sgm_plan_dev& P = *reinterpret_cast<sgm_plan_dev*>(wbase);
    WaveDFS& S = *reinterpret_cast<WaveDFS*>(wbase + enum_plan_bytes());
    uint32_t* op_off = reinterpret_cast<uint32_t*>(wbase + enum_plan_bytes() + sizeof(WaveDFS));
    uint32_t* op_size = op_off + max_ops;
    uint32_t* op_vtx = op_size + max_ops;  // vertex of an adjacency (alias) op, else SGM_INVALID
    uint32_t* ws = stack + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * stack_words;
    // bitmap of a vertex's adjacency list, or null when it has none
    auto vbits = [&](uint32_t v) -> const uint32_t* {
        if (!hub_slot) return nullptr;
        const uint32_t s = hub_slot[v];
        return s == SGM_INVALID ? nullptr : hub_bits + (size_t)s * hub_words;
    };

    clear_nodes(S, nullptr, 0);
    WaveAcc acc = {0, 0, 0, 0, 0, 0, 0, 0};
    if (lane == 0) S.probes = 0;
    uint32_t cur = SGM_INVALID;  // query whose plan sits in this wave's LDS
    uint32_t n = 0, L = 0, root_u = 0;
    const glb<uint32_t>* lastbits = nullptr;  // bitmap of C(u_L) of the current query
    uint32_t embv = 0;  // lane j: the vertex mapped at depth j (as S.emb[j]; readlane is cheaper than LDS)
    const glb<uint32_t>* vmask = nullptr;  // global-typed: loaded from the descriptor, otherwise flat
    uint32_t* stop = nullptr;
    bool overflow = false;
    uint32_t poll = 0;
    uint64_t relay_at = 0;

    // Visiting order (L2 locality): workgroups are dispatched round-robin over the 8 XCDs, so
    // blocks with the same blockIdx.x mod 8 share an L2.  Queries are split into 8 classes
    // (q mod 8); a wave serves its own class first (starting at a wave-dependent query of
    // the class) and only then helps the other classes, so each L2 mostly holds the
    // candidate words of one eighth of the batch.
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + wave;
    const uint32_t ncls = nq < 8u ? nq : 8u;
    auto csize = [&](uint32_t c) { return (nq - c + ncls - 1) / ncls; };
    uint32_t cls = blockIdx.x % ncls;
    uint32_t k = (gwave / ncls) % csize(cls), seen = 0, classes = 0;
    while (classes < ncls) {
        if (__atomic_load_n(dev_stop, __ATOMIC_RELAXED)) break;
        if ((++poll & 63u) == 0) relay_stop(wall_clock64(), relay_at, host_stop, dev_stop);
        const uint32_t q = cls + ncls * k;
        const sgm_equery& Q = qs[q];
        uint32_t it = SGM_INVALID;
        bool split_item = false;
        if (!__atomic_load_n(Q.stop, __ATOMIC_RELAXED)) {
            // split chunks of the heaviest roots first, then the remaining roots
            const uint32_t ns = (SPLIT && Q.split) ? uni(Q.split[0]) : 0u;
            const uint32_t rf = (SPLIT && Q.split) ? uni(Q.split[1]) : 0u;
            if (ns && __atomic_load_n(Q.snext, __ATOMIC_RELAXED) < ns) {
                if (lane == 0) it = atomicAdd(Q.snext, 1u);
                it = uni((uint32_t)__shfl((int)it, 0));
                split_item = it < ns;
                if (!split_item) it = SGM_INVALID;
            }
            if (!split_item && __atomic_load_n(Q.next, __ATOMIC_RELAXED) < Q.n_items - rf) {
                if (lane == 0) it = atomicAdd(Q.next, 1u);
                it = uni((uint32_t)__shfl((int)it, 0));
                it = it < Q.n_items - rf ? it + rf : SGM_INVALID;
            }
        }
        if (it == SGM_INVALID) {
            // exhausted or stopped (cap reached / overflow): try the next query of the class,
            // then the next class; a full pass over every class without work ends the wave
            const uint32_t cs = csize(cls);
            k = k + 1 == cs ? 0 : k + 1;
            if (++seen == cs) {
                cls = cls + 1 == ncls ? 0 : cls + 1;
                k = (gwave / ncls) % csize(cls);
                seen = 0;
                classes++;
            }
            continue;
        }
        seen = 0;
        classes = 0;
        if (q != cur) {
            if (cur != SGM_INVALID) flush_stats(qs[cur], S, acc, n, false);
            const uint32_t* src = reinterpret_cast<const uint32_t*>(Q.plan);
            uint32_t* dst = reinterpret_cast<uint32_t*>(wbase);
            for (uint32_t i = lane; i < sizeof(sgm_plan_dev) / 4; i += 64) dst[i] = src[i];
            // every lane reads the whole copy: order the LDS stores before those loads
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            cur = q;
            n = P.n;
            L = n - 1;
            root_u = Q.root_u;
            vmask = gcast(Q.vmask);
            lastbits = gcast(Q.lastbits);
            stop = Q.stop;
        }
        const unsigned long long cap = Q.cap;
        // a split chunk: root sri, depth-1 positions [lo1, hi1)
        uint32_t v0, lo1 = 0, hi1 = 0;
        if (SPLIT && split_item) {
            const uint32_t* t3 = Q.split + 2 + 2 * SGM_SPLIT_ROOTS + SGM_SPLIT_ITEMS + 3 * it;
            const uint32_t sri = uni(t3[0]);
            lo1 = uni(t3[1]);
            hi1 = uni(t3[2]);
            uint32_t* cutw = Q.split + 2 + SGM_SPLIT_ROOTS + sri;
            v0 = uni(Q.split[2 + sri]);
            // an earlier chunk of this root was cut off: the sequential search never gets here
            if (uni(__atomic_load_n(cutw, __ATOMIC_RELAXED)) < lo1) continue;
            flush_stats(Q, S, acc, n, false);  // the chunk's counters go to its record
            S.sp_lo = lo1;
            S.sp_it = it;
            S.sp_ri = sri;
            S.sp_cut = cutw;
        } else {
            v0 = items[it];
            if (SPLIT) S.sp_lo = SGM_INVALID;
        }
        // not (or no longer) a candidate of u0, or outside the launch's root range
        if (!((vmask[v0] >> root_u) & 1u) || v0 < Q.root_lo || v0 >= Q.root_hi) continue;
        acc.items_done++;
        uint64_t t_mark = wall_clock64();  // the item's time up to here is in acc.busy
        S.emb[0] = v0;
        embv = lane == 0 ? v0 : embv;
        if (lo1 == 0) S.nodes[0] += 1;  // a split root's node is counted by its first chunk
        S.stack_top[0] = 0;
        bool aborted = false;

        // ---- ops of depth 0 (or the whole last level when n == 2)
        if (n == 2) {
            LastRes r = {0, 0, 0, 0};
            const uint32_t o = P.iter_set[1];
            op_off[o] = off[v0];
            op_size[o] = off[v0 + 1] - off[v0];
            last_scan(adj + off[v0], op_size[o], lastbits, S, P.anc, L, r, stop);
            acc.scanned += r.size;
            acc.cnt_pending += r.ok;
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
                    acc.inter++;
                    const uint32_t sz = isect_write(op_list(op_off[p], adj, ws), op_size[p],
                                                    op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr,
                                                    adj + nbo, nbs, vbits(v0), ws + top, acc.elems, acc.outs,
                                                    &S.probes, stop);
                    op_off[o] = top | STACK_FLAG;
                    op_size[o] = sz;
                    op_vtx[o] = SGM_INVALID;
                    top += sz;
                }
                if (op_size[o] == 0) { ok = false; break; }
            }
            if (ok) {
                S.stack_top[1] = top;
                // enter depth 1 (a split chunk walks only its part of the loop)
                S.pos[1] = lo1; S.rem[1] = 0; S.flags[1] = 0;
                S.tot[1] = (SPLIT && split_item) ? hi1 : op_size[P.iter_set[1]];
                S.lst[1] = op_off[P.iter_set[1]];
                S.fail[0] = 0;
                uint32_t d = 1;
                while (true) {
                    if ((++poll & 3u) == 0) {
                        bool end = __atomic_load_n(stop, __ATOMIC_RELAXED) | __atomic_load_n(dev_stop, __ATOMIC_RELAXED);
                        if ((poll & 255u) == 0) {  // every 256 steps: the clock (stop relay, time budget)
                            const uint64_t now = wall_clock64();
                            relay_stop(now, relay_at, host_stop, dev_stop);
                            acc.busy += now - t_mark;
                            t_mark = now;
                            if (Q.budget && acc.busy >= SGM_CHARGE_TICKS) charge_time(Q.st, stop, Q.budget, acc.busy);
                        }
                        const uint32_t slo = SPLIT ? uni(S.sp_lo) : SGM_INVALID;
                        if (slo != SGM_INVALID) {  // an earlier chunk of this split root was cut off
                            const uint32_t* cw = reinterpret_cast<const uint32_t*>(
                                uni64(reinterpret_cast<uint64_t>(S.sp_cut)));
                            end |= uni(__atomic_load_n(cw, __ATOMIC_RELAXED)) < slo;
                        }
                        if (end) {
                            aborted = true;
                            break;
                        }
                    }
                    const uint64_t rem_d = uni64(S.rem[d]);
                    if (rem_d == 0 && ((uni(S.flags[d]) & F_PRUNED) || uni(S.pos[d]) >= uni(S.tot[d]))) {
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
                                acc.fsp++;
                            } else S.fail[d - 1] |= S.fail[d];
                        }
                        continue;
                    }
                    const uint32_t* list = op_list(uni(S.lst[d]), adj, ws);
                    if (rem_d == 0) {
                        // classify the next 64 elements of depth d's iteration set
                        const uint32_t u = uni(P.order[d]);
                        const uint32_t base = uni(S.pos[d]), tot = uni(S.tot[d]);
                        S.pos[d] = base + 64;
                        acc.scanned += (tot - base < 64u) ? (tot - base) : 64u;
                        S.chunk_base[d] = base;
                        const uint32_t i = base + lane;
                        const bool valid = i < tot;
                        const uint32_t w = valid ? list[i] : 0u;
                        const bool f = valid && ((vmask[w] >> u) & 1u);
                        uint32_t c = SGM_INVALID;
                        if (f)
                            for (uint32_t j = 0; j < d; j++)
                                if (rdlane(embv, j) == w) c = j;
                        const uint64_t okb = __ballot(f && c == SGM_INVALID);
                        const uint64_t cfb = __ballot(f && c != SGM_INVALID);
                        if (okb | cfb) S.flags[d] |= F_ANY;
                        if (use_fs && cfb) {
                            const uint64_t m = (f && c != SGM_INVALID) ? (P.anc[d] | P.anc[c]) : 0ull;
                            S.fail[d - 1] |= lanes_or64(cfb, m);
                        }
                        S.rem[d] = okb;
                        continue;
                    }
                    if (d == n - 2 && !serial_last) {
                        // the last level of every remaining extension of this chunk
                        last_chunk(P, S, d, list, off, adj, ws, op_off, op_size, op_vtx, hub_slot, hub_bits,
                                   hub_words, lastbits, embv, use_fs, acc, stop);
                        if (acc.cnt_pending >= FLUSH_AT) flush_count(acc.cnt_pending, &Q.st->count, stop, cap);
                        continue;
                    }
                    // extend with the next candidate of this chunk
                    const uint64_t rm = rem_d;
                    const uint32_t j = (uint32_t)__ffsll((long long)rm) - 1u;
                    S.rem[d] = rm & (rm - 1);
                    const uint32_t v = uni(list[uni(S.chunk_base[d]) + j]);
                    S.emb[d] = v;
                    embv = lane == d ? v : embv;
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
                                last_scan(adj + nbo, nbs, lastbits, S, P.anc, L, r, stop);
                            } else {
                                acc.inter++;
                                last_isect(op_list(op_off[p], adj, ws), op_size[p],
                                           op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr, adj + nbo, nbs,
                                           vbits(v), lastbits, S, P.anc, L, r, acc.elems, &S.probes, stop);
                                acc.outs += r.size;
                            }
                            if (r.size == 0) { empty = true; conf_v = P.op_vdepth[o]; }
                        } else {
                            const uint32_t o = P.iter_set[L];
                            last_scan(op_list(op_off[o], adj, ws), op_size[o], lastbits, S, P.anc, L, r, stop);
                        }
                        acc.cnt_pending += r.ok;
                        if (acc.cnt_pending >= FLUSH_AT) flush_count(acc.cnt_pending, &Q.st->count, stop, cap);
                        S.nodes[L] += r.ok;
                        acc.scanned += r.size;
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
                                acc.fsp++;
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
                            acc.inter++;
                            const uint32_t sz = isect_write(op_list(op_off[p], adj, ws), op_size[p],
                                                            op_vtx[p] != SGM_INVALID ? vbits(op_vtx[p]) : nullptr,
                                                            adj + nbo, nbs, vbits(v), ws + top, acc.elems, acc.outs,
                                                            &S.probes, stop);
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
                                acc.fsp++;
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
        if (SPLIT && uni(S.sp_lo) != SGM_INVALID) {
            // a cut-off at depth 1 ends the loop with F_PRUNED (set only after the loop ran)
            const bool cut1 = use_fs && !aborted && (uni(S.flags[1]) & F_PRUNED);
            const uint32_t sit = uni(S.sp_it);
            // the chunk's record: lane f < 32 its depth-f nodes, lanes 32..36 the other counters
            unsigned long long* rec = Q.rec + sit;
            unsigned long long x = 0;
            if (lane < SGM_MAXQ) x = lane < n ? S.nodes[lane] : 0ull;
            else if (lane == 32) x = acc.inter;
            else if (lane == 33) x = acc.fsp;
            else if (lane == 34) x = acc.elems;
            else if (lane == 35) x = acc.outs;
            else if (lane == 36) x = acc.scanned;
            else if (lane == 37) x = S.probes;
            if (lane < SGM_SPLIT_FIELDS) rec[(size_t)lane * SGM_SPLIT_ITEMS] = x;
            acc.inter = acc.fsp = acc.elems = acc.outs = acc.scanned = 0;
            if (lane == 0) S.probes = 0;
            clear_nodes(S, nullptr, 0);
            if (lane == 0 && !aborted) {
                Q.split[2 + 2 * SGM_SPLIT_ROOTS + sit] = 1u;  // done
                if (cut1) atomicMin(S.sp_cut, S.sp_lo);
            }
        }
        acc.busy += wall_clock64() - t_mark;
        if (Q.budget && acc.busy >= SGM_CHARGE_TICKS) charge_time(Q.st, stop, Q.budget, acc.busy);
        flush_count(acc.cnt_pending, &Q.st->count, stop, cap);
        if (overflow) {
            // this query's result is void (its stop word is raised); the host re-runs it
            flush_stats(Q, S, acc, n, true);
            overflow = false;
            cur = SGM_INVALID;
            continue;
        }
        (void)aborted;  // an aborted item means its query (or the whole batch) was stopped
    }
    if (cur != SGM_INVALID) flush_stats(qs[cur], S, acc, n, false);
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
    unsigned long long probes;  // probe_words not yet published
    uint64_t rem[SGM_MAXQ];
    uint64_t cdm[SGM_MAXQ];     // GuP conflict detection: pruned lanes of the chunk not yet counted
    uint64_t fail[SGM_MAXQ];  // failing sets (use_fs)
    unsigned long long nodes[SGM_MAXQ];
    const uint32_t* lst[SGM_MAXQ];
    uint32_t emb[SGM_MAXQ];
    uint32_t iemb[SGM_MAXQ];
    uint32_t pos[SGM_MAXQ];
    uint32_t tot[SGM_MAXQ];
    uint32_t chunk_base[SGM_MAXQ];
    uint32_t stack_top[SGM_MAXQ + 1];
    // Per depth d >= 1, the rows of its backward neighbours at depths below d-1, intersected
    // once per visit of depth d-1 (fF[d], fFn[d]; null when d's only backward neighbour is
    // at d-1), and its edge from depth d-1 (fE[d]; SGM_INVALID when none, and then fF[d] is
    // the whole list).  Every sibling at d-1 then needs at most one intersection.
    const uint32_t* fF[SGM_MAXQ];
    uint32_t fFn[SGM_MAXQ], fE[SGM_MAXQ];
    uint32_t fFs[SGM_MAXQ];  // hub slot of fF[d] when it is one hub's row, else SGM_INVALID
    // the last level over a chunk of depth-(L-1) siblings (queries of 3+ vertices), when L
    // has no edge from L-1: fF[L]'s matches, size and conflict bits against depths < L-1
    uint32_t baseA, candsA;
    uint64_t cfA;
    uint32_t sok[64], scnt[64];  // per sibling lane: matches and list size (segmented sums)
    uint64_t scf[64];            // per sibling lane: conflict bits
};

// Build the candidate list of depth d into (*list, *size); all but the last intersection
// are materialised on the stack from *top; when `leave_last` is set the last backward
// neighbour's row is returned separately (for a fused count) in (*last, *nlast).
__device__ bool aux_list(const sgm_auxplan_dev& A, const WaveAux& S, uint32_t d, uint32_t* ws, uint32_t stack_words,
                         uint32_t& top, const uint32_t*& list, uint32_t& size, bool leave_last,
                         const uint32_t*& last, uint32_t& nlast, unsigned long long& elems, unsigned long long& outs,
                         unsigned long long* probes, const uint32_t* stop)
{
    const uint32_t f = A.prev_first[d], e_end = A.prev_first[d + 1];
    const uint32_t np = e_end - f;
    {
        const uint32_t* o = A.tab_off[f];
        const uint32_t v = S.emb[A.prev_depth[f]];  // rows are indexed by data vertex
        const uint32_t b = o[v];
        list = A.tab_idx[f] + b;
        size = o[v + 1] - b;
    }
    const uint32_t stop_at = leave_last ? e_end - 1 : e_end;
    for (uint32_t e = f + 1; e < stop_at; e++) {
        const uint32_t* o = A.tab_off[e];
        const uint32_t v = S.emb[A.prev_depth[e]];
        const uint32_t b = o[v], nb = o[v + 1] - b;
        const uint32_t need = size < nb ? size : nb;
        if (top + need > stack_words) return false;
        const uint32_t sz =
            isect_write(list, size, nullptr, A.tab_idx[e] + b, nb, nullptr, ws + top, elems, outs, probes, stop);
        list = ws + top;
        size = sz;
        top += sz;
    }
    if (leave_last && np > 1) {
        const uint32_t e = e_end - 1;
        const uint32_t* o = A.tab_off[e];
        const uint32_t v = S.emb[A.prev_depth[e]];
        const uint32_t b = o[v];
        last = A.tab_idx[e] + b;
        nlast = o[v + 1] - b;
    } else {
        last = nullptr;
        nlast = 0;
    }
    return true;
}

// fF / fFn / fE of depth dn (1 <= dn <= L), at a time when the depths below dn-1 are
// matched: their rows intersected on the stack from *top (a single row is aliased).
__device__ bool aux_prefix(const sgm_auxplan_dev& A, WaveAux& S, uint32_t dn, uint32_t* ws, uint32_t stack_words,
                           uint32_t& top, unsigned long long& elems, unsigned long long& outs,
                           unsigned long long* probes, const uint32_t* stop, const uint32_t* __restrict__ hub_slot,
                           const uint32_t* __restrict__ hub_bits, uint32_t hub_words)
{
    const uint32_t f0 = A.prev_first[dn], f1 = A.prev_first[dn + 1];
    const uint32_t* F = nullptr;
    uint32_t Fn = 0, E = SGM_INVALID, Fs = SGM_INVALID;
    for (uint32_t e = f0; e < f1; e++) {
        const uint32_t pd = A.prev_depth[e];
        if (pd + 1 == dn) {
            E = e;
            continue;
        }
        const uint32_t* o = A.tab_off[e];
        const uint32_t v = S.emb[pd];
        const uint32_t b = o[v], nb = o[v + 1] - b;
        const uint32_t rs = hub_slot ? uni(hub_slot[uni(S.emb[pd])]) : SGM_INVALID;
        if (!F) {
            F = A.tab_idx[e] + b;
            Fn = nb;
            Fs = rs;
            continue;
        }
        const uint32_t need = Fn < nb ? Fn : nb;
        if (top + need > stack_words) return false;
        const uint32_t sz = isect_pos(F, Fn, Fs, A.tab_idx[e] + b, nb, rs, A.cand[dn], hub_bits, hub_words, ws + top,
                                      elems, outs, probes, stop);
        F = ws + top;
        Fn = sz;
        Fs = SGM_INVALID;
        top += sz;
    }
    S.fF[dn] = F;
    S.fFn[dn] = Fn;
    S.fE[dn] = E;
    S.fFs[dn] = Fs;
    return true;
}

// The candidate list of depth d (1 <= d < L): fF[d] intersected with the row of d's edge
// from d-1 at that depth's match (the same set as the intersection of all its backward
// neighbours' rows in order, getNextCandidatesImpl inc/explore.h:284-354).
__device__ bool aux_list_at(const sgm_auxplan_dev& A, const WaveAux& S, uint32_t d, uint32_t* ws,
                            uint32_t stack_words, uint32_t& top, const uint32_t*& list, uint32_t& size,
                            unsigned long long& elems, unsigned long long& outs, unsigned long long* probes,
                            const uint32_t* stop, const uint32_t* __restrict__ hub_slot,
                            const uint32_t* __restrict__ hub_bits, uint32_t hub_words)
{
    const uint32_t E = uni(S.fE[d]);
    const uint32_t* F = S.fF[d];
    const uint32_t Fn = uni(S.fFn[d]);
    if (E == SGM_INVALID) {
        list = F;
        size = Fn;
        return true;
    }
    const uint32_t* o = A.tab_off[E];
    const uint32_t v = S.emb[d - 1];
    const uint32_t b = o[v], nb = o[v + 1] - b;
    if (!F) {
        list = A.tab_idx[E] + b;
        size = nb;
        return true;
    }
    const uint32_t need = Fn < nb ? Fn : nb;
    if (top + need > stack_words) return false;
    const uint32_t rs = hub_slot ? uni(hub_slot[uni(S.emb[d - 1])]) : SGM_INVALID;
    size = isect_pos(F, Fn, uni(S.fFs[d]), A.tab_idx[E] + b, nb, rs, A.cand[d], hub_bits, hub_words, ws + top, elems,
                     outs, probes, stop);
    list = ws + top;
    top += size;
    return true;
}

// count positions whose vertex is not used by depths < nd; conflicts add anc[L] | anc[j]
// (j = the depth holding the vertex) to *cf
__device__ __forceinline__ uint32_t aux_count_chunk(bool valid, uint32_t vidx, const uint32_t* __restrict__ cand,
                                                    const WaveAux& S, uint32_t L, const uint64_t* __restrict__ anc,
                                                    uint64_t* cf, uint32_t nd)
{
    bool ok = false;
    uint32_t c = SGM_INVALID;
    if (valid) {
        const uint32_t v = cand[vidx];
        for (uint32_t j = 0; j < nd; j++)
            if (S.emb[j] == v) c = j;
        ok = c == SGM_INVALID;
    }
    const uint64_t cb = __ballot(valid && !ok);
    if (cf && cb) *cf |= lanes_or64(cb, (valid && !ok) ? (anc[L] | anc[c]) : 0ull);
    return (uint32_t)__popcll(__ballot(ok));
}

// The last level for every sibling of one chunk at depth L-1 at once (the siblings in
// S.rem[L-1], whose list starts at chunk_base): for sibling s the candidates are lastF
// intersected with s's row of the table from depth L-1 (lastE), or lastF alone when the
// last vertex has no backward neighbour at L-1.  The per-sibling results are then applied in
// sibling order exactly as one-at-a-time evaluation would (PruneCheck after each: the first
// sibling whose failing set lacks L-1 cuts the rest of the loop), and S.rem[L-1] is
// consumed.  Counters: nodes[L-1] and nodes[L], one intersection per extra backward
// neighbour of the last vertex for each evaluated sibling, as the reference counts them.
__device__ void aux_last_chunk(const sgm_auxplan_dev& A, WaveAux& S, uint32_t L, uint32_t base, int use_fs,
                               const uint64_t* __restrict__ anc, unsigned long long& inter,
                               unsigned long long& elems, unsigned long long& outs, unsigned long long& scanned,
                               unsigned long long& fsp, unsigned long long& cnt_pending,
                               const uint32_t* __restrict__ hub_slot, const uint32_t* __restrict__ hub_bits,
                               uint32_t hub_words)
{
    const uint32_t lane = lane_id(), d = L - 1;
    const uint64_t sib = uni64(S.rem[d]);
    S.rem[d] = 0;
    const bool me = (sib >> lane) & 1ull;
    const uint32_t idx = me ? S.lst[d][base + lane] : 0u;
    const uint32_t* F = S.fF[L];
    const uint32_t Fn = uni(S.fFn[L]), E = uni(S.fE[L]);
    const uint32_t* candL = A.cand[L];
    const uint32_t nsib = (uint32_t)__popcll(sib);
    uint32_t ok_s = 0, cands_s = 0;
    uint64_t cf_s = 0;
    if (E == SGM_INVALID) {
        // every sibling's list is F: its own vertex is the one extra vertex to exclude
        const uint32_t v = me ? A.cand[d][idx] : 0u;
        uint32_t lo = 0, hi = Fn;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (candL[F[mid]] < v) lo = mid + 1;
            else hi = mid;
        }
        const bool hit = me && lo < Fn && candL[F[lo]] == v;
        cands_s = uni(S.candsA);
        ok_s = uni(S.baseA) - (hit ? 1u : 0u);
        cf_s = uni64(S.cfA) | (hit ? (anc[L] | anc[d]) : 0ull);
        if (lane == 0) S.probes += (unsigned long long)nsib * (1u + probe_cost(Fn, false));
        elems += (unsigned long long)nsib * Fn;
    } else {
        // sibling s's row R from depth L-1, intersected with F when the last vertex has other
        // backward neighbours: the shorter of the two is walked and each key probed in the
        // other; the siblings' walks are concatenated, 64 keys per step
        const uint32_t* o = A.tab_off[E];
        const uint32_t* tidx = A.tab_idx[E];
        const uint32_t vs = me ? A.cand[d][idx] : 0u;  // the sibling's vertex: its row's index
        const uint32_t rb = me ? o[vs] : 0u, rlen = me ? o[vs + 1] - rb : 0u;
        const bool walk_f = F && Fn < rlen;
        // F's keys probe a sibling's row; when the sibling's vertex has an adjacency bitmap,
        // "x in the row" is "C(u_L)[x] adjacent to it": one bitmap word instead of a search
        const uint32_t hslot = (walk_f && hub_slot) ? hub_slot[vs] : SGM_INVALID;
        const uint32_t fslot = (F && hub_slot) ? uni(S.fFs[L]) : SGM_INVALID;  // F as one hub's row
        const uint32_t len = walk_f ? Fn : rlen;
        const uint32_t incl = wave_incl_scan(len), total = rdlane(incl, 63);
        S.sok[lane] = 0;
        S.scnt[lane] = 0;
        S.scf[lane] = 0;
        wave_sync();
        for (uint32_t b0 = 0; b0 < total; b0 += 64) {
            const TabStep st = tab_step(incl, len, total, b0, lane);
            const uint32_t own = st.own & 63u;
            const uint32_t rbj = (uint32_t)__shfl((int)rb, (int)own);
            const uint32_t rlj = (uint32_t)__shfl((int)rlen, (int)own);
            const bool wf = (uint32_t)__shfl((int)walk_f, (int)own) != 0;
            const uint32_t hs = (uint32_t)__shfl((int)hslot, (int)own);
            uint32_t x = 0;
            bool in = false;
            if (st.act) {
                if (wf) {
                    x = F[st.pos];
                    if (hs != SGM_INVALID) {
                        const uint32_t w = candL[x];
                        in = (hub_bits[(size_t)hs * hub_words + (w >> 5)] >> (w & 31u)) & 1u;
                    } else {
                        in = lower_found(tidx + rbj, rlj, x);
                    }
                } else {
                    x = tidx[rbj + st.pos];
                    if (!F) {
                        in = true;
                    } else if (fslot != SGM_INVALID) {
                        const uint32_t w = candL[x];
                        in = (hub_bits[(size_t)fslot * hub_words + (w >> 5)] >> (w & 31u)) & 1u;
                    } else {
                        in = lower_found(F, Fn, x);
                    }
                }
            }
            uint32_t c = SGM_INVALID;
            if (in) {
                const uint32_t v = candL[x];
                for (uint32_t j = 0; j < d; j++)
                    if (S.emb[j] == v) c = j;
            }
            const bool good = in && c == SGM_INVALID;
            const uint64_t bin = __ballot(in), bok = __ballot(good);
            const uint64_t upto = st.below | (1ull << lane);
            if (st.last) {
                S.scnt[st.own] += (uint32_t)__popcll(bin & upto);
                S.sok[st.own] += (uint32_t)__popcll(bok & upto);
            }
            if (use_fs && in && !good) atomicOr((unsigned long long*)&S.scf[st.own & 63u], anc[L] | anc[c]);
        }
        wave_sync();
        cands_s = S.scnt[lane];
        ok_s = S.sok[lane];
        cf_s = S.scf[lane];
        const uint32_t found = wave_sum32(me ? cands_s : 0u);
        if (F) {
            // keys of the shorter list, each with a bitmap word or a binary search in the longer
            const bool pbits = walk_f ? hslot != SGM_INVALID : fslot != SGM_INVALID;
            const uint32_t pw = wave_sum32(me ? len * (1u + probe_cost(walk_f ? rlen : Fn, pbits)) : 0u);
            if (lane == 0) S.probes += pw;
            elems += (unsigned long long)wave_sum32(me ? rlen : 0u) + (unsigned long long)nsib * Fn;
            outs += found;
            scanned += found;
        } else {
            scanned += total;
        }
    }
    // PruneCheck after each sibling, in order (inc/prune.h:119-137): fail = anc[L] for an
    // empty list, all ones after a match, else the conflicts' bits
    uint64_t counted = sib;
    uint64_t gprune = uni64(S.cdm[d]);  // the chunk's guard prunes, up to a cut
    S.cdm[d] = 0;
    if (use_fs) {
        const uint64_t fs = cands_s == 0 ? anc[L] : (ok_s ? ~0ull : cf_s);
        const uint64_t cut = __ballot(me && !((fs >> d) & 1ull));
        if (cut) {
            const uint32_t c = (uint32_t)__ffsll((long long)cut) - 1u;
            counted = sib & ((2ull << c) - 1ull);
            gprune &= (2ull << c) - 1ull;
            S.fail[d - 1] = rdlane64(fs, c);
            S.pos[d] = S.tot[d];
            fsp++;
        } else {
            S.fail[d - 1] |= wave_or64(me ? fs : 0ull);
        }
    }
    const bool cnt_me = (counted >> lane) & 1ull;
    const uint32_t nc = (uint32_t)__popcll(counted);
    const uint32_t okc = wave_sum32(cnt_me ? ok_s : 0u);
    S.nodes[d] += nc;
    S.nodes[L] += okc;
    fsp += (unsigned long long)__popcll(gprune);
    cnt_pending += okc;
    inter += (unsigned long long)nc * (A.prev_first[L + 1] - A.prev_first[L] - 1u);
    wave_sync();
}

// Many queries per launch: items are the depth-0 candidates of every query, query q owning
// item_first[q+1] - item_first[q] of them, handed out by its counter next[q] (zero at
// launch) in the QueryWalk order; plans, statistics and stop words are per query.
// Four waves per SIMD (at most 128 VGPRs, a few spilled): 90 ms per Chung-Lu 200-query
// launch against 98.6 ms at the unconstrained 157 VGPRs (three waves per SIMD).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_enum_aux(const sgm_auxplan_dev* __restrict__ plans,
                                                  const uint32_t* __restrict__ item_first, uint32_t nq,
                                                  uint32_t* __restrict__ stack, uint32_t stack_words,
                                                  uint32_t* __restrict__ next, sgm_stats_dev* __restrict__ stats,
                                                  uint32_t* __restrict__ stops, uint32_t* __restrict__ dev_stop,
                                                  const uint32_t* __restrict__ host_stop, unsigned long long cap,
                                                  unsigned long long budget, uint32_t root_lo, uint32_t root_hi,
                                                  int use_fs, const uint32_t* __restrict__ hub_slot,
                                                  const uint32_t* __restrict__ hub_bits, uint32_t hub_words)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* smem = reinterpret_cast<uint8_t*>(smem_u64);
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    WaveAux& S = *reinterpret_cast<WaveAux*>(smem + wave * ((sizeof(WaveAux) + 15) & ~(size_t)15));
    uint32_t* ws = stack + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * stack_words;
    for (uint32_t d = 0; d < SGM_MAXQ; d++) S.nodes[d] = 0;
    S.probes = 0;
    unsigned long long inter = 0, elems = 0, outs = 0, scanned = 0, cnt_pending = 0, items_done = 0, fsp = 0;
    unsigned long long busy = 0;  // wave time on the current query not yet added to its total
    bool overflow = false;
    uint32_t poll = 0;
    uint64_t relay_at = 0;
    uint32_t cur = SGM_INVALID, n = 0, L = 0;
    const sgm_auxplan_dev* Ap = plans;
    const uint64_t* anc = nullptr;
    sgm_stats_dev* gst = nullptr;
    uint32_t* stop = nullptr;
    // the current query's counters to its statistics (an overflow voids the query: the host
    // re-runs it with a larger stack)
    auto flush_all = [&](bool ovf) {
        if (cur == SGM_INVALID) return;
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (lane == 0) {
            if (inter) atomicAdd(&gst->intersections, inter);
            if (elems) atomicAdd(&gst->elems_read, elems);
            if (outs) atomicAdd(&gst->elems_out, outs);
            if (scanned) atomicAdd(&gst->elems_scanned, scanned);
            if (items_done) atomicAdd(&gst->items_done, items_done);
            if (fsp) atomicAdd(&gst->fs_prunes, fsp);
            if (busy) atomicAdd(&gst->busy_ticks, busy);
            if (S.probes) atomicAdd(&gst->probe_words, S.probes);
            S.probes = 0;
            if (ovf) { atomicAdd(&gst->overflow, 1ull); atomicExch(stop, 1u); }
            for (uint32_t dd = 0; dd < n; dd++)
                if (S.nodes[dd]) atomicAdd(&gst->nodes[dd], S.nodes[dd]);
        }
        inter = elems = outs = scanned = items_done = fsp = busy = 0;
        for (uint32_t dd = 0; dd < SGM_MAXQ; dd++) S.nodes[dd] = 0;
    };
    QueryWalk walk;
    qw_init(walk, nq, blockIdx.x * (blockDim.x >> 6) + wave);
    while (true) {
        uint32_t q = 0, it = 0;
        if (!qw_next(walk, item_first, next, stops, dev_stop, q, it)) break;
        relay_stop(wall_clock64(), relay_at, host_stop, dev_stop);
        if (q != cur) {
            flush_all(false);
            cur = q;
            Ap = plans + q;
            n = Ap->n;
            L = n - 1;
            anc = Ap->anc;
            gst = stats + q;
            stop = stops + q;
        }
        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) continue;  // capped (or voided) query
        const sgm_auxplan_dev& A = *Ap;
        const uint32_t root = A.root_order ? A.root_order[it] : it;  // roots in degree order
        const uint32_t v0 = A.cand[0][root];
        if (v0 < root_lo || v0 >= root_hi) continue;  // outside the launch's root range
        items_done++;
        uint64_t t_mark = wall_clock64();
        bool aborted = false;
        S.emb[0] = v0;
        S.iemb[0] = root;
        S.nodes[0] += 1;
        {
            uint32_t top0 = 0;  // depth 1's backward neighbours are all at depth 0: nothing written
            aux_prefix(A, S, 1, ws, stack_words, top0, elems, outs, &S.probes, stop, hub_slot, hub_bits, hub_words);
        }
        S.stack_top[1] = 0;
        uint32_t d = 1;
        // back from a finished child at depth d to d - 1, then PruneCheck(d - 1)
        // (inc/prune.h:119-137): a failing set without d - 1 cuts d - 1's remaining candidates
        auto back_up = [&]() -> bool {
            d--;
            if (d == 0) return false;
            if (use_fs) {
                const uint64_t f = S.fail[d];
                if (!((f >> d) & 1ull)) {
                    S.fail[d - 1] = f;
                    S.rem[d] = 0;
                    S.cdm[d] = 0;  // candidates after the cut are never tested
                    S.pos[d] = S.tot[d];
                    fsp++;
                } else {
                    S.fail[d - 1] |= f;
                }
            }
            return true;
        };
        // enter depth d (d <= L): build the list; at d == L count directly
        bool entering = true;
        while (true) {
            if ((++poll & 63u) == 0) {
                if ((poll & 1023u) == 0) {  // every 1024 steps: the clock (stop relay, time budget)
                    const uint64_t now = wall_clock64();
                    relay_stop(now, relay_at, host_stop, dev_stop);
                    busy += now - t_mark;
                    t_mark = now;
                    if (budget && busy >= SGM_CHARGE_TICKS) charge_time(gst, stop, budget, busy);
                }
                if (__atomic_load_n(stop, __ATOMIC_RELAXED) | __atomic_load_n(dev_stop, __ATOMIC_RELAXED)) {
                    aborted = true;
                    break;
                }
            }
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
                // the last level is reached one sibling at a time only for 2-vertex queries
                // (otherwise aux_last_chunk takes it a chunk of siblings at a time)
                const bool built = is_last ? aux_list(A, S, d, ws, stack_words, top, list, size, true, last, nlast,
                                                      elems, outs, &S.probes, stop)
                                           : aux_list_at(A, S, d, ws, stack_words, top, list, size, elems, outs,
                                                         &S.probes, stop, hub_slot, hub_bits, hub_words);
                if (!built) {
                    overflow = true; aborted = true; break;
                }
                if (is_last) {
                    const uint32_t* cand = A.cand[d];
                    uint32_t ok = 0, cands = size;
                    uint64_t cf = 0;
                    uint64_t* cfp = use_fs ? &cf : nullptr;
                    if (!last) {
                        scanned += size;
                        for (uint32_t base = 0; base < size; base += 64) {
                            const uint32_t i = base + lane;
                            const bool valid = i < size;
                            ok += aux_count_chunk(valid, valid ? list[i] : 0u, cand, S, L, anc, cfp, L);
                        }
                    } else {
                        const uint32_t* a = list; uint32_t na = size;
                        const uint32_t* b = last; uint32_t nb = nlast;
                        if (na > nb) { const uint32_t* t = a; a = b; b = t; uint32_t tn = na; na = nb; nb = tn; }
                        elems += na + nb;
                        if (nb && lane == 0) S.probes += (unsigned long long)na * (1u + probe_cost(nb, false));
                        cands = 0;
                        for (uint32_t base = 0; base < na; base += 64) {
                            const uint32_t i = base + lane;
                            const uint32_t key = i < na ? a[i] : 0u;
                            const bool hit = i < na && lower_found(b, nb, key);
                            const uint32_t h = (uint32_t)__popcll(__ballot(hit));
                            outs += h;
                            scanned += h;
                            cands += h;
                            ok += aux_count_chunk(hit, key, cand, S, L, anc, cfp, L);
                        }
                    }
                    cnt_pending += ok;
                    if (cnt_pending >= FLUSH_AT) flush_count(cnt_pending, &gst->count, stop, cap);
                    S.nodes[L] += ok;
                    // the last level's failing set: no candidates -> anc[L]; a match -> all;
                    // otherwise the conflicts' anc[L] | anc[j]
                    if (use_fs) S.fail[L - 1] = cands == 0 ? anc[L] : (ok ? ~0ull : cf);
                    if (!back_up()) break;
                    continue;
                }
                if (use_fs) {  // NoCandidatesPruneCheck at the child's entry
                    if (size == 0) {
                        S.fail[d - 1] = anc[d];
                        if (!back_up()) break;
                        continue;
                    }
                    S.fail[d - 1] = 0;
                }
                if (size) {
                    // depth d+1's rows from depths below d are the same for every sibling of
                    // this visit: intersect them once
                    if (!aux_prefix(A, S, d + 1, ws, stack_words, top, elems, outs, &S.probes, stop, hub_slot,
                                    hub_bits, hub_words)) {
                        overflow = true; aborted = true; break;
                    }
                }
                if (d + 1 == L && size) {
                    const uint32_t* F = S.fF[L];
                    const uint32_t Fn = uni(S.fFn[L]);
                    if (uni(S.fE[L]) == SGM_INVALID) {
                        // no edge from L-1: F is every sibling's list; its matches against the
                        // depths below L-1 once, the sibling's own vertex per chunk
                        const uint32_t* cand = A.cand[L];
                        uint32_t ok = 0;
                        uint64_t cf = 0;
                        if (Fn <= 64) {
                            for (uint32_t base = 0; base < Fn; base += 64) {
                                const uint32_t i = base + lane;
                                const bool valid = i < Fn;
                                ok += aux_count_chunk(valid, valid ? F[i] : 0u, cand, S, L, anc,
                                                      use_fs ? &cf : nullptr, L - 1);
                            }
                            scanned += Fn;
                        } else {
                            // a long F (a hub's row): lane j looks for depth j's vertex in F
                            // (F's positions ascend with their vertices) instead of F's every
                            // entry being compared with the matched vertices
                            const uint32_t nd = L - 1;
                            bool hit = false;
                            if (lane < nd) {
                                const uint32_t v = S.emb[lane];
                                uint32_t lo = 0, hi = Fn;
                                while (lo < hi) {
                                    const uint32_t mid = (lo + hi) >> 1;
                                    if (cand[F[mid]] < v) lo = mid + 1;
                                    else hi = mid;
                                }
                                hit = lo < Fn && cand[F[lo]] == v;
                            }
                            const uint64_t hb = __ballot(hit);
                            ok = Fn - (uint32_t)__popcll(hb);
                            if (use_fs && hb) cf = lanes_or64(hb, hit ? (anc[L] | anc[lane]) : 0ull);
                            if (lane == 0) S.probes += (unsigned long long)nd * (1u + 2u * probe_cost(Fn, false));
                        }
                        S.baseA = ok;
                        S.candsA = Fn;
                        S.cfA = cf;
                    }
                }
                S.stack_top[d + 1] = top;
                S.lst[d] = list;
                S.tot[d] = size;
                S.pos[d] = 0;
                S.rem[d] = 0;
                S.cdm[d] = 0;
                continue;
            }
            if (S.rem[d] == 0 && S.cdm[d]) {  // the chunk's remaining guard prunes
                fsp += (unsigned long long)__popcll(uni64(S.cdm[d]));
                S.cdm[d] = 0;
            }
            if (S.rem[d] == 0 && S.pos[d] >= S.tot[d]) {
                if (!back_up()) break;
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
                uint32_t c = SGM_INVALID;
                if (valid) {
                    const uint32_t v = A.cand[d][list[i]];
                    for (uint32_t j = 0; j < d; j++)
                        if (S.emb[j] == v) c = j;
                    ok = c == SGM_INVALID;
                }
                // injectivity conflicts (InjectiveConflict): fail[d-1] |= anc[d] | anc[c]; a
                // later cut-off in this loop overwrites fail[d-1], so the order does not matter
                const uint64_t cb = __ballot(valid && !ok);
                if (use_fs && cb) S.fail[d - 1] |= lanes_or64(cb, (valid && !ok) ? (anc[d] | anc[c]) : 0ull);
                if (const uint32_t* gd = A.guard[d]) {
                    // GUPConflictDetection::matchCheck (inc/prune.h:577-595): a candidate whose
                    // guard is wholly matched is pruned; with failing sets its fail[d] is all
                    // ones (a later cut-off overwrites fail[d-1], so the order does not matter)
                    bool pr = false;
                    if (ok) {
                        const uint32_t* g = gd + 4 * (size_t)list[i];
                        const uint32_t gs = g[0];
                        uint32_t k = 0;
                        for (; k < gs; k++) {
                            const uint32_t w = g[1 + k];
                            bool m = false;
                            for (uint32_t j = 0; j < d; j++) m |= S.emb[j] == w;
                            if (!m) break;
                        }
                        pr = gs && k == gs;
                    }
                    const uint64_t pb = __ballot(pr);
                    if (pb && use_fs) S.fail[d - 1] = ~0ull;
                    ok = ok && !pr;
                    S.cdm[d] = pb;  // counted as the loop reaches them
                }
                S.rem[d] = __ballot(ok);
                if (d + 1 == L && S.rem[d]) {
                    aux_last_chunk(A, S, L, base, use_fs, anc, inter, elems, outs, scanned, fsp, cnt_pending, hub_slot,
                                   hub_bits, hub_words);
                    // a chunk can add many embeddings: publish them and leave at once when the
                    // query's cap (or a stop) is reached, instead of at the next poll
                    if (cnt_pending >= FLUSH_AT) {
                        flush_count(cnt_pending, &gst->count, stop, cap);
                        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) {
                            aborted = true;
                            break;
                        }
                    }
                }
                continue;
            }
            const uint64_t rm = S.rem[d];
            const uint32_t j = (uint32_t)__ffsll((long long)rm) - 1u;
            S.rem[d] = rm & (rm - 1);
            if (S.cdm[d]) {  // guard prunes of the lanes before this sibling
                const uint64_t below = uni64(S.cdm[d]) & ((1ull << j) - 1ull);
                fsp += (unsigned long long)__popcll(below);
                S.cdm[d] &= ~below;
            }
            const uint32_t vidx = list[S.chunk_base[d] + j];
            S.emb[d] = A.cand[d][vidx];
            S.iemb[d] = vidx;
            S.nodes[d] += 1;
            d++;
            entering = true;
        }
        busy += wall_clock64() - t_mark;
        if (budget && busy >= SGM_CHARGE_TICKS) charge_time(gst, stop, budget, busy);
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (overflow) {
            flush_all(true);
            overflow = false;
            cur = SGM_INVALID;
        }
        (void)aborted;  // a stopped query: its remaining items are skipped above
    }
    flush_all(false);
}

// ================================================================== dynamic-order enumeration
// explore_Intersection_DynamicOrder_DAF / _VEQ (inc/explore.h:1674-2162): the query vertex
// matched next is chosen per search node.  A vertex becomes extendable once every previous
// neighbour of it in the start order is matched (updataExtendableV :1726-1770); its
// candidates are then the table row of its first previous neighbour intersected with the
// rows of the others (getCandidatesFor :1691-1724), and it enters a priority queue ranked by
// the sum of its candidates' DAF weights (or by their count), vertices of query degree 1
// last.  The queue of depth d + 1 starts as a copy of depth d's (vec_rank_queue) and its top
// is matched at d + 1.  Ties are left to the queue, so the heap moves are libstdc++'s
// push_heap / pop_heap, replayed by every lane of the wave on its LDS copy.  One wave per
// depth-0 candidate.  A candidate list of one previous neighbour is its table row; the
// intersections of the others go to the wave's HBM stack, above the mark of the depth whose
// match made the vertex extendable.  Injectivity (buffer->vis) is the only check; the last
// level is counted in one pass.
#define SGM_DYN_HEAP 528  // heap of depth d (1 <= d < n) holds at most n - d entries; n - d + 1 reserved

struct WaveDyn {
    unsigned long long probes;
    unsigned long long nodes[SGM_MAXQ];
    uint64_t rem[SGM_MAXQ];           // per depth: the untried extensions of the current chunk
    uint64_t heap[SGM_DYN_HEAP];      // per depth: (vertex << 40) | (degree-1 flag << 32) | weight
    const uint32_t* lst[SGM_MAXQ];    // per query vertex: its candidate list (positions in C(u))
    uint32_t tot[SGM_MAXQ];           // per query vertex: the list's length
    uint32_t iemb[SGM_MAXQ];          // per query vertex: the position of its match in C(u)
    uint32_t ext[SGM_MAXQ];           // per query vertex: previous neighbours not yet matched
    uint32_t emb[SGM_MAXQ];           // per depth: the matched data vertex
    uint32_t mo[SGM_MAXQ];            // per depth: the matched query vertex (match_order)
    uint32_t pos[SGM_MAXQ];
    uint32_t chunk_base[SGM_MAXQ];
    uint32_t stack_top[SGM_MAXQ + 1];
    uint32_t hlen[SGM_MAXQ + 1];
};

// Rank key of a queue entry: larger = lower priority (extendable_vertex_compare,
// inc/explore.h:1675-1686: degree 1 below the rest, then the larger weight below)
__device__ __forceinline__ uint64_t dyn_key(uint64_t e) { return e & 0x1FFFFFFFFull; }

// first word of the heap of depth d (1 <= d < n): depth k reserves n - k + 1 entries
__device__ __forceinline__ uint32_t dyn_heap_off(uint32_t d, uint32_t n)
{
    return (d - 1) * (n + 1) - (d - 1) * d / 2;
}

// std::__push_heap: the value rises from `hole` past every parent that ranks below it
__device__ void dyn_sift_up(uint64_t* f, int hole, uint64_t val)
{
    int parent = (hole - 1) / 2;
    while (hole > 0 && dyn_key(f[parent]) > dyn_key(val)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = val;
}

// std::pop_heap over f[0, len): the top's hole sinks to a leaf through the child that ranks
// higher (the right one unless the left ranks strictly higher), the last entry takes the
// freed slot's place by a sift-up from that leaf
__device__ void dyn_pop(uint64_t* f, int len)
{
    if (len <= 1) return;
    const int m = len - 1;
    const uint64_t val = f[m];
    f[m] = f[0];
    int hole = 0, second = 0;
    while (second < (m - 1) / 2) {
        second = 2 * (second + 1);
        if (dyn_key(f[second]) > dyn_key(f[second - 1])) second--;
        f[hole] = f[second];
        hole = second;
    }
    if ((m & 1) == 0 && second == (m - 2) / 2) {
        second = 2 * (second + 1);
        f[hole] = f[second - 1];
        hole = second - 1;
    }
    dyn_sift_up(f, hole, val);
}

// Many queries per launch as in k_enum_aux: items are the depth-0 candidates of every query.
__global__ void __launch_bounds__(256) k_enum_dyn(const sgm_dynplan_dev* __restrict__ plans,
                                                  const uint32_t* __restrict__ item_first, uint32_t nq,
                                                  uint32_t* __restrict__ stack, uint32_t stack_words,
                                                  uint32_t* __restrict__ next, sgm_stats_dev* __restrict__ stats,
                                                  uint32_t* __restrict__ stops, uint32_t* __restrict__ dev_stop,
                                                  const uint32_t* __restrict__ host_stop, unsigned long long cap,
                                                  unsigned long long budget, uint32_t root_lo, uint32_t root_hi,
                                                  const uint32_t* __restrict__ hub_slot,
                                                  const uint32_t* __restrict__ hub_bits, uint32_t hub_words)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* smem = reinterpret_cast<uint8_t*>(smem_u64);
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    WaveDyn& S = *reinterpret_cast<WaveDyn*>(smem + wave * ((sizeof(WaveDyn) + 15) & ~(size_t)15));
    uint32_t* ws = stack + (size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * stack_words;
    if (lane < SGM_MAXQ) S.nodes[lane] = 0;
    S.probes = 0;
    wave_sync();
    unsigned long long inter = 0, elems = 0, outs = 0, scanned = 0, cnt_pending = 0, items_done = 0, busy = 0;
    bool overflow = false;
    uint32_t poll = 0;
    uint64_t relay_at = 0;
    uint32_t cur = SGM_INVALID, n = 0, L = 0;
    const sgm_dynplan_dev* Pp = plans;
    sgm_stats_dev* gst = nullptr;
    uint32_t* stop = nullptr;
    auto flush_all = [&](bool ovf) {
        if (cur == SGM_INVALID) return;
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (lane == 0) {
            if (inter) atomicAdd(&gst->intersections, inter);
            if (elems) atomicAdd(&gst->elems_read, elems);
            if (outs) atomicAdd(&gst->elems_out, outs);
            if (scanned) atomicAdd(&gst->elems_scanned, scanned);
            if (items_done) atomicAdd(&gst->items_done, items_done);
            if (busy) atomicAdd(&gst->busy_ticks, busy);
            if (S.probes) atomicAdd(&gst->probe_words, S.probes);
            if (ovf) { atomicAdd(&gst->overflow, 1ull); atomicExch(stop, 1u); }
        }
        if (lane < SGM_MAXQ) {
            if (lane < n && S.nodes[lane]) atomicAdd(&gst->nodes[lane], S.nodes[lane]);
            S.nodes[lane] = 0;
        }
        S.probes = 0;
        wave_sync();
        inter = elems = outs = scanned = items_done = busy = 0;
    };
    QueryWalk walk;
    qw_init(walk, nq, blockIdx.x * (blockDim.x >> 6) + wave);
    while (true) {
        uint32_t q = 0, it = 0;
        if (!qw_next(walk, item_first, next, stops, dev_stop, q, it)) break;
        relay_stop(wall_clock64(), relay_at, host_stop, dev_stop);
        if (q != cur) {
            flush_all(false);
            cur = q;
            Pp = plans + q;
            n = Pp->n;
            L = n - 1;
            gst = stats + q;
            stop = stops + q;
        }
        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) continue;
        const sgm_dynplan_dev& P = *Pp;
        const uint32_t u0 = P.root;
        const uint32_t root = P.root_order ? P.root_order[it] : it;  // roots in degree order
        const uint32_t v0 = P.cand[u0][root];
        if (v0 < root_lo || v0 >= root_hi) continue;
        items_done++;
        uint64_t t_mark = wall_clock64();
        bool aborted = false;

        // Match the query vertex of depth d to its chosen candidate's extensions: copy the
        // queue of depth d into depth d + 1, make the following neighbours whose last previous
        // neighbour this was extendable (their candidate lists, weights and queue entries, in
        // following-neighbour order), then pop the vertex of depth d + 1.  False: the stack
        // cannot hold an intersection.
        auto expand = [&](uint32_t d) -> bool {
            const uint32_t u = S.mo[d];
            uint64_t* dst = S.heap + dyn_heap_off(d + 1, n);
            uint32_t len = 0;
            if (d) {
                len = S.hlen[d];
                if (lane < len) dst[lane] = S.heap[dyn_heap_off(d, n) + lane];
                wave_sync();
            }
            uint32_t top = S.stack_top[d];
            for (uint32_t k = P.fol_first[u]; k < P.fol_first[u + 1]; k++) {
                const uint32_t f = P.fol[k];
                const uint32_t left = S.ext[f] - 1u;
                S.ext[f] = left;
                if (left) continue;
                const uint32_t e0 = P.prev_first[f], e1 = P.prev_first[f + 1];
                const uint32_t* list;
                uint32_t size;
                // a row of a hub data vertex is probed through its adjacency bitmap
                auto row_slot = [&](uint32_t v) -> uint32_t { return hub_slot ? uni(hub_slot[v]) : SGM_INVALID; };
                uint32_t lslot;
                {
                    const uint32_t p = P.prev_vtx[e0];
                    const uint32_t vp = uni(P.cand[p][S.iemb[p]]);  // p's match: its row's index
                    const uint32_t* o = P.tab_off[e0];
                    const uint32_t b = o[vp];
                    list = P.tab_idx[e0] + b;
                    size = o[vp + 1] - b;
                    lslot = e1 - e0 > 1 ? row_slot(vp) : SGM_INVALID;
                }
                for (uint32_t e = e0 + 1; e < e1; e++) {
                    const uint32_t p = P.prev_vtx[e];
                    const uint32_t vp = uni(P.cand[p][S.iemb[p]]);
                    const uint32_t* o = P.tab_off[e];
                    const uint32_t b = o[vp], nb = o[vp + 1] - b;
                    const uint32_t need = size < nb ? size : nb;
                    if (top + need > stack_words) return false;
                    const uint32_t sz = isect_pos(list, size, lslot, P.tab_idx[e] + b, nb, row_slot(vp), P.cand[f],
                                                  hub_bits, hub_words, ws + top, elems, outs, &S.probes, stop);
                    list = ws + top;
                    size = sz;
                    lslot = SGM_INVALID;
                    top += sz;
                }
                inter += e1 - e0 - 1;
                S.lst[f] = list;
                S.tot[f] = size;
                uint32_t w = size;
                if (const uint32_t* wf = P.weight[f]) {
                    uint32_t s = 0;
                    for (uint32_t i = lane; i < size; i += 64) s += wf[list[i]];
                    w = wave_sum32(s);  // u32 sum, wrapping as the reference's
                    if (lane == 0) S.probes += 2ull * size;  // the list's positions and their weights
                }
                const uint64_t val = ((uint64_t)f << 40) | ((uint64_t)((P.deg1 >> f) & 1u) << 32) | w;
                dst[len] = val;
                dyn_sift_up(dst, (int)len, val);
                len++;
                wave_sync();
            }
            const uint64_t t = dst[0];
            dyn_pop(dst, (int)len);
            S.hlen[d + 1] = len - 1;
            S.mo[d + 1] = (uint32_t)(t >> 40);
            S.stack_top[d + 1] = top;
            wave_sync();
            return true;
        };

        if (lane < n) S.ext[lane] = P.prev_first[lane + 1] - P.prev_first[lane];
        S.emb[0] = v0;
        S.iemb[u0] = root;
        S.mo[0] = u0;
        S.nodes[0] += 1;
        S.stack_top[0] = 0;
        wave_sync();
        if (!expand(0)) {
            overflow = true;
            aborted = true;
        }
        uint32_t d = 1;
        bool entering = true;
        // back from depth d to d - 1: the match at d - 1 is undone (restoreExtendableV)
        auto back_up = [&]() -> bool {
            d--;
            if (d == 0) return false;
            const uint32_t u = S.mo[d];
            for (uint32_t k = P.fol_first[u]; k < P.fol_first[u + 1]; k++) S.ext[P.fol[k]] += 1u;
            wave_sync();
            return true;
        };
        while (!aborted) {
            if ((++poll & 63u) == 0) {
                if ((poll & 1023u) == 0) {
                    const uint64_t now = wall_clock64();
                    relay_stop(now, relay_at, host_stop, dev_stop);
                    busy += now - t_mark;
                    t_mark = now;
                    if (budget && busy >= SGM_CHARGE_TICKS) charge_time(gst, stop, budget, busy);
                }
                if (__atomic_load_n(stop, __ATOMIC_RELAXED) | __atomic_load_n(dev_stop, __ATOMIC_RELAXED)) {
                    aborted = true;
                    break;
                }
            }
            const uint32_t u = uni(S.mo[d]);
            if (entering) {
                entering = false;
                const uint32_t size = uni(S.tot[u]);
                if (d == L) {
                    const uint32_t* list = S.lst[u];
                    const uint32_t* cand = P.cand[u];
                    uint32_t ok = 0;
                    scanned += size;
                    for (uint32_t base = 0; base < size; base += 64) {
                        const uint32_t i = base + lane;
                        bool good = false;
                        if (i < size) {
                            const uint32_t v = cand[list[i]];
                            good = true;
                            for (uint32_t j = 0; j < L; j++) good &= S.emb[j] != v;
                        }
                        ok += (uint32_t)__popcll(__ballot(good));
                    }
                    cnt_pending += ok;
                    S.nodes[L] += ok;
                    if (cnt_pending >= FLUSH_AT) {
                        flush_count(cnt_pending, &gst->count, stop, cap);
                        if (__atomic_load_n(stop, __ATOMIC_RELAXED)) {
                            aborted = true;
                            break;
                        }
                    }
                    if (!back_up()) break;
                    continue;
                }
                if (size == 0) {
                    if (!back_up()) break;
                    continue;
                }
                S.pos[d] = 0;
                S.rem[d] = 0;
                continue;
            }
            const uint32_t* list = S.lst[u];
            if (S.rem[d] == 0) {
                const uint32_t tot = uni(S.tot[u]);
                const uint32_t base = uni(S.pos[d]);
                if (base >= tot) {
                    if (!back_up()) break;
                    continue;
                }
                S.pos[d] = base + 64;
                S.chunk_base[d] = base;
                const uint32_t i = base + lane;
                scanned += (tot - base < 64u) ? (tot - base) : 64u;
                bool good = false;
                if (i < tot) {
                    const uint32_t v = P.cand[u][list[i]];
                    good = true;
                    for (uint32_t j = 0; j < d; j++) good &= S.emb[j] != v;
                }
                S.rem[d] = __ballot(good);
                continue;
            }
            const uint64_t rm = uni64(S.rem[d]);
            const uint32_t j = (uint32_t)__ffsll((long long)rm) - 1u;
            S.rem[d] = rm & (rm - 1);
            const uint32_t vidx = list[uni(S.chunk_base[d]) + j];
            S.emb[d] = P.cand[u][vidx];
            S.iemb[u] = vidx;
            S.nodes[d] += 1;
            wave_sync();
            if (!expand(d)) {
                overflow = true;
                aborted = true;
                break;
            }
            d++;
            entering = true;
        }
        busy += wall_clock64() - t_mark;
        if (budget && busy >= SGM_CHARGE_TICKS) charge_time(gst, stop, budget, busy);
        flush_count(cnt_pending, &gst->count, stop, cap);
        if (overflow) {
            flush_all(true);
            overflow = false;
            cur = SGM_INVALID;
        }
    }
    flush_all(false);
}

}  // namespace sgm

// ================================================================== launch wrappers
namespace sgm {

size_t enum_smem_bytes(uint32_t max_ops) { return 4 * enum_wave_bytes(max_ops); }

void launch_last_bits(hipStream_t s, const sgm_equery* qs, uint32_t nq, uint32_t V)
{
    if (!nq || !V) return;
    const uint32_t groups = (V + 63) / 64;
    dim3 grid((groups + 3) / 4 < 1024 ? (groups + 3) / 4 : 1024, nq);
    hipLaunchKernelGGL(k_last_bits, grid, dim3(256), 0, s, qs, V);
}

void launch_split_roots(hipStream_t s, const uint32_t* off, const uint32_t* items, const sgm_equery* qs, uint32_t nq,
                        uint32_t V, uint32_t max_chunks)
{
    if (max_chunks == 0 || max_chunks > SGM_SPLIT_CHUNKS) max_chunks = SGM_SPLIT_CHUNKS;
    if (nq) hipLaunchKernelGGL(k_split_roots, dim3(nq), dim3(64), 0, s, off, items, qs, V, max_chunks);
}

void launch_split_reduce(hipStream_t s, const sgm_equery* qs, uint32_t nq)
{
    if (nq) hipLaunchKernelGGL(k_split_reduce, dim3(SGM_SPLIT_ITEMS / 256, nq), dim3(256), 0, s, qs);
}

void launch_enum_cache(hipStream_t s, uint32_t blocks, const uint32_t* off, const uint32_t* adj, const uint32_t* items,
                       const sgm_equery* qs, uint32_t nq, uint32_t max_ops, uint32_t* stack, uint32_t stack_words,
                       uint32_t* dev_stop, const uint32_t* host_stop, int use_fs, const uint32_t* hub_slot,
                       const uint32_t* hub_bits, uint32_t hub_words, int serial_last, bool split)
{
    if (split)
        hipLaunchKernelGGL(k_enum_cache<true>, dim3(blocks), dim3(256), 4 * enum_wave_bytes(max_ops), s, off, adj, items,
                           qs, nq, max_ops, stack, stack_words, dev_stop, host_stop, use_fs, hub_slot, hub_bits,
                           hub_words, serial_last);
    else
        hipLaunchKernelGGL(k_enum_cache<false>, dim3(blocks), dim3(256), 4 * enum_wave_bytes(max_ops), s, off, adj,
                           items, qs, nq, max_ops, stack, stack_words, dev_stop, host_stop, use_fs, hub_slot, hub_bits,
                           hub_words, serial_last);
}

void launch_enum_aux(hipStream_t s, uint32_t blocks, const sgm_auxplan_dev* plans, const uint32_t* item_first,
                     uint32_t nq, uint32_t* stack, uint32_t stack_words, uint32_t* next, sgm_stats_dev* st,
                     uint32_t* stops, uint32_t* dev_stop, const uint32_t* host_stop, unsigned long long cap,
                     unsigned long long budget, uint32_t root_lo, uint32_t root_hi, int use_fs,
                     const uint32_t* hub_slot, const uint32_t* hub_bits, uint32_t hub_words)
{
    const size_t smem = 4 * ((sizeof(WaveAux) + 15) & ~(size_t)15);
    hipLaunchKernelGGL(k_enum_aux, dim3(blocks), dim3(256), smem, s, plans, item_first, nq, stack, stack_words, next,
                       st, stops, dev_stop, host_stop, cap, budget, root_lo, root_hi, use_fs, hub_slot, hub_bits,
                       hub_words);
}

void launch_enum_dyn(hipStream_t s, uint32_t blocks, const sgm_dynplan_dev* plans, const uint32_t* item_first,
                     uint32_t nq, uint32_t* stack, uint32_t stack_words, uint32_t* next, sgm_stats_dev* st,
                     uint32_t* stops, uint32_t* dev_stop, const uint32_t* host_stop, unsigned long long cap,
                     unsigned long long budget, uint32_t root_lo, uint32_t root_hi, const uint32_t* hub_slot,
                     const uint32_t* hub_bits, uint32_t hub_words)
{
    const size_t smem = 4 * ((sizeof(WaveDyn) + 15) & ~(size_t)15);
    hipLaunchKernelGGL(k_enum_dyn, dim3(blocks), dim3(256), smem, s, plans, item_first, nq, stack, stack_words, next,
                       st, stops, dev_stop, host_stop, cap, budget, root_lo, root_hi, hub_slot, hub_bits, hub_words);
}

}  // namespace sgm
