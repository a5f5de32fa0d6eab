// gfx950 kernels for the filter -> order -> enumerate path (wave64, CDNA4).
//
// All matching arithmetic is integer.  Every kernel keeps control flow wave-uniform
// where lanes cooperate (ballots, readlane), so per-wave state lives in LDS or in
// uniform registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sgm_internal.h"
#include "sgm_kernels.h"
#include "sgm_wave.h"

namespace sgm {

// ================================================================== NLF
// Unlabeled NLF (inc/filter.h:24-58): v may take u iff deg(v) >= deg_q(u).  One thread
// per data vertex writes its membership word for every query vertex at once.
__global__ void __launch_bounds__(256) k_nlf(const uint32_t* __restrict__ off, uint32_t V,
                                             const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0 || d->lab_on) return;  // unused batch slot / labelled query (k_nlf_lab)
    uint32_t* __restrict__ vm = d->vmask;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        uint32_t dv = off[v + 1] - off[v], m = 0;
        for (uint32_t u = 0; u < n; u++) m |= (dv >= d->qdeg[u] ? 1u : 0u) << u;
        vm[v] = m;
    }
}

// Labelled NLF (inc/filter.h:24-58): v may take u iff label(v) = label(u), deg(v) >=
// deg_q(u), and for every label among u's neighbours v has at least as many neighbours of
// that label (the NLF maps, inc/graph.h:76-91).  The data side holds each vertex's
// neighbour-label counts as runs ascending by label (run_off[V+1], run_lab, run_cnt), the
// query side u's (label, count) pairs ascending by label (sgm_qdesc::lreq); the count test
// is one merge of the two lists.  Returns the membership word of v.
__device__ __forceinline__ uint32_t nlf_lab_word(const sgm_qdesc* __restrict__ d, uint32_t n, uint32_t v,
                                                 uint32_t dv, uint32_t lv, const uint32_t* __restrict__ run_off,
                                                 const uint32_t* __restrict__ run_lab,
                                                 const uint32_t* __restrict__ run_cnt)
{
    const uint32_t* __restrict__ first = d->lreq;
    const uint32_t* __restrict__ plab = first + n + 1;
    const uint32_t* __restrict__ pcnt = plab + first[n];
    const uint32_t r0 = run_off[v], r1 = run_off[v + 1];
    uint32_t m = 0;
    for (uint32_t u = 0; u < n; u++) {
        if (d->qlab[u] != lv || dv < d->qdeg[u]) continue;
        bool ok = true;
        uint32_t j = r0;
        for (uint32_t p = first[u]; p < first[u + 1] && ok; p++) {
            const uint32_t l = plab[p];
            while (j < r1 && run_lab[j] < l) j++;
            ok = j < r1 && run_lab[j] == l && run_cnt[j] >= pcnt[p];
        }
        if (ok) m |= 1u << u;
    }
    return m;
}

__global__ void __launch_bounds__(256) k_nlf_lab(const uint32_t* __restrict__ off, const uint32_t* __restrict__ lab,
                                                 const uint32_t* __restrict__ run_off,
                                                 const uint32_t* __restrict__ run_lab,
                                                 const uint32_t* __restrict__ run_cnt, uint32_t V,
                                                 const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0 || !d->lab_on) return;
    uint32_t* __restrict__ vm = d->vmask;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
        vm[v] = nlf_lab_word(d, n, v, off[v + 1] - off[v], lab[v], run_off, run_lab, run_cnt);
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
// (inc/filter.h:273-290) for every query of a batch (blockIdx.y = query).  Only bit u
// changes during the phase and u is not its own neighbour, so the checks of one phase are
// independent (the reference's sequential order inside a phase does not matter).
//
// The filter works in degree-rank space: rank r is the r-th vertex of the degree-sorted
// list, the rank-space adjacency lists hold ranks, and the NLF candidates of u are exactly
// the ranks below P_u = ge_cnt[deg_q(u)].
//
// Exact shortcuts (the answers are the reference's, only fewer checks run):
//   * With only NLF bits, left vertex u' is adjacent to w iff deg(w) >= deg_q(u'): nested
//     neighbourhoods, so Hall's condition is "the j-th largest neighbour degree of v is at
//     least the j-th largest requirement" for every j -- eight byte comparisons per 64-bit
//     word of the degree table (SWAR, values <= 31).  The pre-pass evaluates this test once
//     per rank for every u (f0 bit u = fails).
//   * Removals only take edges away, so a rank failing the NLF test fails every later check
//     too: phase (1, u) removes exactly the ranks with f0 bit u, plus those a
//     matcher rejects among the ranks whose neighbours lost a bit of N_q(u) (touch bit u).
//     Every other candidate passes without any work.
//   * Round 2 rechecks v only if a neighbour lost a bit of N_q(u) after v's round-1 check
//     of u (dirty bit u, set by removals in (1, u') for u < u' and in (2, u') for u > u');
//     otherwise every input is unchanged and v passes again.
// Summary words (OR over groups of SGM_GQL_GROUP ranks) of the failures, touch and dirty
// bits let a phase stream a summary word per group and open only the groups that hold work.
// Matchers: lsm_lane (one lane, <= GQL_LANE_DEG neighbours: a greedy matching that
// saturates the left side proves true; a left vertex no right vertex covers, or fewer
// useful right vertices than left vertices, proves false when the whole list was seen),
// then lsm_wave for the undecided; matcher candidates are queued per wave and taken 64 at
// a time, one per lane.
#define GQL_LANE_DEG 128u
#define GQL_VPL 4u
#define GQL_QCAP 512u  // per-wave matcher queue: < 64 left over + 64 * GQL_VPL new per unit

__device__ __forceinline__ int lsm_lane(const uint32_t* __restrict__ adj, uint32_t o, uint32_t deg,
                                        const uint32_t* __restrict__ vm, uint32_t L, uint32_t k)
{
    // returns 1 = passes, 0 = fails, -1 = undecided
    const uint32_t scan = deg < GQL_LANE_DEG ? deg : GQL_LANE_DEG;
    uint32_t orm = 0, useful = 0, freeL = L, matched = 0;
    for (uint32_t j0 = 0; j0 < scan; j0 += 8) {
        uint32_t w[8], m[8];
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = (j0 + t < scan) ? adj[o + j0 + t] : 0u;
#pragma unroll
        for (int t = 0; t < 8; t++) m[t] = (j0 + t < scan) ? (vm[w[t]] & L) : 0u;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            orm |= m[t];
            useful += m[t] != 0u;
            const uint32_t f = m[t] & freeL;
            if (f) { freeL &= ~(f & (0u - f)); matched++; }
        }
        if (matched == k) return 1;
    }
    if (matched == k) return 1;
    if (scan < deg) return -1;  // a partial view proves only success
    if (orm != L || useful < k) return 0;
    return -1;
}

// Hall's test on the degree table: byte j of D (the (j+1)-th largest capped neighbour
// degree) >= byte j of R (the (j+1)-th largest requirement, 0 past the left side) for all
// eight bytes.  Bytes are <= 31, so (D | 0x80..) - R never borrows across bytes.
__device__ __forceinline__ bool hall_word(uint64_t D, uint64_t R)
{
    const uint64_t H = 0x8080808080808080ull;
    return (((D | H) - R) & H) == H;
}

// Removals of the lanes' ranks x with rm set: their neighbours get touch bits tm and dirty
// bits dm, and the summary words of their groups the same bits (each word read first: bits
// are only ever added during a filter, so a word that already holds them needs no atomic).
// The removed lanes' lists are concatenated (wave prefix sum of their lengths) and the wave
// walks the concatenation 64 neighbours at a time, so many short lists share a step and a
// long one is split over all lanes.  Called by the whole wave.
__device__ void gql_mark_nbrs(const sgm_qdesc* __restrict__ d, const uint32_t* __restrict__ roff,
                              const uint32_t* __restrict__ radj, bool rm, uint32_t x, uint32_t tm, uint32_t dm)
{
    const uint32_t o = rm ? roff[x] : 0u;
    const uint32_t dv = rm ? roff[x + 1] - o : 0u;
    const uint32_t incl = wave_incl_scan(dv);
    const uint32_t total = rdlane(incl, 63);
    const uint32_t lane = lane_id();
    for (uint32_t b = 0; b < total; b += 64) {
        const uint32_t t = b + lane;
        const uint32_t j = wave_seg_owner(incl, t);
        const uint32_t oj = (uint32_t)__shfl((int)o, (int)j);
        const uint32_t sj = (uint32_t)__shfl((int)(incl - dv), (int)j);
        if (t >= total) continue;
        const uint32_t w = radj[oj + (t - sj)], gw = w / SGM_GQL_GROUP, gw2 = gw / SGM_GQL_GROUP2;
        if (tm && (d->touch[w] & tm) != tm) {
            atomicOr(&d->touch[w], tm);
            if ((d->tsum[gw] & tm) != tm) atomicOr(&d->tsum[gw], tm);
            if ((d->tsum2[gw2] & tm) != tm) atomicOr(&d->tsum2[gw2], tm);
        }
        if (dm && (d->dirty[w] & dm) != dm) {
            atomicOr(&d->dirty[w], dm);
            if ((d->dsum[gw] & dm) != dm) atomicOr(&d->dsum[gw], dm);
            if ((d->dsum2[gw2] & dm) != dm) atomicOr(&d->dsum2[gw2], dm);
        }
    }
}

// Pre-pass: f0[x] = bits u (x < P_u, N_q(u) non-empty) whose NLF table test rank x fails, and
// fsum[g] = OR of f0 over group g.
__global__ void __launch_bounds__(256) k_gql_prepass(const uint64_t* __restrict__ dtab, uint32_t V,
                                                     const uint32_t* __restrict__ ge_cnt,
                                                     const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0) return;
    uint32_t* __restrict__ flags = d->removed;
    // the query's thresholds and requirement words, staged once per block
    __shared__ uint32_t s_P[SGM_MAXQ];
    __shared__ uint64_t s_req[SGM_MAXQ][4];
    if (threadIdx.x < n) {
        const uint32_t u = threadIdx.x;
        s_P[u] = d->nbr[u] ? ge_cnt[d->qdeg[u]] : 0u;  // 0: no test for an empty left side
#pragma unroll
        for (uint32_t p = 0; p < 4; p++) s_req[u][p] = d->req[u][p];
    }
    __syncthreads();
    uint32_t Pmax = 0, kmax = 0;
    for (uint32_t u = 0; u < n; u++) {
        Pmax = s_P[u] > Pmax ? s_P[u] : Pmax;
        const uint32_t k = __popc(d->nbr[u]);
        kmax = k > kmax ? k : kmax;
    }
    const uint32_t np = (kmax + 7) / 8;
    const uint32_t lane = lane_id();
    uint32_t tests = 0;
    // a block covers the ranks of one level-2 summary word per iteration (block-uniform loop)
    static_assert(SGM_GQL_GROUP * SGM_GQL_GROUP2 == 256, "one level-2 word per 256-thread block");
    __shared__ uint32_t s_or[4];
    const uint32_t span = (Pmax + 255) & ~255u;
    for (uint32_t x0 = blockIdx.x * 256u; x0 < span; x0 += gridDim.x * 256u) {
        const uint32_t x = x0 + threadIdx.x;
        const bool in = x < Pmax;
        uint64_t D[4];
#pragma unroll
        for (uint32_t p = 0; p < 4; p++) D[p] = (in && p < np) ? dtab[(size_t)p * V + x] : 0ull;
        uint32_t f = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (x >= s_P[u]) continue;
            bool ok = true;
#pragma unroll
            for (uint32_t p = 0; p < 4; p++) ok = ok && (p >= np || hall_word(D[p], s_req[u][p]));
            f |= ok ? 0u : (1u << u);
            tests++;
        }
        if (in) d->f0[x] = f;
        uint32_t g = f;
#pragma unroll
        for (int o = 1; o < (int)SGM_GQL_GROUP; o <<= 1) g |= (uint32_t)__shfl_xor((int)g, o);
        if ((lane % SGM_GQL_GROUP) == 0 && in) d->fsum[x / SGM_GQL_GROUP] = g;
        const uint32_t wv = wave_or32(f);
        if (lane == 0) s_or[threadIdx.x >> 6] = wv;
        __syncthreads();
        if (threadIdx.x == 0) d->fsum2[x0 / 256u] = s_or[0] | s_or[1] | s_or[2] | s_or[3];
        __syncthreads();
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) tests += (uint32_t)__shfl_xor((int)tests, o);
    if (lane == 0 && tests) atomicAdd(&flags[SGM_GQL_WORK], tests);
}

__global__ void __launch_bounds__(256) k_gql_phase(const uint32_t* __restrict__ roff,
                                                   const uint32_t* __restrict__ radj, uint32_t V,
                                                   const uint32_t* __restrict__ ge_cnt,
                                                   const sgm_qdesc* __restrict__ qs, uint32_t phase)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (phase >= 2 * n) return;
    const uint32_t r = phase / n, u = phase % n;
    const uint32_t L = d->nbr[u];
    if (L == 0) return;  // empty left side: the matching trivially saturates it
    uint32_t* __restrict__ flags = d->removed;
    if (r == 1 && !((flags[SGM_GQL_DIRTY_ANY] >> u) & 1u)) return;  // no input of any check changed
    const bool use_touch = r == 0 && ((flags[SGM_GQL_TOUCH_ANY] >> u) & 1u);
    // labelled: candidates below P_u need not be members, and every member is checked in
    // round 1 (k_nlf_rank_lab set its touch bit)
    const bool lab = d->lab_on != 0;
    const uint32_t low = L & ((1u << u) - 1u);
    const uint32_t high = u == 31u ? 0u : (L & (~0u << (u + 1u)));
    const uint32_t tm = r == 0 ? high : 0u;  // touch bits only matter to later round-1 phases
    const uint32_t dm = r == 0 ? low : high;
    const uint32_t k = __popc(L);
    const uint32_t P = ge_cnt[d->qdeg[u]];  // NLF candidates of u: ranks [0, P)
    const uint32_t ng = (P + SGM_GQL_GROUP - 1) / SGM_GQL_GROUP;
    const uint32_t* __restrict__ sumA = r == 0 ? d->fsum : d->dsum;
    const uint32_t* __restrict__ sumB = use_touch ? d->tsum : nullptr;
    const uint32_t* __restrict__ word = r == 0 ? d->touch : d->dirty;
    uint32_t* __restrict__ vm = d->vmr;
    const uint32_t* __restrict__ f0 = d->f0;
    const uint32_t lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t w0 = blockIdx.x * waves + (threadIdx.x >> 6);
    const uint32_t stride = gridDim.x * waves;
    uint32_t local_removed = 0, n_lane = 0, n_wave = 0;
    __shared__ uint32_t s_queue[4][GQL_QCAP];
    uint32_t* __restrict__ wq = s_queue[(threadIdx.x >> 6) & 3u];
    uint32_t qh = 0, qt = 0;  // wave-uniform head / tail (ring index = value mod GQL_QCAP)
    // clear bit u of the lanes' ranks with rm set, then mark their neighbours (whole wave)
    auto remove_now = [&](bool rm, uint32_t x) {
        if (rm) {
            atomicAnd(&vm[x], ~(1u << u));
            local_removed++;
        }
        if (__ballot(rm)) gql_mark_nbrs(d, roff, radj, rm, x, tm, dm);
    };
    auto drain = [&](uint32_t cnt) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const bool has = lane < cnt;
        const uint32_t xv = has ? wq[(qh + lane) & (GQL_QCAP - 1u)] : 0u;
        qh += cnt;
        int verdict = -1;
        if (has) {
            const uint32_t o = roff[xv], dv = roff[xv + 1] - o;
            verdict = lsm_lane(radj, o, dv, vm, L, k);
            n_lane++;
        }
        bool rm = has && verdict == 0;
        // the wave: exact matcher for the undecided lanes (only bit u changes in a phase and
        // no check reads it, so the removals can wait until all checks are done)
        uint64_t pend = __ballot(has && verdict < 0);
        while (pend) {
            const uint32_t s2 = (uint32_t)__ffsll((long long)pend) - 1u;
            pend &= pend - 1;
            const uint32_t xs = rdlane(xv, s2);
            const uint32_t os = roff[xs], ds = roff[xs + 1] - os;
            if (lane == 0) n_wave++;
            if (!lsm_wave(radj + os, ds, vm, L, k) && lane == s2) rm = true;
        }
        remove_now(rm, xv);
    };
    // 64 summary words per wave step; the groups whose word holds bit u are then opened
    // 64 / SGM_GQL_GROUP at a time, SGM_GQL_GROUP lanes per group (coalesced rank segments)
    const uint32_t sub = lane / SGM_GQL_GROUP;  // which of the selected groups
    // level-2 words first: lane k of a wave checks word w0 + k * stride (64 words per load,
    // spread over the grid); each word with bit u set opens its 64 level-1 words
    const uint32_t ng2 = (ng + SGM_GQL_GROUP2 - 1) / SGM_GQL_GROUP2;
    const uint32_t* __restrict__ sum2A = r == 0 ? d->fsum2 : d->dsum2;
    const uint32_t* __restrict__ sum2B = use_touch ? d->tsum2 : nullptr;
    for (uint32_t o2 = 0; w0 + o2 * 64u * stride < ng2; o2++) {
        const uint32_t i2 = w0 + (o2 * 64u + lane) * stride;
        const uint32_t sw2 = i2 < ng2 ? (sum2A[i2] | (sum2B ? sum2B[i2] : 0u)) : 0u;
        uint64_t h2 = __ballot((sw2 >> u) & 1u);
        while (h2) {
            const uint32_t k2 = (uint32_t)__ffsll((long long)h2) - 1u;
            h2 &= h2 - 1;
            const uint32_t base = rdlane(i2, k2) * SGM_GQL_GROUP2;
            const uint32_t gi = base + lane;
            const uint32_t sw = gi < ng ? (sumA[gi] | (sumB ? sumB[gi] : 0u)) : 0u;
            uint64_t hm = __ballot((sw >> u) & 1u);
            while (hm) {
                uint32_t mine = SGM_INVALID;
                for (uint32_t j = 0; j < 64 / SGM_GQL_GROUP && hm; j++) {
                    const uint32_t bit = (uint32_t)__ffsll((long long)hm) - 1u;
                    hm &= hm - 1;
                    if (sub == j) mine = base + bit;
                }
                const uint32_t x = mine != SGM_INVALID ? mine * SGM_GQL_GROUP + lane % SGM_GQL_GROUP : SGM_INVALID;
                const bool c = x < P;
                const uint32_t fw = (c && r == 0) ? f0[x] : 0u;
                const uint32_t tw = (c && (r == 1 || use_touch)) ? word[x] : 0u;
                const uint32_t vw = (c && (r == 1 || lab)) ? vm[x] : 0u;
                const bool mem = (r == 0 && !lab) || ((vw >> u) & 1u);  // x is in C(u) now
                const bool rm = mem && ((fw >> u) & 1u);  // fails the NLF table test: fails now too
                const bool need = mem && !rm && ((tw >> u) & 1u);
                remove_now(rm, x);
                const uint64_t bal = __ballot(need);
                if (need) wq[(qt + mbcnt64(bal)) & (GQL_QCAP - 1u)] = x;
                qt += (uint32_t)__popcll(bal);
                while (qt - qh >= 64) drain(64);
            }
        }
    }
    while (qt != qh) drain(qt - qh < 64 ? qt - qh : 64u);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        local_removed += (uint32_t)__shfl_xor((int)local_removed, o);
        n_lane += (uint32_t)__shfl_xor((int)n_lane, o);
        n_wave += (uint32_t)__shfl_xor((int)n_wave, o);
    }
    if (lane == 0) {
        if (local_removed) {
            atomicAdd(&flags[phase], local_removed);
            atomicAdd(&flags[SGM_GQL_WORK + 3], local_removed);
            if (tm) atomicOr(&flags[SGM_GQL_TOUCH_ANY], tm);
            if (dm) atomicOr(&flags[SGM_GQL_DIRTY_ANY], dm);
        }
        if (n_lane) atomicAdd(&flags[SGM_GQL_WORK + 1], n_lane);
        if (n_wave) atomicAdd(&flags[SGM_GQL_WORK + 2], n_wave);
    }
}

// NLF in rank space: degrees are non-increasing in rank, so rank x takes u iff
// x < ge_cnt[deg_q(u)].
__global__ void __launch_bounds__(256) k_nlf_rank(const uint32_t* __restrict__ ge_cnt, uint32_t V,
                                                  const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0 || d->lab_on) return;
    __shared__ uint32_t s_P[SGM_MAXQ];  // thresholds, staged once per block
    if (threadIdx.x < n) s_P[threadIdx.x] = ge_cnt[d->qdeg[threadIdx.x]];
    __syncthreads();
    uint32_t* __restrict__ vmr = d->vmr;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < V; x += gridDim.x * blockDim.x) {
        uint32_t m = 0;
        for (uint32_t u = 0; u < n; u++) m |= (x < s_P[u] ? 1u : 0u) << u;
        vmr[x] = m;
    }
}

// Labelled NLF in rank space, the start of a labelled GQL.  The degree-table shortcut (a
// candidate that no removal has touched passes its round-1 check unexamined) holds only
// for degree-nested sets, so every member gets a check: touch[x] = vmr[x], with the touch
// summaries and the any-touch flags to match (those words start clear).
__global__ void __launch_bounds__(256) k_nlf_rank_lab(const uint32_t* __restrict__ by_deg,
                                                      const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ lab,
                                                      const uint32_t* __restrict__ run_off,
                                                      const uint32_t* __restrict__ run_lab,
                                                      const uint32_t* __restrict__ run_cnt, uint32_t V,
                                                      const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0 || !d->lab_on) return;
    const uint32_t lane = lane_id();
    uint32_t any = 0;
    // block-uniform loop: every lane reaches the shuffles; rank blocks of 64 stay inside one
    // level-2 group (SGM_GQL_GROUP * SGM_GQL_GROUP2 = 256 ranks)
    for (uint32_t base = blockIdx.x * blockDim.x; base < V; base += gridDim.x * blockDim.x) {
        const uint32_t x = base + threadIdx.x;
        uint32_t m = 0;
        if (x < V) {
            const uint32_t v = by_deg[x];
            m = nlf_lab_word(d, n, v, off[v + 1] - off[v], lab[v], run_off, run_lab, run_cnt);
            d->vmr[x] = m;
            d->touch[x] = m;
        }
        uint32_t g1 = m;
#pragma unroll
        for (uint32_t s = 1; s < SGM_GQL_GROUP; s <<= 1) g1 |= (uint32_t)__shfl_xor((int)g1, (int)s);
        if (x < V && x % SGM_GQL_GROUP == 0) d->tsum[x / SGM_GQL_GROUP] = g1;
        const uint32_t w = wave_or32(m);
        if (lane == 0 && w) atomicOr(&d->tsum2[x / (SGM_GQL_GROUP * SGM_GQL_GROUP2)], w);
        any |= w;
    }
    if (lane == 0 && any) atomicOr(&d->removed[SGM_GQL_TOUCH_ANY], any);
}

// Back to vertex ids after the last phase: vmask[v] = vmr[rank[v]].
__global__ void __launch_bounds__(256) k_unrank(const uint32_t* __restrict__ rank, uint32_t V,
                                                const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    if (d->n == 0) return;
    const uint32_t* __restrict__ vmr = d->vmr;
    uint32_t* __restrict__ vm = d->vmask;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) vm[v] = vmr[rank[v]];
}

// Rank-space adjacency: list of rank x = the ranks of by_deg[x]'s neighbours (in the
// vertex-id order of the original list).  One wave per rank.
__global__ void __launch_bounds__(256) k_rank_adj(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                  const uint32_t* __restrict__ by_deg, const uint32_t* __restrict__ rank,
                                                  const uint32_t* __restrict__ roff, uint32_t V,
                                                  uint32_t* __restrict__ radj)
{
    const uint32_t lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t x = blockIdx.x * waves + (threadIdx.x >> 6); x < V; x += gridDim.x * waves) {
        const uint32_t v = by_deg[x];
        const uint32_t o = off[v], dv = off[v + 1] - o, ro = roff[x];
        for (uint32_t j = lane; j < dv; j += 64) radj[ro + j] = rank[adj[o + j]];
    }
}

// Degree table for the GQL shortcut, by rank: byte j of dtab[(j/8)*V + x] is the (j+1)-th
// largest min(deg(w), 31) over the neighbours w of rank x, 0 if x has at most j neighbours
// (j < 32).  Query degrees are at most 31, so the cap loses nothing.  One wave per rank: a
// 31-bin histogram by ballots, then lane j reads off its order statistic.
__global__ void __launch_bounds__(256) k_nbr_deg(const uint32_t* __restrict__ roff, const uint32_t* __restrict__ radj,
                                                 uint32_t V, uint8_t* __restrict__ dtab)
{
    const uint32_t lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t x = blockIdx.x * waves + (threadIdx.x >> 6); x < V; x += gridDim.x * waves) {
        const uint32_t o = roff[x], dv = roff[x + 1] - o;
        uint32_t hist = 0;  // lane b: neighbours with min(deg, 31) == b
        for (uint32_t base = 0; base < dv; base += 64) {
            uint32_t c = 0;  // neighbours have degree >= 1, so bin 0 stays empty
            if (base + lane < dv) {
                const uint32_t w = radj[o + base + lane];
                const uint32_t dw = roff[w + 1] - roff[w];
                c = dw < 31u ? dw : 31u;
            }
            for (uint32_t b = 1; b < 32; b++) {
                const uint32_t cnt = (uint32_t)__popcll(__ballot(c == b));
                if (lane == b) hist += cnt;
            }
        }
        uint32_t ge = 0, Dj = 0;
        for (uint32_t b = 31; b >= 1; b--) {
            ge += rdlane(hist, b);
            if (Dj == 0 && ge >= lane + 1) Dj = b;
        }
        if (lane < 32) dtab[((size_t)(lane >> 3) * V + x) * 8 + (lane & 7u)] = (uint8_t)Dj;
    }
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
    if (n == 0) return;
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

// Candidate statistics of a batch without lists: cnt[q*32 + u] = |C(u)| and
// degsum[q*32 + u] = sum of deg(v) over C(u) (the cost model's mean candidate degree).
// Both start zeroed.  Each lane accumulates all 32 query vertices in registers over its
// strided vertices (compile-time indices, no cross-lane work in the loop); waves reduce
// once, the block sums its waves in LDS and adds one atomic per query vertex (same-address
// atomics from every wave of every block would serialise in L2).
__global__ void __launch_bounds__(256) k_cand_stats(const uint32_t* __restrict__ off, uint32_t V,
                                                    const sgm_qdesc* __restrict__ qs, uint32_t* __restrict__ cnt,
                                                    unsigned long long* __restrict__ degsum)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t n = d->n;
    if (n == 0) return;
    const uint32_t* __restrict__ vm = d->vmask;
    const uint32_t lane = lane_id();
    uint32_t c[SGM_MAXQ];
    unsigned long long sm[SGM_MAXQ];
#pragma unroll
    for (int u = 0; u < SGM_MAXQ; u++) { c[u] = 0; sm[u] = 0; }
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const uint32_t m = vm[v];
        if (!m) continue;
        const unsigned long long dv = off[v + 1] - off[v];
#pragma unroll
        for (int u = 0; u < SGM_MAXQ; u++) {
            const uint32_t bit = (m >> u) & 1u;
            c[u] += bit;
            sm[u] += bit ? dv : 0ull;
        }
    }
    // wave sums, then block sums in LDS: one pair of atomics per block and query vertex
    // (all blocks of a query add into the same 32 counters)
    __shared__ uint32_t s_c[4][SGM_MAXQ];
    __shared__ unsigned long long s_s[4][SGM_MAXQ];
    const uint32_t wv = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < SGM_MAXQ; u++) {
        uint32_t x = c[u];
        long long y = (long long)sm[u];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            x += (uint32_t)__shfl_xor((int)x, o);
            y += __shfl_xor(y, o);
        }
        if (lane == 0) { s_c[wv][u] = x; s_s[wv][u] = (unsigned long long)y; }
    }
    __syncthreads();
    const uint32_t u = threadIdx.x;
    if (u < n) {
        uint32_t x = 0;
        unsigned long long y = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) { x += s_c[w][u]; y += s_s[w][u]; }
        const size_t base = (size_t)blockIdx.y * SGM_MAXQ;
        if (x) {
            atomicAdd(&cnt[base + u], x);
            atomicAdd(&degsum[base + u], y);
        }
    }
}

// Per-query housekeeping of a filter batch, one launch each instead of one host call per
// query: clear the flag words of every query (block = query) ...
__global__ void __launch_bounds__(64) k_clear_flags(const sgm_qdesc* __restrict__ qs)
{
    uint32_t* f = qs[blockIdx.x].removed;
    for (uint32_t i = threadIdx.x; i < SGM_GQL_FLAG_WORDS; i += blockDim.x) f[i] = 0;
}

// ... and gather the candidate counts into one array (query i at out[i * 32]) for a single
// copy back.
__global__ void __launch_bounds__(64) k_gather_counts(const sgm_cdesc* __restrict__ cs, uint32_t* __restrict__ out)
{
    const sgm_cdesc* c = cs + blockIdx.x;
    if (threadIdx.x < SGM_MAXQ) out[(size_t)blockIdx.x * SGM_MAXQ + threadIdx.x] = threadIdx.x < c->n ? c->cnt[threadIdx.x] : 0u;
}

// Strided word fill: rep blockIdx.y clears p[rep * stride ... + len) (per-query scratch).
__global__ void __launch_bounds__(256) k_fill_strided(uint32_t* __restrict__ p, size_t len, size_t stride,
                                                      uint32_t val)
{
    uint32_t* q = p + (size_t)blockIdx.y * stride;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) q[i] = val;
}

// Word fill with vector stores (graph-capturable, unlike a memset node).
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t val)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = val;
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
// Table of directed query edge (a,b) (inc/table.h:7-137): for the i-th candidate v of a, the
// ascending positions k in C(b) of v's data neighbours that are candidates of b.  That row
// depends on (b, v) only, not on a, so a build makes each query's rows once per (b, v) in a
// row pool -- for the v in U_b, the union of C(a) over b's query neighbours a -- and gives
// edge (a,b) either the pool of b itself (when its rows are all in C(a)) or a copy of the
// pool rows of C(a).  Passes: row lengths per (b, v) in one walk over N(v) for all b at once,
// a scan per b into pool offsets, the row counts and offsets of every edge, then the pool
// rows and the copies.  Positions are ranks in b's bitmap, so rows come out sorted because
// adjacency lists are.
__device__ __forceinline__ uint32_t cand_rank(const glb<uint64_t>* __restrict__ bm,
                                              const glb<uint32_t>* __restrict__ rank, uint32_t w)
{
    const uint32_t word = w >> 6, bit = w & 63u;
    return rank[word] + (uint32_t)__popcll(bm[word] & ((1ull << bit) - 1ull));  // both loads in flight together
}

// U_v: the query vertices b whose pool holds a row of v (v is a candidate of a neighbour of b)
__device__ __forceinline__ uint32_t pool_umask(const sgm_pooldesc& D, uint32_t vm)
{
    uint32_t u = 0;
    for (uint32_t m = vm; m; m &= m - 1) u |= D.nbr[__ffs(m) - 1];
    return u;
}

// A wave takes 64 consecutive data vertices (one per lane) and walks their concatenated
// neighbour lists 64 entries at a time (rows average a few dozen entries, so one wave per
// vertex would leave most lanes idle behind a chain of dependent loads).  Per (vertex, b)
// state lives in LDS: slot (lane, b) at tl[(wave * 64 + lane) * kmax + b].  In each step the
// entries of one vertex are a contiguous lane range; its last lane carries the vertex's hits
// of the step into the slot.
// Row lengths: poff[b * (V + 1) + v] = |N(v) n C(b)| when b is in U_v, else 0 (blockIdx.y =
// the query).
__global__ void __launch_bounds__(256) k_pool_count(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                    uint32_t V, const sgm_pooldesc* __restrict__ ds, uint32_t kmax)
{
    extern __shared__ uint32_t tl[];
    const sgm_pooldesc& D = ds[blockIdx.y];
    const glb<uint32_t>* __restrict__ vmask = gcast(D.vmask);
    glb<uint32_t>* __restrict__ poff = gcast(D.poff);
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = D.n;
    uint32_t* mine = tl + (wave * 64 + lane) * kmax;
    uint32_t* slots = tl + wave * 64 * kmax;
    for (uint32_t g = blockIdx.x * waves + wave; (uint64_t)g * 64 < V; g += gridDim.x * waves) {
        const uint32_t v = g * 64 + lane;
        const bool valid = v < V;
        const uint32_t um = valid ? pool_umask(D, vmask[v]) : 0u;
        const uint32_t o = um ? off[v] : 0u, len = um ? off[v + 1] - o : 0u;
        for (uint32_t b = 0; b < n; b++) mine[b] = 0;
        wave_sync();
        const uint32_t ub = uni(wave_or32(um));
        const uint32_t incl = wave_incl_scan(len), total = rdlane(incl, 63);
        for (uint32_t b0 = 0; b0 < total; b0 += 64) {
            const TabStep st = tab_step(incl, len, total, b0, lane);
            const uint32_t own = st.own & 63u;
            const uint32_t oj = (uint32_t)__shfl((int)o, (int)own);
            const uint32_t uj = (uint32_t)__shfl((int)um, (int)own);
            const uint32_t m = st.act ? vmask[adj[oj + st.pos]] & uj : 0u;
            const uint64_t upto = st.below | (1ull << lane);
            for (uint32_t bs = ub; bs; bs &= bs - 1) {
                const uint32_t b = __ffs(bs) - 1;
                const uint64_t bal = __ballot((m >> b) & 1u);
                if (bal && st.last) slots[own * kmax + b] += (uint32_t)__popcll(bal & upto);
            }
        }
        wave_sync();
        if (valid)
            for (uint32_t b = 0; b < n; b++) poff[(size_t)b * (V + 1) + v] = mine[b];
        wave_sync();
    }
}

// The depth-0 candidates of each query (blockIdx.x) in degree order: by_deg walked in tiles
// of 256, candidates placed at their block prefix, each as its position in C(u).  Walking the
// roots of a capped search from the highest degree reaches the cap with fewer descents, as
// the cache + filter search does with its degree-ordered items.
__global__ void __launch_bounds__(256) k_root_order(const uint32_t* __restrict__ by_deg, uint32_t V,
                                                    const sgm_rootdesc* __restrict__ ds)
{
    const sgm_rootdesc& D = ds[blockIdx.x];
    const glb<uint32_t>* __restrict__ vmask = gcast(D.vmask);
    const glb<uint64_t>* __restrict__ bm = gcast(D.bm);
    const glb<uint32_t>* __restrict__ rank = gcast(D.rank);
    glb<uint32_t>* __restrict__ out = gcast(D.out);
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    for (uint32_t base = 0; base < V; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < V ? by_deg[i] : 0u;
        const bool c = i < V && ((vmask[v] >> D.u) & 1u);
        const uint64_t bal = __ballot(c);
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t pre = carry;
        for (uint32_t k = 0; k < wv; k++) pre += wsum[k];
        if (c) out[pre + mbcnt64(bal)] = cand_rank(bm, rank, v);
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

// Rank records of one query (blockIdx.y): rrec[w * rstride + b] = the position of w in C(b),
// or SGM_INVALID when w is not a candidate of b (b < rstride, a multiple of 4 >= n), so the
// pool pass reads every query vertex's rank of a neighbour in one 16-byte-aligned record.
__global__ void __launch_bounds__(256) k_pool_ranks(uint32_t V, const sgm_pooldesc* __restrict__ ds)
{
    const sgm_pooldesc& D = ds[blockIdx.y];
    const glb<uint32_t>* __restrict__ vmask = gcast(D.vmask);
    const glb<uint64_t>* __restrict__ bm = gcast(D.bm);
    const glb<uint32_t>* __restrict__ rank = gcast(D.rank);
    glb<uint32_t>* __restrict__ rrec = gcast(D.rrec);
    const uint32_t n = D.n, rs = D.rstride, W = (V + 63) / 64;
    const uint64_t total = (uint64_t)V * rs;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = (uint32_t)(i / rs), b = (uint32_t)(i % rs);
        uint32_t r = SGM_INVALID;
        if (b < n && ((vmask[w] >> b) & 1u)) r = cand_rank(bm + (size_t)b * W, rank + (size_t)b * W, w);
        rrec[i] = r;
    }
}

// The pool rows: entries of (b, v) at pool + pbase[b] + poff[b][v] (blockIdx.y = the query).
// A step loads each lane's neighbour's rank record once (rstride / 4 16-byte loads) and
// takes every query vertex from registers.  For each b the rows of a wave's 64 consecutive
// vertices are one contiguous run of b's pool, and the walk visits their entries in that
// order, so a step's hits of b go to the run's next free positions: lane b keeps that
// position, and a hit's place is it plus the hits of lower lanes (no per-row state).
__global__ void __launch_bounds__(256) k_pool_fill(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                   uint32_t V, const sgm_pooldesc* __restrict__ ds)
{
    const sgm_pooldesc& D = ds[blockIdx.y];
    const glb<uint32_t>* __restrict__ vmask = gcast(D.vmask);
    const glb<uint32_t>* __restrict__ poff = gcast((const uint32_t*)D.poff);
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const glb<u32x4>* __restrict__ rrec = gcast((const u32x4*)D.rrec);
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = D.n, nl = D.rstride >> 2;
    // lane b holds b's pool base
    const uint64_t my_base = lane < n ? (uint64_t)(size_t)(D.pool + D.pbase[lane]) : 0ull;
    for (uint32_t g = blockIdx.x * waves + wave; (uint64_t)g * 64 < V; g += gridDim.x * waves) {
        const uint32_t v = g * 64 + lane;
        const bool valid = v < V;
        const uint32_t um = valid ? pool_umask(D, vmask[v]) : 0u;
        const uint32_t o = um ? off[v] : 0u, len = um ? off[v + 1] - o : 0u;
        uint32_t my_ptr = lane < n ? poff[(size_t)lane * (V + 1) + g * 64] : 0u;  // lane b: b's next position
        const uint32_t ub = uni(wave_or32(um));
        const uint32_t incl = wave_incl_scan(len), total = rdlane(incl, 63);
        for (uint32_t b0 = 0; b0 < total; b0 += 64) {
            const TabStep st = tab_step(incl, len, total, b0, lane);
            const uint32_t own = st.own & 63u;
            const uint32_t oj = (uint32_t)__shfl((int)o, (int)own);
            // every lane takes part in the shuffle (the owner lane may hold no entry this step)
            const uint32_t uj = (uint32_t)__shfl((int)um, (int)own);
            const uint32_t w = st.act ? adj[oj + st.pos] : 0u;
            uint32_t rec[SGM_MAXQ];
#pragma unroll
            for (uint32_t i = 0; i < SGM_MAXQ / 4; i++) {
                u32x4 x = {SGM_INVALID, SGM_INVALID, SGM_INVALID, SGM_INVALID};
                if (i < nl && st.act) x = rrec[(size_t)w * nl + i];
                rec[4 * i] = x.x;
                rec[4 * i + 1] = x.y;
                rec[4 * i + 2] = x.z;
                rec[4 * i + 3] = x.w;
            }
#pragma unroll
            for (uint32_t b = 0; b < SGM_MAXQ; b++) {
                if (!((ub >> b) & 1u)) continue;
                const bool hit = ((uj >> b) & 1u) && rec[b] != SGM_INVALID;
                const uint64_t bal = __ballot(hit);
                if (!bal) continue;
                if (hit) {
                    const uint32_t p = rdlane(my_ptr, b) + mbcnt64(bal);
                    // streaming store: the pools are written once here
                    __builtin_nontemporal_store(rec[b], gcast((uint32_t*)rdlane64(my_base, b)) + p);
                }
                if (lane == b) my_ptr += (uint32_t)__popcll(bal);
            }
        }
    }
}

// Entry total and longest row of every directed edge (a, b) over the pool rows of C(a)
// (blockIdx.y = the edge; tot and maxv cleared before the launch).
__global__ void __launch_bounds__(256) k_edge_stats(const sgm_edgedesc* __restrict__ ds, uint32_t* __restrict__ tot,
                                                    uint32_t* __restrict__ maxv)
{
    const sgm_edgedesc& D = ds[blockIdx.y];
    const glb<uint32_t>* __restrict__ rows = gcast(D.rows);
    const glb<uint32_t>* __restrict__ voff = gcast(D.voff);
    uint32_t sum = 0, mx = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < D.nrows; r += gridDim.x * blockDim.x) {
        const uint32_t v = rows[r];
        const uint32_t len = voff[v + 1] - voff[v];
        sum += len;
        mx = len > mx ? len : mx;
    }
    sum = wave_sum32(sum);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)mx, o);
        mx = y > mx ? y : mx;
    }
    if (lane_id() == 0) {
        if (sum) atomicAdd(tot + D.slot, sum);
        if (mx) atomicMax(maxv + D.slot, mx);
    }
}

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total) and the maximum into *maxv, by
// one block of 256 threads; *tot (when set) also receives the total.
__device__ void scan_block(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
                           uint32_t* __restrict__ maxv, uint32_t* __restrict__ tot);

// Many scans at once, one block each (the row counts of every directed query edge of many
// queries into their row offsets, totals and maxima).
__global__ void __launch_bounds__(256) k_scan_segments(const sgm_scandesc* __restrict__ ds)
{
    const sgm_scandesc& D = ds[blockIdx.x];
    scan_block(D.in, D.n, D.out, D.maxv, D.tot);
}

__device__ void scan_block(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out,
                           uint32_t* __restrict__ maxv, uint32_t* __restrict__ tot)
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
    if (threadIdx.x == 0) {
        out[n] = carry;
        *maxv = cmax;
        if (tot) *tot = carry;
    }
}

// ================================================================== BFS-tree filters
// One refinement step of the CFL, CECI and DAF filters (inc/filter.h:317-425 generate /
// pruneCandidates, :602-788 CECI's frontier invalidation, :942-1020 DAF's passes) for every
// query of a batch.  Query q's step s is (w, P) = prog[q][1 + 2s], prog[q][2 + 2s]: data
// vertex v keeps bit `w & 31` of its membership word iff the OR over its neighbours of the
// source words covers P.  Source vmask (w >> 8 == 0): "v has a neighbour in C(p) for every
// pivot p in P" -- pruneCandidates, and generateCandidates when the bit still holds NLF(u).
// Source vmr = the NLF words (w >> 8 == 1): "v has a neighbour in NLF(u) for every u in P"
// -- CECI's invalidation of a frontier candidate.  A step changes only its own bit, which P
// never holds, so the words other lanes read are unaffected by its stores.  Lanes scan short
// lists alone; lists of more than 32 neighbours are scanned by the whole wave, 64 at a time,
// and every scan stops as soon as P is covered.
__global__ void __launch_bounds__(256) k_refine(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                uint32_t V, const sgm_qdesc* __restrict__ qs,
                                                const uint32_t* __restrict__ prog, uint32_t step)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    const uint32_t* __restrict__ pr = prog + (size_t)blockIdx.y * SGM_RPROG_WORDS;
    if (d->n == 0 || step >= pr[0]) return;
    const uint32_t w = pr[1 + 2 * step], P = pr[2 + 2 * step];
    const uint32_t bit = w & 31u;
    const uint32_t* __restrict__ src = (w >> 8) ? d->vmr : d->vmask;
    uint32_t* vm = d->vmask;
    const uint32_t lane = lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t base = wave * 64u; base < V; base += nwaves * 64u) {
        const uint32_t v = base + lane;
        const uint32_t m = v < V ? vm[v] : 0u;
        const bool mem = (m >> bit) & 1u;
        uint32_t b = 0, e = 0;
        if (mem) {
            b = off[v];
            e = off[v + 1];
        }
        bool keep = true;
        const bool alone = mem && e - b <= 32u;
        if (alone) {
            uint32_t acc = 0;
            for (uint32_t p = b; p < e && (acc & P) != P; p++) acc |= src[adj[p]];
            keep = (acc & P) == P;
        }
        uint64_t wide = __ballot(mem && !alone);
        while (wide) {
            const uint32_t l = (uint32_t)__ffsll((long long)wide) - 1u;
            wide &= wide - 1;
            const uint32_t lb = rdlane(b, l), le = rdlane(e, l);
            uint32_t acc = 0;
            for (uint32_t p = lb; p < le; p += 64u) {
                const uint32_t x = p + lane < le ? src[adj[p + lane]] : 0u;
                acc |= wave_or32(x);
                if ((acc & P) == P) break;
            }
            if (lane == l) keep = (acc & P) == P;
        }
        if (mem && !keep) vm[v] = m & ~(1u << bit);
    }
}

// vmr[v] = vmask[v] for every query of a batch (CECI keeps the NLF words beside its sets).
__global__ void __launch_bounds__(256) k_copy_vmask(uint32_t V, const sgm_qdesc* __restrict__ qs)
{
    const sgm_qdesc* d = qs + blockIdx.y;
    if (d->n == 0) return;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) d->vmr[v] = d->vmask[v];
}

// Row sums of one aux table for the CFL and GuP order estimates (inc/order.h:197-239
// estimatePathEmbeddsingsNum, :386-410 count_embed): s_j = sum over the entries k of row j of
// in[idx[k]] (in == null: 1 each, the row length), wrapping mod 2^64 as the reference's
// size_t sums do; out[j] = s_j, or out[j] *= s_j when mul is set; *total += sum of s_j.
__global__ void __launch_bounds__(256) k_row_gather(const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tidx,
                                                    const uint32_t* __restrict__ rowv, uint32_t rows,
                                                    const unsigned long long* __restrict__ in,
                                                    unsigned long long* __restrict__ out, int mul,
                                                    unsigned long long* __restrict__ total)
{
    unsigned long long t = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        const uint32_t v = rowv[j];
        const uint32_t b = toff[v], e = toff[v + 1];
        unsigned long long s = 0;
        if (in)
            for (uint32_t k = b; k < e; k++) s += in[tidx[k]];
        else
            s = e - b;
        out[j] = mul ? out[j] * s : s;
        t += s;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)t, o);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(t >> 32), o);
        t += ((unsigned long long)hi << 32) | lo;
    }
    if (lane_id() == 0 && t) atomicAdd(total, t);
}

// computeWeightArray's row passes (inc/order.h:1003-1022), one descriptor per (vertex, tree
// child) of one reverse-BFS step of many queries (blockIdx.y): a row's weight is the u32 sum
// (wrapping) of the child's weights at the row's positions, kept when below the current one
// (atomicMin: a vertex's children share the step).  A wave takes 64 rows, one per lane;
// rows longer than 16 entries (power-law hubs) are summed by the whole wave, one at a time.
__global__ void __launch_bounds__(256) k_weight_step(const sgm_wdesc* __restrict__ ds)
{
    const sgm_wdesc& D = ds[blockIdx.y];
    const uint32_t rows = D.rows, lane = lane_id();
    const uint32_t* __restrict__ toff = D.toff;
    const uint32_t* __restrict__ tidx = D.tidx;
    const uint32_t* __restrict__ in = D.in;
    uint32_t* __restrict__ out = D.out;
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t g = blockIdx.x * waves + (threadIdx.x >> 6); (uint64_t)g * 64 < rows; g += gridDim.x * waves) {
        const uint32_t j = g * 64 + lane;
        const bool valid = j < rows;
        if (!in) {
            if (valid) out[j] = D.fill;
            continue;
        }
        uint32_t b = 0, e = 0;
        if (valid) {
            const uint32_t v = D.rowv[j];
            b = toff[v];
            e = toff[v + 1];
        }
        const bool lng = valid && e - b > 16;
        uint32_t s = 0;
        if (valid && !lng)
            for (uint32_t k = b; k < e; k++) s += in[tidx[k]];
        uint64_t lm = __ballot(lng);
        while (lm) {
            const uint32_t l = (uint32_t)__ffsll((long long)lm) - 1u;
            lm &= lm - 1;
            const uint32_t bl = rdlane(b, l), el = rdlane(e, l);
            uint32_t ws = 0;
            for (uint32_t k = bl + lane; k < el; k += 64) ws += in[tidx[k]];
            ws = wave_sum32(ws);
            if (lane == l) s = ws;
        }
        if (valid) atomicMin(out + j, s);
    }
}

__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long* __restrict__ p, size_t n, unsigned long long val)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = val;
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

void launch_gql_prepass(hipStream_t s, const uint64_t* dtab, uint32_t V, const uint32_t* ge_cnt, const sgm_qdesc* qs,
                        uint32_t nq)
{
    dim3 grid(clampg(((uint64_t)V + 1023) / 1024, 1, 16384 / (nq ? nq : 1) + 1), nq);
    hipLaunchKernelGGL(k_gql_prepass, grid, dim3(256), 0, s, dtab, V, ge_cnt, qs);
}

void launch_gql_phase(hipStream_t s, const uint32_t* roff, const uint32_t* radj, uint32_t V, const uint32_t* ge_cnt,
                      const sgm_qdesc* qs, uint32_t nq, uint32_t phase)
{
    // the phase visits short lists and skips idle rank blocks by their level-2 summaries: a
    // modest number of blocks per query (4096 in all measured best at 64 and 200 queries,
    // 8192 for launches of several hundred)
    const uint32_t total = nq > 256 ? 8192u : 4096u;
    dim3 grid(clampg(total / (nq ? nq : 1), 8, 256), nq);
    hipLaunchKernelGGL(k_gql_phase, grid, dim3(256), 0, s, roff, radj, V, ge_cnt, qs, phase);
}

void launch_fill_strided(hipStream_t s, uint32_t* p, size_t len, size_t stride, uint32_t reps, uint32_t val)
{
    dim3 grid(clampg((len + 255) / 256, 1, 4096), reps);
    hipLaunchKernelGGL(k_fill_strided, grid, dim3(256), 0, s, p, len, stride, val);
}

void launch_nlf_rank(hipStream_t s, const uint32_t* ge_cnt, uint32_t V, const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_nlf_rank, grid, dim3(256), 0, s, ge_cnt, V, qs);
}

void launch_nlf_lab(hipStream_t s, const uint32_t* off, const uint32_t* lab, const uint32_t* run_off,
                    const uint32_t* run_lab, const uint32_t* run_cnt, uint32_t V, const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_nlf_lab, grid, dim3(256), 0, s, off, lab, run_off, run_lab, run_cnt, V, qs);
}

void launch_nlf_rank_lab(hipStream_t s, const uint32_t* by_deg, const uint32_t* off, const uint32_t* lab,
                         const uint32_t* run_off, const uint32_t* run_lab, const uint32_t* run_cnt, uint32_t V,
                         const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_nlf_rank_lab, grid, dim3(256), 0, s, by_deg, off, lab, run_off, run_lab, run_cnt, V, qs);
}

void launch_unrank(hipStream_t s, const uint32_t* rank, uint32_t V, const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    hipLaunchKernelGGL(k_unrank, grid, dim3(256), 0, s, rank, V, qs);
}

void launch_rank_adj(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* by_deg,
                     const uint32_t* rank, const uint32_t* roff, uint32_t V, uint32_t* radj)
{
    hipLaunchKernelGGL(k_rank_adj, dim3(clampg(((uint64_t)V + 3) / 4, 1, 16384)), dim3(256), 0, s, off, adj, by_deg,
                       rank, roff, V, radj);
}

void launch_cand_stats(hipStream_t s, const uint32_t* off, uint32_t V, const sgm_qdesc* qs, uint32_t nq,
                       uint32_t* cnt, unsigned long long* degsum)
{
    // a few blocks per query, many vertices per thread: the per-block atomics stay few
    dim3 grid(clampg(((uint64_t)V + 4095) / 4096, 1, 32), nq);
    hipLaunchKernelGGL(k_cand_stats, grid, dim3(256), 0, s, off, V, qs, cnt, degsum);
}

void launch_clear_flags(hipStream_t s, const sgm_qdesc* qs, uint32_t nq)
{
    if (nq) hipLaunchKernelGGL(k_clear_flags, dim3(nq), dim3(64), 0, s, qs);
}

void launch_gather_counts(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t* out)
{
    if (nq) hipLaunchKernelGGL(k_gather_counts, dim3(nq), dim3(64), 0, s, cs, out);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t val)
{
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(clampg((n + 255) / 256, 1, 8192)), dim3(256), 0, s, p, n, val);
}

void launch_nbr_deg(hipStream_t s, const uint32_t* roff, const uint32_t* radj, uint32_t V, uint64_t* dtab)
{
    hipLaunchKernelGGL(k_nbr_deg, dim3(clampg(((uint64_t)V + 3) / 4, 1, 16384)), dim3(256), 0, s, roff, radj, V,
                       reinterpret_cast<uint8_t*>(dtab));
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

// grid.y / grid.x hold at most 65535 descriptors per launch: longer lists go in pieces
static const uint32_t kMaxGridY = 65535;

void launch_root_order(hipStream_t s, const uint32_t* by_deg, uint32_t V, const sgm_rootdesc* ds, uint32_t nq)
{
    if (nq && V) hipLaunchKernelGGL(k_root_order, dim3(nq), dim3(256), 0, s, by_deg, V, ds);
}

void launch_pool_count(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_pooldesc* ds,
                       uint32_t nq, uint32_t kmax)
{
    if (!nq || !V || !kmax) return;
    const uint32_t xcap = 16384 / nq > 8 ? 16384 / nq : 8u;
    const size_t smem = sizeof(uint32_t) * 4 * 64 * kmax;
    for (uint32_t s0 = 0; s0 < nq; s0 += kMaxGridY) {
        const uint32_t k = nq - s0 < kMaxGridY ? nq - s0 : kMaxGridY;
        dim3 grid(clampg(((uint64_t)V + 255) / 256, 1, xcap), k);
        hipLaunchKernelGGL(k_pool_count, grid, dim3(256), smem, s, off, adj, V, ds + s0, kmax);
    }
}

void launch_pool_fill(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_pooldesc* ds,
                      uint32_t nq)
{
    if (!nq || !V) return;
    for (uint32_t s0 = 0; s0 < nq; s0 += kMaxGridY) {
        const uint32_t k = nq - s0 < kMaxGridY ? nq - s0 : kMaxGridY;
        dim3 grid(clampg(((uint64_t)V * 4 + 255) / 256, 1, 16384 / k + 1), k);
        hipLaunchKernelGGL(k_pool_ranks, grid, dim3(256), 0, s, V, ds + s0);
    }
    const uint32_t xcap = 16384 / nq > 8 ? 16384 / nq : 8u;
    for (uint32_t s0 = 0; s0 < nq; s0 += kMaxGridY) {
        const uint32_t k = nq - s0 < kMaxGridY ? nq - s0 : kMaxGridY;
        dim3 grid(clampg(((uint64_t)V + 255) / 256, 1, xcap), k);
        hipLaunchKernelGGL(k_pool_fill, grid, dim3(256), 0, s, off, adj, V, ds + s0);
    }
}

void launch_edge_stats(hipStream_t s, const sgm_edgedesc* ds, uint32_t m, uint32_t max_rows, uint32_t* tot,
                       uint32_t* maxv)
{
    if (!m || !max_rows) return;
    for (uint32_t s0 = 0; s0 < m; s0 += kMaxGridY) {
        const uint32_t k = m - s0 < kMaxGridY ? m - s0 : kMaxGridY;
        dim3 grid(clampg(((uint64_t)max_rows + 255) / 256, 1, 64), k);
        hipLaunchKernelGGL(k_edge_stats, grid, dim3(256), 0, s, ds + s0, tot, maxv);
    }
}

void launch_scan_segments(hipStream_t s, const sgm_scandesc* ds, uint32_t m)
{
    for (uint32_t s0 = 0; s0 < m; s0 += kMaxGridY) {
        const uint32_t k = m - s0 < kMaxGridY ? m - s0 : kMaxGridY;
        hipLaunchKernelGGL(k_scan_segments, dim3(k), dim3(256), 0, s, ds + s0);
    }
}

void launch_hub_bits(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* hubs, uint32_t n_hubs,
                     uint32_t* bits, uint32_t words)
{
    hipLaunchKernelGGL(k_hub_bits, dim3(clampg(n_hubs, 1, 65535)), dim3(256), 0, s, off, adj, hubs, n_hubs, bits, words);
}

void launch_refine(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_qdesc* qs,
                   uint32_t nq, const uint32_t* prog, uint32_t step)
{
    dim3 grid(clampg((V + 255) / 256, 1, 2048), nq);
    if (nq && V) hipLaunchKernelGGL(k_refine, grid, dim3(256), 0, s, off, adj, V, qs, prog, step);
}

void launch_copy_vmask(hipStream_t s, uint32_t V, const sgm_qdesc* qs, uint32_t nq)
{
    dim3 grid(clampg((V + 255) / 256, 1, 4096), nq);
    if (nq && V) hipLaunchKernelGGL(k_copy_vmask, grid, dim3(256), 0, s, V, qs);
}

void launch_row_gather(hipStream_t s, const uint32_t* toff, const uint32_t* tidx, const uint32_t* rowv, uint32_t rows,
                       const unsigned long long* in, unsigned long long* out, int mul, unsigned long long* total)
{
    if (rows)
        hipLaunchKernelGGL(k_row_gather, dim3(clampg(((uint64_t)rows + 255) / 256, 1, 1024)), dim3(256), 0, s, toff,
                           tidx, rowv, rows, in, out, mul, total);
}

void launch_weight_step(hipStream_t s, const sgm_wdesc* ds, uint32_t m, uint32_t max_rows)
{
    if (!m || !max_rows) return;
    for (uint32_t s0 = 0; s0 < m; s0 += kMaxGridY) {
        const uint32_t k = m - s0 < kMaxGridY ? m - s0 : kMaxGridY;
        dim3 grid(clampg(((uint64_t)max_rows + 255) / 256, 1, 1024), k);
        hipLaunchKernelGGL(k_weight_step, grid, dim3(256), 0, s, ds + s0);
    }
}

void launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long val)
{
    if (n) hipLaunchKernelGGL(k_fill_u64, dim3(clampg((n + 255) / 256, 1, 4096)), dim3(256), 0, s, p, n, val);
}

}  // namespace sgm
