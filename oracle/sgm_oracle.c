/*
 * sgm_oracle.c -- CPU restatement of the reference's filter -> order -> enumerate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product
 * library (subgraphmatching_amd/).  Every function cites the reference file:line it
 * restates (paths relative to the Orrinn/SubgraphMatching source tree).
 *
 * Pinning (what checks this restatement independently of the GPU path):
 *   - orc_nlf:            closed form deg(v) >= deg_q(u)                 (tests/test_oracle_pins.py)
 *   - orc_gql:            brute-force Hall's-condition restatement with the same 2-round
 *                         schedule on tiny graphs; GQL subset-of NLF; soundness (every
 *                         vertex used by a brute-force embedding survives)
 *   - orc_table:          definition via adjacency test on every pair
 *   - orc_order_ri/gql:   hand-worked known answers exercising every tie-break rule
 *   - orc_plan:           hand-worked known answers; the tie order of the op sort is
 *                         checked against libstdc++ std::sort on random inputs
 *   - orc_enum_*:         counts vs brute-force injective homomorphism enumeration and
 *                         closed forms (triangles, paths, stars); per-depth node counts and
 *                         intersection counts with pruning off vs brute-force prefix
 *                         enumeration.
 *   - DAF prune decisions (node counts with failing sets on) are pinned only by a
 *     hand-worked example and by the invariant nodes(fs on) <= nodes(fs off): the exact
 *     reference behaviour on large inputs is "parity unpinned" (the reference does not
 *     build in this image).
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_INVALID 0xFFFFFFFFu
#define ORC_MAXQ 32
#define ORC_MAXOPS 512

static inline uint32_t dg(const uint32_t *off, uint32_t v) { return off[v + 1] - off[v]; }

/* ------------------------------------------------------------------------------------
 * Sorted-set intersection.  isect_merge restates IntersectionMerge (inc/common_ops.h:
 * 154-199).  The reference's Intersection (inc/common_ops.h:248-254) calls its AVX2 variant
 * when built with INTERSECTION_AVX2 (inc/common_ops.h:18-150), as its cmake build does; so
 * that the CPU baseline is timed on the same kind of intersection, isect uses an AVX2 one
 * when this file is compiled with -mavx2.  The output list is the same as the merge's (sets
 * are strictly increasing), checked by orc_isect_test against isect_merge.
 * ---------------------------------------------------------------------------------- */
static uint32_t isect_merge(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out)
{
    uint32_t i = 0, j = 0, c = 0;
    if (na == 0 || nb == 0) return 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) i++;
        else if (a[i] > b[j]) j++;
        else { out[c++] = a[i]; i++; j++; }
    }
    return c;
}

#ifdef __AVX2__
#include <immintrin.h>
/* The reference's strategy split: with the longer list more than twice the shorter, each
 * key of the shorter list is broadcast and compared with 8 consecutive elements of the
 * longer one, the window moving on once the key passes its last element; otherwise two keys
 * at a time are compared with four elements at a time (one 8-lane compare), advancing
 * whichever side has the smaller last element.  A scalar merge finishes the tails. */
static uint32_t isect_avx2(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out)
{
    if (na == 0 || nb == 0) return 0;
    if (na > nb) { const uint32_t *t = a; a = b; b = t; uint32_t tn = na; na = nb; nb = tn; }
    uint32_t i = 0, j = 0, c = 0;
    if (nb / na > 2) {
        while (i < na && j + 8 <= nb) {
            const __m256i w = _mm256_loadu_si256((const __m256i *)(b + j));
            const uint32_t last = b[j + 7];
            while (i < na && a[i] <= last) {
                const __m256i k = _mm256_set1_epi32((int)a[i]);
                if (_mm256_movemask_epi8(_mm256_cmpeq_epi32(k, w))) out[c++] = a[i];
                i++;
            }
            j += 8;
        }
    } else {
        /* lanes 0-3: key a[i] against b[j..j+3]; lanes 4-7: key a[i+1] against the same */
        const __m256i kidx = _mm256_set_epi32(1, 1, 1, 1, 0, 0, 0, 0);
        const __m256i widx = _mm256_set_epi32(3, 2, 1, 0, 3, 2, 1, 0);
        while (i + 2 <= na && j + 4 <= nb) {
            const __m128i k2 = _mm_loadl_epi64((const __m128i *)(a + i));          /* a[i], a[i+1] */
            const __m128i w4 = _mm_loadu_si128((const __m128i *)(b + j));          /* b[j..j+3] */
            const __m256i kk = _mm256_permutevar8x32_epi32(_mm256_castsi128_si256(k2), kidx);
            const __m256i ww = _mm256_permutevar8x32_epi32(_mm256_castsi128_si256(w4), widx);
            const uint32_t m = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi32(kk, ww));
            if (m & 0xFFFFu) out[c++] = a[i];
            if (m >> 16) out[c++] = a[i + 1];
            const uint32_t ka = a[i + 1], wb = b[j + 3];
            if (ka == wb) { i += 2; j += 4; }
            else if (ka > wb) j += 4;
            else i += 2;
        }
    }
    /* tails: a scalar merge of what is left (both sides strictly increasing) */
    while (i < na && j < nb) {
        if (a[i] < b[j]) i++;
        else if (a[i] > b[j]) j++;
        else { out[c++] = a[i]; i++; j++; }
    }
    return c;
}
#endif

static uint32_t isect(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out)
{
#ifdef __AVX2__
    return isect_avx2(a, na, b, nb, out);
#else
    return isect_merge(a, na, b, nb, out);
#endif
}

/* Checks the intersection used by the oracle against the merge (1 = same output). */
int orc_isect_test(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out1, uint32_t *out2,
                   uint32_t *n_out)
{
    const uint32_t c1 = isect(a, na, b, nb, out1), c2 = isect_merge(a, na, b, nb, out2);
    *n_out = c1;
    if (c1 != c2) return 0;
    for (uint32_t k = 0; k < c1; k++)
        if (out1[k] != out2[k]) return 0;
    return 1;
}

/* The same intersection for the other oracle files (the dynamic-order searches). */
uint32_t orc_isect(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out)
{
    return isect(a, na, b, nb, out);
}

/* 1 when isect is the AVX2 intersection. */
int orc_isect_avx2(void)
{
#ifdef __AVX2__
    return 1;
#else
    return 0;
#endif
}

/* ------------------------------------------------------------------------------------
 * NLF (inc/filter.h:24-58 NLFByVertex, :80-93 _NLF_Filter).  v may take u iff
 * label(v) = label(u) (the scan runs over getVertexByLabel, inc/graph.h:96-112),
 * deg(v) >= deg_q(u), and for every label l among u's neighbours v has at least as many
 * neighbours of label l as u has (the NLF maps, inc/graph.h:76-91).  dlab / qlab NULL
 * means every label is 0; then the test reduces to deg(v) >= deg_q(u) (a query vertex of
 * degree 0 has an empty NLF map and accepts every data vertex).  Candidates are written in
 * ascending vertex id into cand[u*V ...] (sortCandidates, :13-22).  Returns 1, or 0 at the
 * first empty set (the reference stops there, :87-90).
 * ---------------------------------------------------------------------------------- */
static uint32_t lab_of(const uint32_t *lab, uint32_t v) { return lab ? lab[v] : 0u; }

int orc_nlf(uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab, uint32_t n,
            const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab, uint32_t *cand, uint32_t *cnt)
{
    for (uint32_t u = 0; u < n; u++) {
        uint32_t qd = qoff[u + 1] - qoff[u], c = 0;
        /* u's NLF map as (label, count) pairs */
        uint32_t ql[ORC_MAXQ], qc[ORC_MAXQ], np = 0;
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
            uint32_t l = lab_of(qlab, qadj[p]), k = 0;
            while (k < np && ql[k] != l) k++;
            if (k == np) { ql[np] = l; qc[np++] = 0; }
            qc[k]++;
        }
        for (uint32_t v = 0; v < V; v++) {
            if (lab_of(dlab, v) != lab_of(qlab, u) || dg(off, v) < qd) continue;
            int ok = 1;
            for (uint32_t k = 0; k < np && ok; k++) {
                uint32_t have = 0;
                for (uint32_t p = off[v]; p < off[v + 1]; p++) have += lab_of(dlab, adj[p]) == ql[k];
                ok = have >= qc[k];
            }
            if (ok) cand[(size_t)u * V + c++] = v;
        }
        cnt[u] = c;
        if (c == 0) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------------------
 * GQL local semi-perfect matching (inc/filter.h:168-204 verifyExactTwigIso).  Left side
 * = N_q(u), right side = N_G(v), edge (i,j) iff valid[N_q(u)[i]][N_G(v)[j]].  True iff
 * a matching saturates the left side.  The reference computes a maximum matching with
 * a cheap greedy start plus BFS augmentation (inc/filter.h:95-166); any exact maximum
 * matching gives the same boolean, so this restatement uses Kuhn's DFS augmentation.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t *l_off;   /* [qmax+1] */
    uint32_t *l_edges; /* [qmax * dmax] right positions */
    int32_t *r_match;  /* [dmax] */
    uint32_t *r_seen;  /* [dmax] stamp */
    uint32_t stamp;
} gql_scratch;

static int kuhn(gql_scratch *s, uint32_t i)
{
    for (uint32_t e = s->l_off[i]; e < s->l_off[i + 1]; e++) {
        uint32_t r = s->l_edges[e];
        if (s->r_seen[r] == s->stamp) continue;
        s->r_seen[r] = s->stamp;
        if (s->r_match[r] < 0 || kuhn(s, (uint32_t)s->r_match[r])) {
            s->r_match[r] = (int32_t)i;
            return 1;
        }
    }
    return 0;
}

static int lsm_check(uint32_t V, const uint32_t *off, const uint32_t *adj,
                     const uint32_t *qoff, const uint32_t *qadj, const uint8_t *valid,
                     uint32_t u, uint32_t v, gql_scratch *s)
{
    uint32_t L = qoff[u + 1] - qoff[u], R = dg(off, v), ec = 0;
    const uint32_t *qn = qadj + qoff[u], *dn = adj + off[v];
    for (uint32_t i = 0; i < L; i++) {
        s->l_off[i] = ec;
        const uint8_t *row = valid + (size_t)qn[i] * V;
        for (uint32_t j = 0; j < R; j++)
            if (row[dn[j]]) s->l_edges[ec++] = j;
    }
    s->l_off[L] = ec;
    for (uint32_t j = 0; j < R; j++) s->r_match[j] = -1;
    for (uint32_t i = 0; i < L; i++) {
        s->stamp++;
        if (s->stamp == 0) { memset(s->r_seen, 0, sizeof(uint32_t) * (R ? R : 1)); s->stamp = 1; }
        if (!kuhn(s, i)) return 0;
    }
    return 1;
}

/* GQL filter (inc/filter.h:233-313 _GQL_Filter): NLF, then exactly two refinement rounds;
 * each round visits u = 0..n-1 and, for each still-valid v in C(u), removes v if the
 * local matching fails.  A removal is visible at once to later checks (valid[u][v] is
 * cleared immediately, :285-287).  Then compaction in ascending order (:206-220) and the
 * empty-set check (:222-230).  stats (optional, 4 entries): checks executed, removals
 * in round 0, removals in round 1, reserved. */
int orc_gql(uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab, uint32_t n,
            const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab, uint32_t *cand, uint32_t *cnt,
            uint64_t *stats)
{
    if (stats) memset(stats, 0, 4 * sizeof(uint64_t));
    if (!orc_nlf(V, off, adj, dlab, n, qoff, qadj, qlab, cand, cnt)) return 0;
    uint8_t *valid = (uint8_t *)calloc((size_t)n * V, 1);
    uint32_t qmax = 0, dmax = 0;
    for (uint32_t u = 0; u < n; u++) if (qoff[u + 1] - qoff[u] > qmax) qmax = qoff[u + 1] - qoff[u];
    for (uint32_t v = 0; v < V; v++) if (dg(off, v) > dmax) dmax = dg(off, v);
    gql_scratch s;
    s.l_off = (uint32_t *)malloc(sizeof(uint32_t) * (qmax + 1));
    s.l_edges = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)qmax * dmax + 1));
    s.r_match = (int32_t *)malloc(sizeof(int32_t) * (dmax + 1));
    s.r_seen = (uint32_t *)calloc(dmax + 1, sizeof(uint32_t));
    s.stamp = 0;
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = 0; j < cnt[u]; j++) valid[(size_t)u * V + cand[(size_t)u * V + j]] = 1;
    for (uint32_t l = 0; l < 2; l++)
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t j = 0; j < cnt[u]; j++) {
                uint32_t v = cand[(size_t)u * V + j];
                if (v == ORC_INVALID) continue;
                if (stats) stats[0]++;
                if (!lsm_check(V, off, adj, qoff, qadj, valid, u, v, &s)) {
                    cand[(size_t)u * V + j] = ORC_INVALID;
                    valid[(size_t)u * V + v] = 0;
                    if (stats) stats[1 + l]++;
                }
            }
    int ok = 1;
    for (uint32_t u = 0; u < n; u++) {
        uint32_t c = 0;
        for (uint32_t j = 0; j < cnt[u]; j++) {
            uint32_t v = cand[(size_t)u * V + j];
            if (v != ORC_INVALID) cand[(size_t)u * V + c++] = v;
        }
        cnt[u] = c;
        if (c == 0) ok = 0;
    }
    free(valid); free(s.l_off); free(s.l_edges); free(s.r_match); free(s.r_seen);
    return ok;
}

/* ------------------------------------------------------------------------------------
 * Auxiliary edge table for one directed query edge (a,b) (inc/table.h:7-137 buildTables;
 * inc/graph.h:13-49 Edges).  out_off[i] .. out_off[i+1] lists, ascending, the positions k
 * in C(b) with (C(a)[i], C(b)[k]) in E(G); out_v holds C(b)[k].  The reference builds one
 * direction by scanning and the other by a counting-sort transpose; both give this same
 * canonical content.  With out_idx == NULL only the edge count is returned.  *max_deg
 * receives the longest list (Edges::_max_degree).  pos_scratch must hold V entries.
 * ---------------------------------------------------------------------------------- */
uint64_t orc_table(uint32_t V, const uint32_t *off, const uint32_t *adj,
                   const uint32_t *cand_a, uint32_t cnt_a, const uint32_t *cand_b, uint32_t cnt_b,
                   uint32_t *out_off, uint32_t *out_idx, uint32_t *out_v, uint32_t *max_deg,
                   int32_t *pos_scratch)
{
    for (uint32_t v = 0; v < V; v++) pos_scratch[v] = -1;
    for (uint32_t k = 0; k < cnt_b; k++) pos_scratch[cand_b[k]] = (int32_t)k;
    uint64_t e = 0;
    uint32_t md = 0;
    for (uint32_t i = 0; i < cnt_a; i++) {
        uint32_t v = cand_a[i], ld = 0;
        if (out_off) out_off[i] = (uint32_t)e;
        for (uint32_t p = off[v]; p < off[v + 1]; p++) {
            int32_t k = pos_scratch[adj[p]];
            if (k >= 0) {
                if (out_idx) { out_idx[e] = (uint32_t)k; out_v[e] = adj[p]; }
                e++; ld++;
            }
        }
        if (ld > md) md = ld;
    }
    if (out_off) out_off[cnt_a] = (uint32_t)e;
    if (max_deg) *max_deg = md;
    return e;
}

/* ------------------------------------------------------------------------------------
 * RI order (inc/order.h:1138-1281 Order_RI).
 * ---------------------------------------------------------------------------------- */
static int q_has_edge(const uint32_t *qoff, const uint32_t *qadj, uint32_t a, uint32_t b)
{
    for (uint32_t p = qoff[a]; p < qoff[a + 1]; p++) if (qadj[p] == b) return 1;
    return 0;
}

void orc_order_ri(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, uint32_t *order)
{
    uint8_t visited[ORC_MAXQ] = {0};
    uint32_t tie[ORC_MAXQ], tmp[ORC_MAXQ], nt, ntmp;
    /* root: maximum degree, first (lowest id) on ties (:1145-1151) */
    order[0] = 0;
    for (uint32_t i = 1; i < n; i++)
        if (qoff[i + 1] - qoff[i] > qoff[order[0] + 1] - qoff[order[0]]) order[0] = i;
    visited[order[0]] = 1;
    for (uint32_t i = 1; i < n; i++) {
        /* most backward neighbours (:1159-1179); count 0 ties are kept */
        uint32_t max_bn = 0;
        nt = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (visited[u]) continue;
            uint32_t bn = 0;
            for (uint32_t j = 0; j < i; j++) if (q_has_edge(qoff, qadj, u, order[j])) bn++;
            if (bn > max_bn) { max_bn = bn; nt = 0; tie[nt++] = u; }
            else if (bn == max_bn) tie[nt++] = u;
        }
        if (nt != 1) {
            /* tie-break 1 (:1182-1221): number of ordered vertices that share a neighbour
             * with u's unvisited neighbours */
            memcpy(tmp, tie, sizeof(uint32_t) * nt); ntmp = nt; nt = 0;
            uint32_t best = 0;
            for (uint32_t t = 0; t < ntmp; t++) {
                uint32_t u = tmp[t], ufn[ORC_MAXQ], nf = 0;
                for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) if (!visited[qadj[p]]) ufn[nf++] = qadj[p];
                uint32_t cc = 0, scratch[ORC_MAXQ];
                for (uint32_t j = 0; j < i; j++) {
                    uint32_t uu = order[j];
                    if (isect(qadj + qoff[uu], qoff[uu + 1] - qoff[uu], ufn, nf, scratch) > 0) cc++;
                }
                if (cc > best) { best = cc; nt = 0; tie[nt++] = u; }
                else if (cc == best) tie[nt++] = u;
            }
        }
        if (nt != 1) {
            /* tie-break 2 (:1223-1267): number of u's unvisited neighbours adjacent to no
             * ordered vertex */
            memcpy(tmp, tie, sizeof(uint32_t) * nt); ntmp = nt; nt = 0;
            uint32_t best = 0;
            for (uint32_t t = 0; t < ntmp; t++) {
                uint32_t u = tmp[t], cc = 0;
                for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
                    uint32_t uu = qadj[p];
                    if (visited[uu]) continue;
                    int ok = 1;
                    for (uint32_t j = 0; j < i; j++) if (q_has_edge(qoff, qadj, uu, order[j])) { ok = 0; break; }
                    if (ok) cc++;
                }
                if (cc > best) { best = cc; nt = 0; tie[nt++] = u; }
                else if (cc == best) tie[nt++] = u;
            }
        }
        order[i] = tie[0];
        visited[order[i]] = 1;
    }
}

/* GQL order (inc/order.h:563-597 Order_GQL, :11-35 GQLStartVertex). */
void orc_order_gql(uint32_t V, uint32_t n, const uint32_t *qoff, const uint32_t *qadj,
                   const uint32_t *cnt, uint32_t *order)
{
    uint8_t visited[ORC_MAXQ] = {0}, adjacent[ORC_MAXQ] = {0};
    uint32_t s = 0;
    for (uint32_t i = 1; i < n; i++) {
        if (cnt[i] < cnt[s]) s = i;
        else if (cnt[i] == cnt[s] && qoff[i + 1] - qoff[i] > qoff[s + 1] - qoff[s]) s = i;
    }
    order[0] = s;
    visited[s] = 1;
    for (uint32_t p = qoff[s]; p < qoff[s + 1]; p++) adjacent[qadj[p]] = 1;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t next = 0, minv = V + 1;
        for (uint32_t j = 0; j < n; j++) {
            if (visited[j] || !adjacent[j]) continue;
            if (cnt[j] < minv) { minv = cnt[j]; next = j; }
            else if (cnt[j] == minv && qoff[j + 1] - qoff[j] > qoff[next + 1] - qoff[next]) next = j;
        }
        order[i] = next;
        visited[next] = 1;
        for (uint32_t p = qoff[next]; p < qoff[next + 1]; p++) adjacent[qadj[p]] = 1;
    }
}

/* ------------------------------------------------------------------------------------
 * Intersection-cache plan (inc/code_gen.h:573-847 VertexSetIR / PlanIR).
 *
 * The reference sorts candidate ops with std::sort (introsort, not stable); ops that
 * compare equal keep an implementation-defined relative order, which decides the
 * _v_depth kept for a shared op.  orc_sort_ops below restates libstdc++'s std::sort step
 * by step (median-of-three introsort with a 16-element threshold, heap-sort fallback,
 * final insertion sort), so the kept _v_depth matches a g++/libstdc++ build of the
 * reference.  tests/test_oracle_pins.py compares this sort with std::sort itself.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t depth;   /* depth at which the set is computed */
    uint32_t label;   /* label of the follow vertex (VertexSetIR::_v_label) */
    uint32_t vdepth;  /* depth of the follow vertex that produced it */
    uint32_t size;
    uint32_t e[ORC_MAXQ]; /* sorted depth indices of the backward neighbours */
} orc_op;

static int op_less(const orc_op *a, const orc_op *b)
{
    if (a == b) return 0;
    if (a->depth != b->depth) return a->depth < b->depth;
    if (a->label != b->label) return a->label < b->label;
    if (a->size != b->size) return a->size < b->size;
    for (uint32_t i = 0; i < a->size; i++)
        if (a->e[i] != b->e[i]) return a->e[i] < b->e[i];
    return 0;
}

static void op_swap(orc_op *a, orc_op *b) { orc_op t = *a; *a = *b; *b = t; }

static void ins_sort(orc_op *f, long n)
{
    for (long i = 1; i < n; i++) {
        orc_op val = f[i];
        if (op_less(&val, &f[0])) {
            memmove(f + 1, f, sizeof(orc_op) * i);
            f[0] = val;
        } else {
            long j = i;
            while (op_less(&val, &f[j - 1])) { f[j] = f[j - 1]; j--; }
            f[j] = val;
        }
    }
}

static void unguarded_ins_sort(orc_op *f, long from, long n)
{
    for (long i = from; i < n; i++) {
        orc_op val = f[i];
        long j = i;
        while (op_less(&val, &f[j - 1])) { f[j] = f[j - 1]; j--; }
        f[j] = val;
    }
}

static void sift_down(orc_op *f, long hole, long len, orc_op val)
{
    /* heap adjust as in libstdc++ __adjust_heap: move the hole to a leaf, then push up */
    long top = hole, child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (op_less(&f[child], &f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    long parent = (hole - 1) / 2;
    while (hole > top && op_less(&f[parent], &val)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = val;
}

static void heap_sort_range(orc_op *f, long n)
{
    /* std::__partial_sort(first, last, last): make_heap then sort_heap */
    if (n < 2) return;
    for (long parent = (n - 2) / 2;; parent--) {
        orc_op v = f[parent];
        sift_down(f, parent, n, v);
        if (parent == 0) break;
    }
    for (long last = n - 1; last > 0; last--) {
        orc_op v = f[last];
        f[last] = f[0];
        sift_down(f, 0, last, v);
    }
}

static void median_to_first(orc_op *r, orc_op *a, orc_op *b, orc_op *c)
{
    if (op_less(a, b)) {
        if (op_less(b, c)) op_swap(r, b);
        else if (op_less(a, c)) op_swap(r, c);
        else op_swap(r, a);
    } else if (op_less(a, c)) op_swap(r, a);
    else if (op_less(b, c)) op_swap(r, c);
    else op_swap(r, b);
}

static long unguarded_partition(orc_op *f, long first, long last, long pivot)
{
    for (;;) {
        while (op_less(&f[first], &f[pivot])) first++;
        last--;
        while (op_less(&f[pivot], &f[last])) last--;
        if (!(first < last)) return first;
        op_swap(&f[first], &f[last]);
        first++;
    }
}

static void introsort_loop(orc_op *f, long first, long last, long depth_limit)
{
    while (last - first > 16) {
        if (depth_limit == 0) { heap_sort_range(f + first, last - first); return; }
        depth_limit--;
        long mid = first + (last - first) / 2;
        median_to_first(&f[first], &f[first + 1], &f[mid], &f[last - 1]);
        long cut = unguarded_partition(f, first + 1, last, first);
        introsort_loop(f, cut, last, depth_limit);
        last = cut;
    }
}

void orc_sort_ops(orc_op *f, long n)
{
    if (n <= 1) return;
    long lg = 0;
    while ((1L << (lg + 1)) <= n) lg++;
    introsort_loop(f, 0, n, 2 * lg);
    if (n > 16) { ins_sort(f, 16); unguarded_ins_sort(f, 16, n); }
    else ins_sort(f, n);
}

/* Test hook: sort records of (depth, size, e[0..size)) with orc_sort_ops and return the
 * permutation (payload = input position carried in vdepth). */
void orc_sort_ops_test(long n, const uint32_t *depth, const uint32_t *size, const uint32_t *elems,
                       uint32_t *perm_out)
{
    orc_op *ops = (orc_op *)calloc((size_t)(n ? n : 1), sizeof(orc_op));
    for (long i = 0; i < n; i++) {
        ops[i].depth = depth[i]; ops[i].label = 0; ops[i].size = size[i]; ops[i].vdepth = (uint32_t)i;
        for (uint32_t k = 0; k < size[i]; k++) ops[i].e[k] = elems[i * ORC_MAXQ + k];
    }
    orc_sort_ops(ops, n);
    for (long i = 0; i < n; i++) perm_out[i] = ops[i].vdepth;
    free(ops);
}

typedef struct {
    uint32_t n, n_ops;
    uint32_t depth[ORC_MAXOPS], parent[ORC_MAXOPS], vdepth[ORC_MAXOPS], emask[ORC_MAXOPS];
    uint32_t iter_set[ORC_MAXQ];
} orc_plan;

static int op_equal(const orc_op *a, const orc_op *b)
{
    if (a->label != b->label || a->size != b->size) return 0;
    for (uint32_t i = 0; i < a->size; i++) if (a->e[i] != b->e[i]) return 0;
    return 1;
}

/* PlanIR(order, query) (inc/code_gen.h:699-811).  Returns 0 on success, -1 when an
 * iteration set is empty (the order is not connected; the reference exits, :743-746). */
int orc_plan_build(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab,
                   const uint32_t *order, orc_plan *pl)
{
    uint32_t dep_of[ORC_MAXQ];
    orc_op vir[ORC_MAXQ];
    static orc_op all[ORC_MAXOPS], uniq[ORC_MAXOPS];
    orc_op iter_tmp[ORC_MAXQ];
    memset(pl, 0, sizeof(*pl));
    pl->n = n;
    for (uint32_t i = 0; i < n; i++) dep_of[order[i]] = i;
    for (uint32_t q = 0; q < n; q++) {
        /* VertexSetIR per vertex: sorted depths of its neighbours (:713-724) */
        memset(&vir[q], 0, sizeof(orc_op));
        for (uint32_t p = qoff[q]; p < qoff[q + 1]; p++) vir[q].e[vir[q].size++] = dep_of[qadj[p]];
        for (uint32_t a = 1; a < vir[q].size; a++)
            for (uint32_t b = a; b > 0 && vir[q].e[b - 1] > vir[q].e[b]; b--) {
                uint32_t t = vir[q].e[b]; vir[q].e[b] = vir[q].e[b - 1]; vir[q].e[b - 1] = t;
            }
    }
    long na = 0;
    for (uint32_t i = 0; i < ORC_MAXQ; i++) { memset(&iter_tmp[i], 0, sizeof(orc_op)); iter_tmp[i].label = ORC_INVALID; }
    for (uint32_t dep = 0; dep + 1 < n; dep++) {
        /* (:732-748) restrict each follow vertex's set to depths <= dep */
        for (uint32_t fd = dep + 1; fd < n; fd++) {
            const orc_op *src = &vir[order[fd]];
            orc_op o; memset(&o, 0, sizeof(o));
            o.depth = dep; o.vdepth = fd; o.label = lab_of(qlab, order[fd]);
            for (uint32_t k = 0; k < src->size; k++) if (src->e[k] <= dep) o.e[o.size++] = src->e[k];
            if (o.size) all[na++] = o;
        }
        const orc_op *it = &vir[order[dep + 1]];
        orc_op o; memset(&o, 0, sizeof(o));
        o.depth = dep; o.vdepth = dep + 1; o.label = lab_of(qlab, order[dep + 1]);
        for (uint32_t k = 0; k < it->size; k++) if (it->e[k] <= dep) o.e[o.size++] = it->e[k];
        if (o.size == 0) return -1;
        iter_tmp[dep + 1] = o;
    }
    if (na == 0) return -1;
    orc_sort_ops(all, na);
    /* dedup keeping the first occurrence (:752-772); equality ignores depth */
    long nu = 0;
    for (long i = 0; i < na; i++) {
        int first = 1;
        for (long j = 0; j < i; j++) if (op_equal(&all[i], &all[j])) { first = 0; break; }
        if (first) uniq[nu++] = all[i];
    }
    pl->n_ops = (uint32_t)nu;
    for (long i = 0; i < nu; i++) {
        pl->depth[i] = uniq[i].depth;
        pl->vdepth[i] = uniq[i].vdepth;
        pl->parent[i] = ORC_INVALID;
        uint32_t m = 0;
        for (uint32_t k = 0; k < uniq[i].size; k++) m |= 1u << uniq[i].e[k];
        pl->emask[i] = m;
    }
    /* parent = first op whose edge list equals this one minus its last element (:777-787) */
    for (long i = 1; i < nu; i++) {
        orc_op pre = uniq[i];
        pre.size--;
        for (long j = 0; j < i; j++)
            if (op_equal(&pre, &uniq[j])) { pl->parent[i] = (uint32_t)j; break; }
    }
    /* iteration set per depth (:800-808); depth 0 has none */
    for (uint32_t d = 0; d < n; d++) {
        pl->iter_set[d] = ORC_INVALID;
        if (d == 0) continue;
        for (long j = 0; j < nu; j++)
            if (op_equal(&iter_tmp[d], &uniq[j])) { pl->iter_set[d] = (uint32_t)j; break; }
    }
    return 0;
}

/* Flattened plan export for ctypes: arrays of n_ops entries plus iter_set[n]. */
int orc_plan_export(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab,
                    const uint32_t *order, uint32_t *n_ops, uint32_t *depth, uint32_t *parent,
                    uint32_t *vdepth, uint32_t *emask, uint32_t *iter_set)
{
    static orc_plan pl;
    int rc = orc_plan_build(n, qoff, qadj, qlab, order, &pl);
    if (rc) return rc;
    *n_ops = pl.n_ops;
    for (uint32_t i = 0; i < pl.n_ops; i++) {
        depth[i] = pl.depth[i]; parent[i] = pl.parent[i]; vdepth[i] = pl.vdepth[i]; emask[i] = pl.emask[i];
    }
    for (uint32_t d = 0; d < n; d++) iter_set[d] = pl.iter_set[d];
    return 0;
}

/* ------------------------------------------------------------------------------------
 * Enumeration statistics returned to callers.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint64_t count;              /* min(#embeddings, cap) */
    uint64_t intersections;      /* Profiler::total_intersection (inc/common_type.h:185) */
    uint64_t nodes[ORC_MAXQ];    /* extensions per depth (inc/explore.h:555-557 semantics) */
    uint64_t fs_prunes;          /* failing-set sibling cut-offs */
    double enum_ms;              /* thread CPU time of enumeration (inc/explore.h:1093,1126) */
    int32_t timed_out;           /* deadline reached: count is partial and must not be used */
    int32_t capped;              /* the cap stopped the search */
} orc_stats;

static double thread_ms(void)
{
    struct timespec t;
    clock_gettime(CLOCK_THREAD_CPUTIME_ID, &t);
    return t.tv_sec * 1000.0 + t.tv_nsec / 1e6;
}

/* ------------------------------------------------------------------------------------
 * Cache + filter + DAF failing-set enumeration (inc/explore.h:1073-1129
 * explore_Intersection_Cache_DAF_Filter, :968-1071 _explore_recur_cache_daf_filter,
 * :200-243 CacheBuffer::getNextCandidatesImpl(conf_dep), inc/prune.h:82-171
 * DAFFailingPrune).  use_fs = 0 disables the failing-set bookkeeping (cache + filter
 * only), which makes node counts schedule independent.  Labels enter through the plan
 * (ops are keyed by the follow vertex's label) and the candidate sets only: this overload
 * of getNextCandidatesImpl reads whole neighbour lists (data.getNeb, :214), not the
 * label-partitioned ones of the overload without conf_dep (:170-171).
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t V, n;
    const uint32_t *off, *adj, *order;
    const orc_plan *pl;
    const uint8_t *exist;  /* n x V by query vertex (candidates_exsist) */
    uint32_t emb[ORC_MAXQ];
    uint8_t *vis;
    uint32_t *revdep;
    uint32_t *opbuf[ORC_MAXOPS];
    const uint32_t *opptr[ORC_MAXOPS];
    uint32_t opsize[ORC_MAXOPS];
    uint64_t fail[ORC_MAXQ], anc[ORC_MAXQ];
    uint64_t cap;
    int stop, use_fs;
    double deadline;
    uint64_t tick;
    orc_stats *st;
} cache_ctx;

static void rec_cache(cache_ctx *c, uint32_t d)
{
    const orc_plan *pl = c->pl;
    uint32_t u = c->order[d];
    uint32_t vprev = c->emb[c->order[d - 1]];
    int valid = 1;
    uint32_t conf = 0;
    if (c->deadline > 0 && ((++c->tick) & 0xFFF) == 0 && thread_ms() > c->deadline) {
        c->stop = 1; c->st->timed_out = 1; return;
    }
    for (uint32_t o = 0; o < pl->n_ops; o++) {
        if (pl->depth[o] != d - 1) continue;
        const uint32_t *nb = c->adj + c->off[vprev];
        uint32_t nn = dg(c->off, vprev);
        if (pl->parent[o] == ORC_INVALID) {
            c->opptr[o] = nb; c->opsize[o] = nn;
        } else {
            uint32_t p = pl->parent[o];
            c->st->intersections++;
            c->opsize[o] = isect(c->opptr[p], c->opsize[p], nb, nn, c->opbuf[o]);
            c->opptr[o] = c->opbuf[o];
        }
        if (c->opsize[o] == 0) { valid = 0; conf = pl->vdepth[o]; break; }
    }
    uint32_t it = pl->iter_set[d];
    uint32_t tot = c->opsize[it];
    if (!valid) {
        if (c->use_fs) c->fail[d - 1] = c->anc[conf];
        return;
    }
    if (c->use_fs) c->fail[d - 1] = 0;
    uint32_t skip = 0;
    const uint32_t *list = c->opptr[it];
    for (uint32_t pos = 0; pos < tot; pos++) {
        uint32_t v = list[pos];
        if (!c->exist[(size_t)u * c->V + v]) { skip++; continue; }
        if (c->vis[v]) {
            if (c->use_fs) {
                c->fail[d] = c->anc[d] | c->anc[c->revdep[v]];
                c->fail[d - 1] |= c->fail[d];
            }
            goto prune_update;
        }
        c->emb[u] = v;
        c->revdep[v] = d;
        c->vis[v] = 1;
        c->st->nodes[d]++;
        if (d == c->n - 1) {
            c->st->count++;
            if (c->use_fs) { c->fail[d] = ~0ull; c->fail[d - 1] |= ~0ull; }
            if (c->st->count >= c->cap) { c->stop = 1; c->st->capped = 1; return; }
        } else {
            rec_cache(c, d + 1);
            if (c->stop) break;
        }
        c->vis[v] = 0;
    prune_update:
        if (c->use_fs) {
            if (!((c->fail[d] >> d) & 1ull)) {
                c->fail[d - 1] = c->fail[d];
                c->st->fs_prunes++;
                break;
            } else c->fail[d - 1] |= c->fail[d];
        }
    }
    if (c->use_fs && skip == tot) c->fail[d - 1] = c->anc[d];
}

static void daf_ancestors(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, const uint32_t *order,
                          uint64_t *anc)
{
    /* DAFFailingPrune::Init (inc/prune.h:88-117): anc[i] = {i} U anc[j] over backward
     * neighbours j of the vertex at depth i */
    uint32_t dep_of[ORC_MAXQ];
    for (uint32_t i = 0; i < n; i++) dep_of[order[i]] = i;
    for (uint32_t i = 0; i < n; i++) {
        anc[i] = 1ull << i;
        uint32_t q = order[i];
        for (uint32_t p = qoff[q]; p < qoff[q + 1]; p++) {
            uint32_t j = dep_of[qadj[p]];
            if (j < i) anc[i] |= anc[j];
        }
    }
}

/* root_lo / root_hi: only depth-0 candidates v with root_lo <= v < root_hi are searched
 * (root_hi == 0: all).  The reference's depth-0 loop never prunes and every deeper
 * failing-set slot is reset before it is read, so the subtree of each depth-0 candidate is
 * searched alike whatever else runs: counts and statistics over disjoint ranges add up to
 * those of the whole search (a restriction of the reference's loop, not a new search). */
int orc_enum_cache(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n,
                   const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab, const uint32_t *cand,
                   const uint32_t *cnt, uint32_t cand_stride, const uint32_t *order, uint64_t cap,
                   int use_fs, double time_limit_ms, uint32_t root_lo, uint32_t root_hi, orc_stats *st)
{
    memset(st, 0, sizeof(*st));
    if (n < 2 || n > ORC_MAXQ) return -2;
    static orc_plan pl;
    if (orc_plan_build(n, qoff, qadj, qlab, order, &pl)) return -1;
    cache_ctx c;
    memset(&c, 0, sizeof(c));
    c.V = V; c.n = n; c.off = off; c.adj = adj; c.order = order; c.pl = &pl; c.cap = cap;
    c.use_fs = use_fs; c.st = st;
    uint32_t dmax = 1;
    for (uint32_t v = 0; v < V; v++) if (dg(off, v) > dmax) dmax = dg(off, v);
    uint8_t *exist = (uint8_t *)calloc((size_t)n * V, 1);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t j = 0; j < cnt[u]; j++) exist[(size_t)u * V + cand[(size_t)u * cand_stride + j]] = 1;
    c.exist = exist;
    c.vis = (uint8_t *)calloc(V, 1);
    c.revdep = (uint32_t *)malloc(sizeof(uint32_t) * V);
    memset(c.revdep, 0xFF, sizeof(uint32_t) * V);
    for (uint32_t o = 0; o < pl.n_ops; o++)
        c.opbuf[o] = pl.parent[o] == ORC_INVALID ? NULL : (uint32_t *)malloc(sizeof(uint32_t) * dmax);
    daf_ancestors(n, qoff, qadj, order, c.anc);
    double t0 = thread_ms();
    c.deadline = time_limit_ms > 0 ? t0 + time_limit_ms : 0;
    uint32_t u0 = order[0];
    for (uint32_t i = 0; i < cnt[u0]; i++) {
        uint32_t v0 = cand[(size_t)u0 * cand_stride + i];
        if (v0 < root_lo || (root_hi && v0 >= root_hi)) continue;
        c.emb[u0] = v0; c.vis[v0] = 1; c.revdep[v0] = 0;
        st->nodes[0]++;
        rec_cache(&c, 1);
        if (c.stop) break;
        c.vis[v0] = 0;
        /* PruneCheck(0) never cuts (inc/prune.h:120-121) */
    }
    st->enum_ms = thread_ms() - t0;
    for (uint32_t o = 0; o < pl.n_ops; o++) free(c.opbuf[o]);
    free(exist); free(c.vis); free(c.revdep);
    return 0;
}

/* ------------------------------------------------------------------------------------
 * Candidate-space enumeration over the auxiliary tables (inc/explore.h:2757-2848
 * explore_Intersection_All<fs,false,false,false,false>, :2610-2747 _explore_recur_All,
 * :284-354 CandidatesBuffer::getNextCandidatesImpl).  Candidates at depth d are the
 * intersection, over the backward neighbours p of u_d in order position, of the table
 * lists edge_matrix[p][u_d].getNeb(idx(p)); each extra neighbour is one intersection.
 * use_fs: DAF failing sets through PruningOpt (inc/prune.h:2099-2260 on DAFFailingPrune,
 * :82-171): at a child's entry an empty candidate set gives fail[d-1] = anc[d] and ends the
 * child (NoCandidatesPruneCheck), a non-empty one clears fail[d-1]; an injectivity
 * conflict with the vertex mapped at depth c gives fail[d] = anc[d] | anc[c], then
 * fail[d-1] |= fail[d]; a full match gives fail[d] = all, fail[d-1] |= all; after each
 * candidate (BacktrackingPrune) depth d > 0 cuts its remaining candidates when fail[d]
 * lacks d (fail[d-1] = fail[d]), else fail[d-1] |= fail[d].
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t *off, *idx; /* CSR over C(a) of positions in C(b) */
} orc_tab;

typedef struct {
    uint32_t n;
    const uint32_t *cand, *order;
    uint32_t stride;
    uint32_t prev[ORC_MAXQ][ORC_MAXQ], nprev[ORC_MAXQ];
    orc_tab *tab; /* tab[a*n+b] */
    uint32_t *buf[ORC_MAXQ], *tmp;
    uint32_t emb[ORC_MAXQ], iemb[ORC_MAXQ];
    uint8_t *vis;         /* 1 + the depth that maps the vertex, 0 = free */
    uint64_t cap;
    int stop, use_fs;
    uint64_t fail[ORC_MAXQ], anc[ORC_MAXQ];
    double deadline;
    uint64_t tick;
    orc_stats *st;
    /* GuP conflict detection (sgm_oracle_gup.c): 4 words per (depth, position), from goff */
    const uint32_t *guard, *goff;
} all_ctx;

static void rec_all(all_ctx *c, uint32_t d)
{
    uint32_t u = c->order[d];
    if (c->deadline > 0 && ((++c->tick) & 0xFFF) == 0 && thread_ms() > c->deadline) {
        c->stop = 1; c->st->timed_out = 1; return;
    }
    uint32_t p0 = c->prev[u][0];
    orc_tab *t0 = &c->tab[p0 * c->n + u];
    uint32_t i0 = c->iemb[p0];
    uint32_t tot = t0->off[i0 + 1] - t0->off[i0];
    memcpy(c->buf[d], t0->idx + t0->off[i0], sizeof(uint32_t) * tot);
    for (uint32_t k = 1; k < c->nprev[u]; k++) {
        uint32_t p = c->prev[u][k];
        orc_tab *t = &c->tab[p * c->n + u];
        uint32_t ip = c->iemb[p];
        c->st->intersections++;
        uint32_t m = isect(t->idx + t->off[ip], t->off[ip + 1] - t->off[ip], c->buf[d], tot, c->tmp);
        uint32_t *sw = c->buf[d]; c->buf[d] = c->tmp; c->tmp = sw;
        tot = m;
    }
    if (c->use_fs) {  /* NoCandidatesPruneCheck */
        if (tot == 0) { c->fail[d - 1] = c->anc[d]; return; }
        c->fail[d - 1] = 0;
    }
    for (uint32_t pos = 0; pos < tot; pos++) {
        uint32_t vi = c->buf[d][pos];
        uint32_t v = c->cand[(size_t)u * c->stride + vi];
        const uint32_t *g = c->guard ? c->guard + 4 * (size_t)(c->goff[d] + vi) : NULL;
        uint32_t gk = 0;
        if (g && g[0] && !c->vis[v])
            while (gk < g[0] && c->vis[g[1 + gk]]) gk++;
        if (c->vis[v]) {
            if (!c->use_fs) continue;
            c->fail[d] = c->anc[d] | c->anc[c->vis[v] - 1];  /* InjectiveConflict */
            c->fail[d - 1] |= c->fail[d];
        } else if (g && g[0] && gk == g[0]) {
            /* GUPConflictDetection::matchCheck (inc/prune.h:577-595): every vertex of v's
             * guard is matched, so v is pruned; with failing sets ForwardPrune sets fail[d]
             * to all (inc/prune.h:2296-2300) */
            c->st->fs_prunes++;
            if (c->use_fs) { c->fail[d] = ~0ull; c->fail[d - 1] |= ~0ull; }
        } else {
            c->emb[u] = v; c->iemb[u] = vi; c->vis[v] = (uint8_t)(d + 1);
            c->st->nodes[d]++;
            if (d == c->n - 1) {
                c->st->count++;
                if (c->use_fs) { c->fail[d] = ~0ull; c->fail[d - 1] |= ~0ull; }  /* SuccessMatch */
                if (c->st->count >= c->cap) { c->stop = 1; c->st->capped = 1; return; }
            } else {
                rec_all(c, d + 1);
                if (c->stop) return;
            }
            c->vis[v] = 0;
        }
        if (c->use_fs) {  /* BacktrackingPrune -> DAFFailingPrune::PruneCheck(d), d > 0 */
            if (!((c->fail[d] >> d) & 1ull)) {
                c->fail[d - 1] = c->fail[d];
                c->st->fs_prunes++;
                break;
            }
            c->fail[d - 1] |= c->fail[d];
        }
    }
}

/* explore_Intersection_All<fs, false, cd, false, false>: guard/goff from orc_gup_guards for
 * GuP's conflict detection (cd), or null; its prunes are counted in fs_prunes. */
int orc_enum_all_cd(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n,
                    const uint32_t *qoff, const uint32_t *qadj, const uint32_t *cand,
                    const uint32_t *cnt, uint32_t cand_stride, const uint32_t *order, uint64_t cap,
                    int use_fs, double time_limit_ms, uint32_t root_lo, uint32_t root_hi, const uint32_t *goff,
                    const uint32_t *guard, orc_stats *st)
{
    memset(st, 0, sizeof(*st));
    if (n < 2 || n > ORC_MAXQ) return -2;
    all_ctx c;
    memset(&c, 0, sizeof(c));
    c.n = n; c.cand = cand; c.order = order; c.stride = cand_stride; c.cap = cap; c.st = st;
    c.use_fs = use_fs;
    c.guard = guard;
    c.goff = goff;
    daf_ancestors(n, qoff, qadj, order, c.anc);
    uint32_t pos_of[ORC_MAXQ];
    for (uint32_t i = 0; i < n; i++) pos_of[order[i]] = i;
    /* GetPreviousNeb (inc/common_ops.h:259-313): backward neighbours sorted by position */
    for (uint32_t q = 0; q < n; q++) {
        c.nprev[q] = 0;
        for (uint32_t p = qoff[q]; p < qoff[q + 1]; p++)
            if (pos_of[qadj[p]] < pos_of[q]) c.prev[q][c.nprev[q]++] = qadj[p];
        for (uint32_t a = 1; a < c.nprev[q]; a++)
            for (uint32_t b = a; b > 0 && pos_of[c.prev[q][b - 1]] > pos_of[c.prev[q][b]]; b--) {
                uint32_t t = c.prev[q][b]; c.prev[q][b] = c.prev[q][b - 1]; c.prev[q][b - 1] = t;
            }
        if (pos_of[q] > 0 && c.nprev[q] == 0) return -1;
    }
    c.tab = (orc_tab *)calloc((size_t)n * n, sizeof(orc_tab));
    int32_t *scratch = (int32_t *)malloc(sizeof(int32_t) * (V ? V : 1));
    uint32_t maxc = 1;
    for (uint32_t q = 0; q < n; q++) if (cnt[q] > maxc) maxc = cnt[q];
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = qoff[a]; p < qoff[a + 1]; p++) {
            uint32_t b = qadj[p];
            orc_tab *t = &c.tab[a * n + b];
            const uint32_t *ca = cand + (size_t)a * cand_stride, *cb = cand + (size_t)b * cand_stride;
            uint64_t e = orc_table(V, off, adj, ca, cnt[a], cb, cnt[b], NULL, NULL, NULL, NULL, scratch);
            t->off = (uint32_t *)malloc(sizeof(uint32_t) * (cnt[a] + 1));
            t->idx = (uint32_t *)malloc(sizeof(uint32_t) * (e ? e : 1));
            uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * (e ? e : 1));
            orc_table(V, off, adj, ca, cnt[a], cb, cnt[b], t->off, t->idx, tv, NULL, scratch);
            free(tv);
        }
    for (uint32_t d = 0; d < n; d++) c.buf[d] = (uint32_t *)malloc(sizeof(uint32_t) * maxc);
    c.tmp = (uint32_t *)malloc(sizeof(uint32_t) * maxc);
    c.vis = (uint8_t *)calloc(V, 1);
    double t0 = thread_ms();
    c.deadline = time_limit_ms > 0 ? t0 + time_limit_ms : 0;
    uint32_t u0 = order[0];
    for (uint32_t i = 0; i < cnt[u0]; i++) {
        uint32_t v0 = cand[(size_t)u0 * cand_stride + i];
        if (v0 < root_lo || (root_hi && v0 >= root_hi)) continue;  /* (see orc_enum_cache) */
        c.emb[u0] = v0; c.iemb[u0] = i; c.vis[v0] = 1;  /* depth 0; depth 0 never prunes */
        st->nodes[0]++;
        rec_all(&c, 1);
        if (c.stop) break;
        c.vis[v0] = 0;
    }
    st->enum_ms = thread_ms() - t0;
    for (uint32_t a = 0; a < n * n; a++) { free(c.tab[a].off); free(c.tab[a].idx); }
    free(c.tab); free(scratch);
    for (uint32_t d = 0; d < n; d++) free(c.buf[d]);
    free(c.tmp); free(c.vis);
    return 0;
}

int orc_enum_all(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n,
                 const uint32_t *qoff, const uint32_t *qadj, const uint32_t *cand,
                 const uint32_t *cnt, uint32_t cand_stride, const uint32_t *order, uint64_t cap,
                 int use_fs, double time_limit_ms, uint32_t root_lo, uint32_t root_hi, orc_stats *st)
{
    return orc_enum_all_cd(V, off, adj, n, qoff, qadj, cand, cnt, cand_stride, order, cap, use_fs, time_limit_ms,
                           root_lo, root_hi, NULL, NULL, st);
}

uint32_t orc_abi_version(void) { return 4u; }
uint32_t orc_stats_size(void) { return (uint32_t)sizeof(orc_stats); }
