/*
 * sgm_oracle_tree.c -- CPU restatement of the reference's BFS-tree filters (CFL, CECI,
 * DAF) and of the orders built on the BFS tree or on the auxiliary edge tables (QSI, CFL,
 * CECI, GuP), with the query k-core and BFS helpers they share.
 *
 * TEST INFRASTRUCTURE ONLY.  Like sgm_oracle.c this file is the parity checker, linked into
 * liboracle.so and loaded by tests/ (never by the product library).  Every function cites
 * the reference file:line it restates (paths relative to the Orrinn/SubgraphMatching tree).
 *
 * The filters are restated in the reference's own shape: candidate arrays with INVALID
 * holes, the flag-counting neighbour scans of generateCandidates / pruneCandidates, the
 * frontier invalidation of CECI, then compaction and an ascending sort (FilterParamConstruct).
 *
 * Pinning (tests/test_oracle_pins.py):
 *   - orc_kcore:        core numbers from their definition (repeated peeling for each k)
 *   - orc_bfs:          hand-worked trees; children in adjacency order; levels = distances
 *   - filters:          an independent set-level restatement of each schedule in Python
 *                       (membership sets, "v has a neighbour in C(p)" tests), CFL/CECI/DAF
 *                       subset-of NLF, soundness (every vertex of a brute-force embedding
 *                       survives), and hand-worked cases in which the schedule matters
 *   - the tie order of CFLStartVertex's priority queue: std::priority_queue itself on
 *     tie-heavy inputs (oracle/sort_probe.cpp)
 *   - orders:           hand-worked known answers for every rule (QSI edge ties, CFL core /
 *                       tree paths and leaves, GuP core weights and tree phase), permutation
 *                       and connectivity of every order on random queries
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define T_INVALID 0xFFFFFFFFu
#define T_MAXQ 32

static inline uint32_t tdeg(const uint32_t *off, uint32_t v) { return off[v + 1] - off[v]; }
static inline uint32_t tlab(const uint32_t *lab, uint32_t v) { return lab ? lab[v] : 0u; }

/* ------------------------------------------------------------------------------------
 * Query k-core (inc/graph.h:387-458 Query::KCore, :460-470 buildKCore): the bin-sorted
 * peeling of Batagelj and Zaversnik.  core[v] = the largest k such that v lies in the
 * k-core.  Returns the number of vertices with core > 1 (Query::getKCoreLength).
 * ---------------------------------------------------------------------------------- */
uint32_t orc_kcore(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, int32_t *core)
{
    int32_t vert[T_MAXQ], pos[T_MAXQ], bin[T_MAXQ + 1], start[T_MAXQ + 1];
    int32_t md = 0;
    for (uint32_t v = 0; v < n; v++) if ((int32_t)tdeg(qoff, v) > md) md = (int32_t)tdeg(qoff, v);
    for (int32_t d = 0; d <= md; d++) bin[d] = 0;
    for (uint32_t v = 0; v < n; v++) { core[v] = (int32_t)tdeg(qoff, v); bin[core[v]]++; }
    int32_t s = 0;
    for (int32_t d = 0; d <= md; d++) { start[d] = s; s += bin[d]; }
    for (uint32_t v = 0; v < n; v++) { pos[v] = start[core[v]]; vert[pos[v]] = (int32_t)v; start[core[v]]++; }
    for (int32_t d = md; d > 0; d--) start[d] = start[d - 1];
    start[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        int32_t v = vert[i];
        for (uint32_t p = qoff[v]; p < qoff[v + 1]; p++) {
            int32_t u = (int32_t)qadj[p];
            if (core[u] > core[v]) {
                int32_t du = core[u], pu = pos[u], pw = start[du], w = vert[pw];
                if (u != w) { pos[u] = pw; pos[w] = pu; vert[pu] = w; vert[pw] = u; }
                start[du]++;
                core[u]--;
            }
        }
    }
    uint32_t len = 0;
    for (uint32_t v = 0; v < n; v++) len += core[v] > 1;
    return len;
}

/* ------------------------------------------------------------------------------------
 * BFS tree of the query (inc/common_ops.h:318-350 BFS; TreeNode inc/graph.h:497-519):
 * children in adjacency order, level = BFS distance, parent of the start vertex 0 (the
 * TreeNode default).  Returns the number of vertices reached.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t order[T_MAXQ], parent[T_MAXQ], level[T_MAXQ], nch[T_MAXQ], ch[T_MAXQ][T_MAXQ];
    uint32_t cnt;
} tbfs;

static void bfs_tree(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, uint32_t s, tbfs *t)
{
    uint8_t seen[T_MAXQ] = {0};
    memset(t, 0, sizeof *t);
    uint32_t head = 0;
    t->order[t->cnt++] = s;
    seen[s] = 1;
    while (head < t->cnt) {
        uint32_t u = t->order[head++];
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
            uint32_t w = qadj[p];
            if (seen[w]) continue;
            seen[w] = 1;
            t->order[t->cnt++] = w;
            t->parent[w] = u;
            t->level[w] = t->level[u] + 1;
            t->ch[u][t->nch[u]++] = w;
        }
    }
}

uint32_t orc_bfs(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, uint32_t start, uint32_t *order,
                 uint32_t *parent, uint32_t *level)
{
    tbfs t;
    bfs_tree(n, qoff, qadj, start, &t);
    for (uint32_t i = 0; i < t.cnt; i++) order[i] = t.order[i];
    for (uint32_t v = 0; v < n; v++) { parent[v] = t.parent[v]; level[v] = t.level[v]; }
    return t.cnt;
}

/* Previous / following neighbours of each vertex in order position (inc/common_ops.h:
 * 259-316 _GetNeb, GetPreviousNeb, GetFollowingNeb): neighbours earlier (later) in the
 * order, sorted by position. */
static void order_nebs(uint32_t n, const uint32_t *qoff, const uint32_t *qadj, const uint32_t *order, int following,
                       uint32_t nb[T_MAXQ][T_MAXQ], uint32_t *cnt)
{
    uint32_t pos[T_MAXQ];
    for (uint32_t i = 0; i < n; i++) pos[order[i]] = i;
    for (uint32_t v = 0; v < n; v++) {
        cnt[v] = 0;
        /* by position: walk the positions in order and keep the neighbours of v */
        for (uint32_t i = 0; i < n; i++) {
            uint32_t w = order[i];
            int adj = 0;
            for (uint32_t p = qoff[v]; p < qoff[v + 1]; p++) adj |= qadj[p] == w;
            if (!adj) continue;
            if (following ? i > pos[v] : i < pos[v]) nb[v][cnt[v]++] = w;
        }
    }
}

/* ------------------------------------------------------------------------------------
 * NLF of one query vertex (inc/filter.h:24-58 NLFByVertex): label(v) = label(u),
 * deg(v) >= deg_q(u), and v's neighbour-label counts cover u's.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t V;
    const uint32_t *off, *adj, *dlab;
    uint32_t n;
    const uint32_t *qoff, *qadj, *qlab;
    uint32_t np[T_MAXQ], pl[T_MAXQ][T_MAXQ], pc[T_MAXQ][T_MAXQ];
} tin;

static void tin_init(tin *I, uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab, uint32_t n,
                     const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab)
{
    I->V = V; I->off = off; I->adj = adj; I->dlab = dlab;
    I->n = n; I->qoff = qoff; I->qadj = qadj; I->qlab = qlab;
    for (uint32_t u = 0; u < n; u++) {
        I->np[u] = 0;
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++) {
            uint32_t l = tlab(qlab, qadj[p]), k = 0;
            while (k < I->np[u] && I->pl[u][k] != l) k++;
            if (k == I->np[u]) { I->pl[u][k] = l; I->pc[u][k] = 0; I->np[u]++; }
            I->pc[u][k]++;
        }
    }
}

static int nlf_ok(const tin *I, uint32_t u, uint32_t v)
{
    if (tlab(I->dlab, v) != tlab(I->qlab, u) || tdeg(I->off, v) < tdeg(I->qoff, u)) return 0;
    if (!I->dlab && !I->qlab) return 1; /* one label: the count test is the degree test */
    for (uint32_t k = 0; k < I->np[u]; k++) {
        uint32_t have = 0;
        for (uint32_t p = I->off[v]; p < I->off[v + 1]; p++) have += tlab(I->dlab, I->adj[p]) == I->pl[u][k];
        if (have < I->pc[u][k]) return 0;
    }
    return 1;
}

static uint32_t nlf_vertex(const tin *I, uint32_t u, uint32_t *out)
{
    uint32_t c = 0;
    for (uint32_t v = 0; v < I->V; v++)
        if (nlf_ok(I, u, v)) { if (out) out[c] = v; c++; }
    return c;
}

/* ------------------------------------------------------------------------------------
 * Candidate arrays: cand[u*V ...], cnt[u]; INVALID marks a removed entry until the final
 * compaction.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    tin *I;
    uint32_t *cand, *cnt;
    uint32_t *flag, *upd;
} tstate;

static uint32_t *C_(tstate *S, uint32_t u) { return S->cand + (size_t)u * S->I->V; }

/* generateCandidates (inc/filter.h:317-378): C(u) = the vertices of u's label and degree
 * adjacent to a valid candidate of every pivot (flag counting over the pivots in order),
 * then the NLF test, in the order of first discovery. */
static void generate(tstate *S, uint32_t u, const uint32_t *piv, uint32_t npiv)
{
    const tin *I = S->I;
    uint32_t count = 0, nupd = 0, lu = tlab(I->qlab, u), du = tdeg(I->qoff, u);
    for (uint32_t k = 0; k < npiv; k++) {
        uint32_t p = piv[k], *cp = C_(S, p);
        for (uint32_t j = 0; j < S->cnt[p]; j++) {
            uint32_t v = cp[j];
            if (v == T_INVALID) continue;
            for (uint32_t e = I->off[v]; e < I->off[v + 1]; e++) {
                uint32_t w = I->adj[e];
                if (S->flag[w] == count && tlab(I->dlab, w) == lu && tdeg(I->off, w) >= du) {
                    S->flag[w]++;
                    if (count == 0) S->upd[nupd++] = w;
                }
            }
        }
        count++;
    }
    uint32_t *cu = C_(S, u);
    for (uint32_t i = 0; i < nupd; i++) {
        uint32_t v = S->upd[i];
        if (S->flag[v] == count && nlf_ok(I, u, v)) cu[S->cnt[u]++] = v;
    }
    for (uint32_t i = 0; i < nupd; i++) S->flag[S->upd[i]] = 0;
}

/* pruneCandidates (inc/filter.h:380-425): a candidate of u survives iff it is adjacent to a
 * valid candidate of every pivot (the same flag counting). */
static void prune(tstate *S, uint32_t u, const uint32_t *piv, uint32_t npiv)
{
    const tin *I = S->I;
    uint32_t count = 0, nupd = 0, lu = tlab(I->qlab, u), du = tdeg(I->qoff, u);
    for (uint32_t k = 0; k < npiv; k++) {
        uint32_t p = piv[k], *cp = C_(S, p);
        for (uint32_t j = 0; j < S->cnt[p]; j++) {
            uint32_t v = cp[j];
            if (v == T_INVALID) continue;
            for (uint32_t e = I->off[v]; e < I->off[v + 1]; e++) {
                uint32_t w = I->adj[e];
                if (S->flag[w] == count && tlab(I->dlab, w) == lu && tdeg(I->off, w) >= du) {
                    S->flag[w]++;
                    if (count == 0) S->upd[nupd++] = w;
                }
            }
        }
        count++;
    }
    uint32_t *cu = C_(S, u);
    for (uint32_t i = 0; i < S->cnt[u]; i++)
        if (cu[i] != T_INVALID && S->flag[cu[i]] != count) cu[i] = T_INVALID;
    for (uint32_t i = 0; i < nupd; i++) S->flag[S->upd[i]] = 0;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* compactCandidates (inc/filter.h:206-220), sortCandidates (:18-22, via FilterParamConstruct
 * :60-78), isCandidateSetValid (:222-230). */
static int finish(tstate *S)
{
    int ok = 1;
    for (uint32_t u = 0; u < S->I->n; u++) {
        uint32_t *cu = C_(S, u), k = 0;
        for (uint32_t i = 0; i < S->cnt[u]; i++) if (cu[i] != T_INVALID) cu[k++] = cu[i];
        S->cnt[u] = k;
        qsort(cu, k, sizeof(uint32_t), cmp_u32);
        ok &= k > 0;
    }
    return ok;
}

/* ------------------------------------------------------------------------------------
 * std::priority_queue<pair<VertexID,double>, vector<...>, cmp> with cmp(l, r) = l.second <
 * r.second (a max-heap on the second member), as libstdc++ implements push (push_back +
 * sift-up) and pop (the last element moved into the hole left by sifting the root's hole
 * down to a leaf, always towards the larger child, then sifted up).  Ties between equal
 * keys are resolved by exactly these moves, and CFLStartVertex depends on them.
 * ---------------------------------------------------------------------------------- */
typedef struct { uint32_t v; double key; } theap_e;
typedef struct { theap_e a[T_MAXQ + 1]; long len; } theap;

static int hless(const theap_e *x, const theap_e *y) { return x->key < y->key; }

static void h_sift_up(theap_e *f, long hole, long top, theap_e val)
{
    long parent = (hole - 1) / 2;
    while (hole > top && hless(&f[parent], &val)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = val;
}

static void h_push(theap *h, theap_e x)
{
    h->a[h->len] = x;
    h->len++;
    h_sift_up(h->a, h->len - 1, 0, h->a[h->len - 1]);
}

static void h_pop(theap *h)
{
    if (h->len > 1) {
        theap_e *f = h->a;
        long len = h->len - 1, hole = 0, second = 0;
        theap_e val = f[len];
        f[len] = f[0];
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            if (hless(&f[second], &f[second - 1])) second--;
            f[hole] = f[second];
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            f[hole] = f[second - 1];
            hole = second - 1;
        }
        h_sift_up(f, hole, 0, val);
    }
    h->len--;
}

/* Test hook: push (v, keys[v]) for v < count, pop while more than keep remain, then pop
 * the rest into out[] (the order CFLStartVertex visits them). */
void orc_heap_test(uint32_t count, const double *keys, uint32_t keep, uint32_t *out)
{
    theap h;
    h.len = 0;
    for (uint32_t v = 0; v < count; v++) { theap_e e = {v, keys[v]}; h_push(&h, e); }
    while (h.len > (long)keep) h_pop(&h);
    uint32_t k = 0;
    while (h.len > 0) { out[k++] = h.a[0].v; h_pop(&h); }
}

/* ------------------------------------------------------------------------------------
 * Start vertices.
 * ---------------------------------------------------------------------------------- */
static void label_freq(const tin *I, uint32_t l, uint32_t *freq, uint32_t *maxfreq)
{
    /* Graph::getLabelFreq / getMaxLabelFreq (inc/graph.h:339-345) */
    uint32_t f = 0, mf = 0;
    if (!I->dlab) { f = l == 0 ? I->V : 0; mf = I->V; }
    else {
        for (uint32_t v = 0; v < I->V; v++) f += I->dlab[v] == l;
        /* the largest label class */
        uint32_t *lab = (uint32_t *)malloc(sizeof(uint32_t) * (I->V ? I->V : 1));
        memcpy(lab, I->dlab, sizeof(uint32_t) * I->V);
        qsort(lab, I->V, sizeof(uint32_t), cmp_u32);
        for (uint32_t i = 0; i < I->V;) {
            uint32_t j = i;
            while (j < I->V && lab[j] == lab[i]) j++;
            if (j - i > mf) mf = j - i;
            i = j;
        }
        free(lab);
    }
    if (freq) *freq = f;
    if (maxfreq) *maxfreq = mf;
}

/* CFLStartVertex (inc/filter.h:437-478): the three query vertices of smallest
 * label-frequency / degree among those of core value > 1 (all if the query has no 2-core),
 * kept by the max-heap above; then, in the heap's pop order, the first of minimum
 * NLF-count / degree. */
static uint32_t cfl_start(const tin *I)
{
    int32_t core[T_MAXQ];
    uint32_t klen = orc_kcore(I->n, I->qoff, I->qadj, core), maxf = 0;
    theap h;
    h.len = 0;
    for (uint32_t u = 0; u < I->n; u++) {
        if (klen == 0 || core[u] > 1) {
            uint32_t f;
            label_freq(I, tlab(I->qlab, u), &f, NULL);
            theap_e e = {u, f / (double)tdeg(I->qoff, u)};
            h_push(&h, e);
        }
    }
    while (h.len > 3) h_pop(&h);
    label_freq(I, 0, NULL, &maxf);
    uint32_t start = 0;
    double min_score = (double)(uint32_t)(maxf + 1u);
    while (h.len > 0) {
        uint32_t u = h.a[0].v;
        double cur = nlf_vertex(I, u, NULL) / (double)tdeg(I->qoff, u);
        if (cur < min_score) { start = u; min_score = cur; }
        h_pop(&h);
    }
    return start;
}

/* CECIStartVertex (inc/filter.h:585-600) and DAFStartVertex (:922-940): minimum NLF-count /
 * degree, first on ties; DAF skips vertices of degree <= 1. */
static uint32_t ceci_daf_start(const tin *I, int daf)
{
    double min_score = I->V;
    uint32_t start = 0;
    for (uint32_t u = 0; u < I->n; u++) {
        uint32_t d = tdeg(I->qoff, u);
        if (daf && d <= 1) continue;
        double cur = nlf_vertex(I, u, NULL) / (double)d;
        if (cur < min_score) { min_score = cur; start = u; }
    }
    return start;
}

/* kind: 3 CFL, 4 CECI, 5 DAF (the C ABI's SGM_FILTER_* numbers). */
uint32_t orc_start_vertex(int kind, uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab,
                          uint32_t n, const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab)
{
    tin I;
    tin_init(&I, V, off, adj, dlab, n, qoff, qadj, qlab);
    return kind == 3 ? cfl_start(&I) : ceci_daf_start(&I, kind == 5);
}

/* ------------------------------------------------------------------------------------
 * _CFL_Filter (inc/filter.h:481-583): BFS from CFLStartVertex; C(start) = NLF; level by
 * level, generate each vertex from its previous neighbours, then prune the level in
 * reverse order by following neighbours on the same level; then bottom-up, prune each
 * vertex by its following neighbours on deeper levels.
 * ---------------------------------------------------------------------------------- */
static int cfl_filter(tstate *S)
{
    const tin *I = S->I;
    const uint32_t n = I->n;
    tbfs t;
    bfs_tree(n, I->qoff, I->qadj, cfl_start(I), &t);
    if (t.cnt != n) return 0;
    uint32_t prev[T_MAXQ][T_MAXQ], pc[T_MAXQ], fol[T_MAXQ][T_MAXQ], fc[T_MAXQ];
    order_nebs(n, I->qoff, I->qadj, t.order, 0, prev, pc);
    order_nebs(n, I->qoff, I->qadj, t.order, 1, fol, fc);
    uint32_t same[T_MAXQ][T_MAXQ], sc[T_MAXQ], diff[T_MAXQ][T_MAXQ], dc[T_MAXQ], maxl = 0;
    for (uint32_t v = 0; v < n; v++) {
        if (t.level[v] > maxl) maxl = t.level[v];
        sc[v] = dc[v] = 0;
        for (uint32_t k = 0; k < fc[v]; k++) {
            uint32_t w = fol[v][k];
            if (t.level[w] == t.level[v]) same[v][sc[v]++] = w;
            else if (t.level[w] > t.level[v]) diff[v][dc[v]++] = w;
        }
    }
    const uint32_t levels = maxl + 1;
    uint32_t at[T_MAXQ][T_MAXQ], ac[T_MAXQ] = {0};
    for (uint32_t i = 0; i < n; i++) { uint32_t v = t.order[i]; at[t.level[v]][ac[t.level[v]]++] = v; }
    const uint32_t s = t.order[0];
    S->cnt[s] = nlf_vertex(I, s, C_(S, s));
    for (uint32_t l = 1; l < levels; l++) {
        for (uint32_t k = 0; k < ac[l]; k++) generate(S, at[l][k], prev[at[l][k]], pc[at[l][k]]);
        for (uint32_t k = ac[l]; k-- > 0;) {
            uint32_t u = at[l][k];
            if (sc[u]) prune(S, u, same[u], sc[u]);
        }
    }
    for (int l = (int)levels - 2; l >= 0; l--)
        for (uint32_t k = 0; k < ac[l]; k++) {
            uint32_t u = at[l][k];
            if (dc[u]) prune(S, u, diff[u], dc[u]);
        }
    return finish(S);
}

/* ------------------------------------------------------------------------------------
 * _CECI_Filter (inc/filter.h:602-920).  Candidate sets come from the TE and NTE
 * construction alone: C(root) = NLF; for each later vertex u in BFS order, C(u) = the NLF
 * candidates of u adjacent to a valid candidate of u's tree parent, and a parent candidate
 * with no NLF(u) neighbour becomes invalid (:644-708); then, for every vertex u and every
 * previous neighbour u_p, a candidate of u_p with no NLF(u) neighbour becomes invalid
 * (:716-788).  The reverse-BFS refinement (:790-883) and the final clean-up (:895-917)
 * update only the function-local TE / NTE maps and cardinalities and write no candidate,
 * so they are not restated.  Where the reference ends the process on an empty C(u)
 * (:700-703) this returns 0.
 * ---------------------------------------------------------------------------------- */
static int ceci_filter(tstate *S)
{
    const tin *I = S->I;
    const uint32_t n = I->n;
    tbfs t;
    bfs_tree(n, I->qoff, I->qadj, ceci_daf_start(I, 0), &t);
    if (t.cnt != n) return 0;
    uint32_t prev[T_MAXQ][T_MAXQ], pc[T_MAXQ];
    order_nebs(n, I->qoff, I->qadj, t.order, 0, prev, pc);
    const uint32_t root = t.order[0];
    S->cnt[root] = nlf_vertex(I, root, C_(S, root));
    if (S->cnt[root] == 0) return 0;
    for (uint32_t i = 1; i < n; i++) {
        uint32_t u = t.order[i], up = t.parent[u], *cu = C_(S, u), *fr = C_(S, up);
        S->cnt[u] = 0;
        for (uint32_t j = 0; j < S->cnt[up]; j++) {
            uint32_t vf = fr[j];
            if (vf == T_INVALID) continue;
            int any = 0;
            for (uint32_t e = I->off[vf]; e < I->off[vf + 1]; e++) {
                uint32_t v = I->adj[e];
                if (!nlf_ok(I, u, v)) continue;
                any = 1;
                if (S->flag[v] == 0) { S->flag[v] = 1; cu[S->cnt[u]++] = v; }
            }
            if (!any) fr[j] = T_INVALID;
        }
        if (S->cnt[u] == 0) return 0;
        for (uint32_t j = 0; j < S->cnt[u]; j++) S->flag[cu[j]] = 0;
    }
    for (uint32_t i = 1; i < n; i++) {
        uint32_t u = t.order[i];
        for (uint32_t k = 0; k < pc[u]; k++) {
            uint32_t up = prev[u][k], *fr = C_(S, up);
            for (uint32_t j = 0; j < S->cnt[up]; j++) {
                uint32_t vf = fr[j];
                if (vf == T_INVALID) continue;
                int any = 0;
                for (uint32_t e = I->off[vf]; e < I->off[vf + 1] && !any; e++) any = nlf_ok(I, u, I->adj[e]);
                if (!any) fr[j] = T_INVALID;
            }
        }
    }
    return finish(S);
}

/* ------------------------------------------------------------------------------------
 * _DAF_Filter (inc/filter.h:942-1020): NLF (false at the first empty set); BFS from
 * DAFStartVertex; three refinements: forward over BFS order pruning by previous
 * neighbours, backward by following neighbours, forward again.
 * ---------------------------------------------------------------------------------- */
static int daf_filter(tstate *S)
{
    const tin *I = S->I;
    const uint32_t n = I->n;
    for (uint32_t u = 0; u < n; u++) {
        S->cnt[u] = nlf_vertex(I, u, C_(S, u));
        if (S->cnt[u] == 0) return 0;
    }
    tbfs t;
    bfs_tree(n, I->qoff, I->qadj, ceci_daf_start(I, 1), &t);
    if (t.cnt != n) return 0;
    uint32_t prev[T_MAXQ][T_MAXQ], pc[T_MAXQ], fol[T_MAXQ][T_MAXQ], fc[T_MAXQ];
    order_nebs(n, I->qoff, I->qadj, t.order, 0, prev, pc);
    order_nebs(n, I->qoff, I->qadj, t.order, 1, fol, fc);
    for (uint32_t k = 0; k < 3; k++) {
        if (k % 2 == 0) {
            for (uint32_t i = 1; i < n; i++) { uint32_t u = t.order[i]; prune(S, u, prev[u], pc[u]); }
        } else {
            for (int i = (int)n - 2; i >= 0; i--) { uint32_t u = t.order[i]; prune(S, u, fol[u], fc[u]); }
        }
    }
    return finish(S);
}

/* kind: 3 CFL, 4 CECI, 5 DAF.  cand[n*V] receives each C(u) ascending, cnt[n] the sizes.
 * Returns 1 (every set non-empty), 0 (the reference returns false or exits), -1 (bad kind
 * or a disconnected query). */
int orc_filter_tree(int kind, uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab, uint32_t n,
                    const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab, uint32_t *cand, uint32_t *cnt)
{
    if (n == 0 || n > T_MAXQ || kind < 3 || kind > 5) return -1;
    tin I;
    tin_init(&I, V, off, adj, dlab, n, qoff, qadj, qlab);
    tstate S;
    S.I = &I;
    S.cand = cand;
    S.cnt = cnt;
    memset(cnt, 0, sizeof(uint32_t) * n);
    S.flag = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    S.upd = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    int ok = kind == 3 ? cfl_filter(&S) : kind == 4 ? ceci_filter(&S) : daf_filter(&S);
    free(S.flag);
    free(S.upd);
    return ok;
}

/* ------------------------------------------------------------------------------------
 * Auxiliary edge tables of the orders, built from the candidate sets (the content of
 * buildTables, inc/table.h:7-137): row i of (a, b) lists the positions k in C(b) with
 * (C(a)[i], C(b)[k]) an edge, ascending.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t *off[T_MAXQ][T_MAXQ], *idx[T_MAXQ][T_MAXQ];
    uint32_t ecnt[T_MAXQ][T_MAXQ]; /* Edges::_e_cnt is a u32 (inc/graph.h:19) */
} ttables;

static void tables_build(const tin *I, const uint32_t *cand, const uint32_t *cnt, uint32_t stride, ttables *T)
{
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (I->V ? I->V : 1));
    memset(T, 0, sizeof *T);
    for (uint32_t a = 0; a < I->n; a++)
        for (uint32_t p = I->qoff[a]; p < I->qoff[a + 1]; p++) {
            uint32_t b = I->qadj[p];
            const uint32_t *ca = cand + (size_t)a * stride, *cb = cand + (size_t)b * stride;
            for (uint32_t v = 0; v < I->V; v++) pos[v] = -1;
            for (uint32_t k = 0; k < cnt[b]; k++) pos[cb[k]] = (int32_t)k;
            uint64_t e = 0;
            for (uint32_t i = 0; i < cnt[a]; i++)
                for (uint32_t q = I->off[ca[i]]; q < I->off[ca[i] + 1]; q++) e += pos[I->adj[q]] >= 0;
            uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * (cnt[a] + 1));
            uint32_t *ti = (uint32_t *)malloc(sizeof(uint32_t) * (e ? e : 1));
            e = 0;
            for (uint32_t i = 0; i < cnt[a]; i++) {
                to[i] = (uint32_t)e;
                for (uint32_t q = I->off[ca[i]]; q < I->off[ca[i] + 1]; q++)
                    if (pos[I->adj[q]] >= 0) ti[e++] = (uint32_t)pos[I->adj[q]];
            }
            to[cnt[a]] = (uint32_t)e;
            T->off[a][b] = to;
            T->idx[a][b] = ti;
            T->ecnt[a][b] = (uint32_t)e;
        }
    free(pos);
}

static void tables_free(ttables *T)
{
    for (uint32_t a = 0; a < T_MAXQ; a++)
        for (uint32_t b = 0; b < T_MAXQ; b++) { free(T->off[a][b]); free(T->idx[a][b]); }
}

/* ------------------------------------------------------------------------------------
 * Order_QSI (inc/order.h:610-668) with QSIStartEdge (:146-179): the start edge has the
 * fewest table entries (tie: smaller degree sum), taken from its higher-degree end (the
 * first end on equal degrees); then repeatedly the unvisited end of an edge from the
 * visited set with the fewest entries, ties by fewer visited neighbours, then larger degree.
 * ---------------------------------------------------------------------------------- */
static void order_qsi(const tin *I, const ttables *T, uint32_t *order)
{
    const uint32_t n = I->n;
    uint8_t vis[T_MAXQ] = {0};
    uint32_t minv = 0xFFFFFFFFu, mind = 0xFFFFFFFFu, s0 = 0, s1 = 1;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t p = I->qoff[i]; p < I->qoff[i + 1]; p++) {
            uint32_t e = I->qadj[p], cur = T->ecnt[i][e], ds = tdeg(I->qoff, i) + tdeg(I->qoff, e);
            if (cur < minv || (cur == minv && ds < mind)) {
                minv = cur;
                mind = ds;
                if (tdeg(I->qoff, i) < tdeg(I->qoff, e)) { s0 = e; s1 = i; } else { s0 = i; s1 = e; }
            }
        }
    order[0] = s0;
    order[1] = s1;
    vis[s0] = vis[s1] = 1;
    for (uint32_t l = 2; l < n; l++) {
        uint32_t mv = 0xFFFFFFFFu, maxd = 0, maxadj = 0, sel = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!vis[i]) continue;
            for (uint32_t p = I->qoff[i]; p < I->qoff[i + 1]; p++) {
                uint32_t e = I->qadj[p];
                if (vis[e]) continue;
                uint32_t cur = T->ecnt[i][e], cd = tdeg(I->qoff, e), adjs = 0;
                for (uint32_t q = I->qoff[e]; q < I->qoff[e + 1]; q++) adjs += vis[I->qadj[q]];
                if (cur < mv || (cur == mv && adjs < maxadj) || (cur == mv && adjs == maxadj && cd > maxd)) {
                    sel = e;
                    mv = cur;
                    maxd = cd;
                    maxadj = adjs;
                }
            }
        }
        order[l] = sel;
        vis[sel] = 1;
    }
}

/* ------------------------------------------------------------------------------------
 * Order_CFL (inc/order.h:781-976).  Root = CFLStartVertex; core paths (:252-269) when the
 * root has core > 1, tree paths (:271-289) from every core vertex (or the root); the path
 * with the smallest estimated embeddings per non-tree edge first, then the paths by
 * estimated embeddings per candidate of their attachment vertex, then the leaves by
 * candidate count.  Embedding estimates (estimatePathEmbeddsingsNum, :197-239) multiply
 * along the path's tables with size_t arithmetic (wrapping mod 2^64).
 * ---------------------------------------------------------------------------------- */
typedef struct { uint32_t len; uint32_t v[T_MAXQ]; } tpath;

static void core_paths_rec(const int32_t *core, const tbfs *t, uint32_t u, tpath *cur, tpath *out, uint32_t *nout)
{
    cur->v[cur->len++] = u;
    int leaf = 1;
    for (uint32_t k = 0; k < t->nch[u]; k++)
        if (core[t->ch[u][k]] > 1) { core_paths_rec(core, t, t->ch[u][k], cur, out, nout); leaf = 0; }
    if (leaf) out[(*nout)++] = *cur;
    cur->len--;
}

static void tree_paths_rec(const tin *I, const tbfs *t, uint32_t u, tpath *cur, tpath *out, uint32_t *nout)
{
    cur->v[cur->len++] = u;
    int leaf = 1;
    for (uint32_t k = 0; k < t->nch[u]; k++)
        if (tdeg(I->qoff, t->ch[u][k]) > 1) { tree_paths_rec(I, t, t->ch[u][k], cur, out, nout); leaf = 0; }
    if (leaf && cur->len > 1) out[(*nout)++] = *cur;
    cur->len--;
}

static void estimate_path(const ttables *T, const uint32_t *cnt, const tpath *p, uint64_t *est)
{
    const uint32_t L = p->len;
    uint32_t a = p->v[L - 2], b = p->v[L - 1];
    uint64_t *child = (uint64_t *)malloc(sizeof(uint64_t) * (cnt[a] ? cnt[a] : 1)), sum = 0;
    for (uint32_t i = 0; i < cnt[a]; i++) { child[i] = T->off[a][b][i + 1] - T->off[a][b][i]; sum += child[i]; }
    est[L - 2] = sum;
    for (int i = (int)L - 2; i >= 1; i--) {
        uint32_t s = p->v[i - 1], e = p->v[i];
        uint64_t *par = (uint64_t *)malloc(sizeof(uint64_t) * (cnt[s] ? cnt[s] : 1));
        sum = 0;
        for (uint32_t j = 0; j < cnt[s]; j++) {
            uint64_t loc = 0;
            for (uint32_t k = T->off[s][e][j]; k < T->off[s][e][j + 1]; k++) loc += child[T->idx[s][e][k]];
            par[j] = loc;
            sum += loc;
        }
        est[i - 1] = sum;
        free(child);
        child = par;
    }
    free(child);
}

static int order_cfl(const tin *I, const ttables *T, const uint32_t *cnt, uint32_t *order)
{
    const uint32_t n = I->n;
    int32_t core[T_MAXQ];
    orc_kcore(n, I->qoff, I->qadj, core);
    tbfs t;
    bfs_tree(n, I->qoff, I->qadj, cfl_start(I), &t);
    if (t.cnt != n) return -1;
    const uint32_t root = t.order[0];
    uint8_t vis[T_MAXQ] = {0};
    uint32_t leaves[T_MAXQ], nleaves = 0, sel_cnt = 0;
    for (uint32_t u = 0; u < n; u++) if (tdeg(I->qoff, u) == 1) leaves[nleaves++] = u;
    static tpath cpaths[4096];
    static tpath forests[T_MAXQ][512];
    uint32_t ncp = 0, nfor = 0, fsz[T_MAXQ];
    tpath cur;
    cur.len = 0;
    if (core[root] > 1) {
        core_paths_rec(core, &t, root, &cur, cpaths, &ncp);
        for (uint32_t u = 0; u < n; u++) {
            if (core[u] <= 1) continue;
            uint32_t k = 0;
            cur.len = 0;
            tree_paths_rec(I, &t, u, &cur, forests[nfor], &k);
            if (k) fsz[nfor++] = k;
        }
    } else {
        uint32_t k = 0;
        tree_paths_rec(I, &t, root, &cur, forests[nfor], &k);
        if (k) fsz[nfor++] = k;
    }
    order[sel_cnt++] = root;
    vis[root] = 1;
    if (ncp) {
        static uint64_t est[4096][T_MAXQ];
        uint32_t ntree[4096], alive[4096];
        for (uint32_t i = 0; i < ncp; i++) {
            /* generateNoneTreeEdgesCount (:241-250), + 1 */
            uint32_t nt = tdeg(I->qoff, cpaths[i].v[0]) - t.nch[cpaths[i].v[0]];
            for (uint32_t j = 1; j < cpaths[i].len; j++)
                nt += tdeg(I->qoff, cpaths[i].v[j]) - t.nch[cpaths[i].v[j]] - 1;
            ntree[i] = nt + 1;
            estimate_path(T, cnt, &cpaths[i], est[i]);
            alive[i] = i;
        }
        uint32_t nalive = ncp, sel = 0;
        double minv = 1.7976931348623157e308;
        for (uint32_t i = 0; i < nalive; i++) {
            double cv = est[alive[i]][0] / (double)ntree[alive[i]];
            if (cv < minv) { minv = cv; sel = i; }
        }
        for (;;) {
            const tpath *p = &cpaths[alive[sel]];
            for (uint32_t j = 1; j < p->len; j++) {
                if (vis[p->v[j]]) continue;
                order[sel_cnt++] = p->v[j];
                vis[p->v[j]] = 1;
            }
            for (uint32_t i = sel; i + 1 < nalive; i++) alive[i] = alive[i + 1];
            nalive--;
            if (!nalive) break;
            minv = 1.7976931348623157e308;
            sel = 0;
            for (uint32_t i = 0; i < nalive; i++) {
                const tpath *q = &cpaths[alive[i]];
                uint32_t ridx = 0;
                for (uint32_t j = 0; j < q->len; j++) {
                    if (vis[q->v[j]]) continue;
                    ridx = j - 1;
                    break;
                }
                double cv = est[alive[i]][ridx] / (double)cnt[q->v[ridx]];
                if (cv < minv) { minv = cv; sel = i; }
            }
        }
    }
    for (uint32_t f = 0; f < nfor; f++) {
        static uint64_t est[512][T_MAXQ];
        uint32_t alive[512], nalive = fsz[f];
        for (uint32_t i = 0; i < nalive; i++) { estimate_path(T, cnt, &forests[f][i], est[i]); alive[i] = i; }
        while (nalive) {
            double minv = 1.7976931348623157e308;
            uint32_t sel = 0;
            for (uint32_t i = 0; i < nalive; i++) {
                const tpath *q = &forests[f][alive[i]];
                uint32_t ridx = 0;
                for (uint32_t j = 0; j < q->len; j++) {
                    if (vis[q->v[j]]) continue;
                    ridx = j == 0 ? j : j - 1;
                    break;
                }
                double cv = est[alive[i]][ridx] / (double)cnt[q->v[ridx]];
                if (cv < minv) { minv = cv; sel = i; }
            }
            const tpath *p = &forests[f][alive[sel]];
            for (uint32_t j = 0; j < p->len; j++) {
                if (vis[p->v[j]]) continue;
                order[sel_cnt++] = p->v[j];
                vis[p->v[j]] = 1;
            }
            for (uint32_t i = sel; i + 1 < nalive; i++) alive[i] = alive[i + 1];
            nalive--;
        }
    }
    while (nleaves) {
        double minv = 1.7976931348623157e308;
        uint32_t sel = 0;
        for (uint32_t i = 0; i < nleaves; i++) {
            double cv = cnt[leaves[i]];
            if (cv < minv) { minv = cv; sel = i; }
        }
        if (!vis[leaves[sel]]) { order[sel_cnt++] = leaves[sel]; vis[leaves[sel]] = 1; }
        for (uint32_t i = sel; i + 1 < nleaves; i++) leaves[i] = leaves[i + 1];
        nleaves--;
    }
    return sel_cnt == n ? 0 : -1;
}

/* ------------------------------------------------------------------------------------
 * Order_GuP (inc/order.h:537-561): root by _GuP_start_vertex (:481-535) over the NLF
 * candidate counts; gupOrderCore (:299-384) over the 2-core with edge weights
 * e(a,b)/|C(a)|/|C(b)| and a min-heap of (cost, vertex); gupOrderTree (:412-450) with
 * path counts from count_embed (:386-410) and a min-heap of (-depth, estimate, vertex).
 * Both heaps order distinct (key, vertex) tuples totally, so a linear minimum search pops
 * them in the same order.  Returns -1 where the reference asserts (a query whose vertices
 * all have degree <= 2, :484-492).
 * ---------------------------------------------------------------------------------- */
static uint32_t gup_start(const tin *I, const int32_t *core, const uint32_t *nlfcnt)
{
    const uint32_t n = I->n;
    uint32_t maxk = 0, cand[T_MAXQ], nc = 0;
    for (uint32_t u = 0; u < n; u++) if ((uint32_t)core[u] > maxk) maxk = (uint32_t)core[u];
    for (uint32_t u = 0; u < n; u++) if ((uint32_t)core[u] == maxk) cand[nc++] = u;
    for (int k = (int)maxk; k >= 1; k--) {
        uint32_t pu[T_MAXQ], np = 0;
        double pw[T_MAXQ];
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t u = cand[i];
            int cd = 0;
            for (uint32_t p = I->qoff[u]; p < I->qoff[u + 1]; p++) cd += (int)core[I->qadj[p]] >= k;
            if (cd == 0) continue;
            pu[np] = u;
            pw[np++] = nlfcnt[u] / (double)cd;
        }
        if (!np) continue;
        double minp = pw[0];
        for (uint32_t i = 1; i < np; i++) if (pw[i] < minp) minp = pw[i];
        uint32_t mins[T_MAXQ], nm = 0;
        for (uint32_t i = 0; i < np; i++) {
            double ax = pw[i] < 0 ? -pw[i] : pw[i], am = minp < 0 ? -minp : minp;
            double sc = 1.0;
            if (ax > sc) sc = ax;
            if (am > sc) sc = am;
            double df = pw[i] - minp;
            if (df < 0) df = -df;
            if (df <= 1e-12 * sc) mins[nm++] = pu[i];
        }
        if (nm == 1 || k == 1) return mins[0];
    }
    return cand[0];
}

typedef struct {
    const tin *I;
    const ttables *T;
    const uint32_t *cnt;
    const tbfs *t;
    const uint8_t *ordered;
    double *nembeds;
} tgup;

/* count_embed (inc/order.h:386-410): path counts of u's candidates over its unordered BFS
 * subtree; nembeds[child] = (sum of the child's row sums) / |C(u)|. */
static uint64_t *count_embed(tgup *G, uint32_t u)
{
    const uint32_t cu = G->cnt[u];
    uint64_t *np = (uint64_t *)malloc(sizeof(uint64_t) * (cu ? cu : 1));
    for (uint32_t i = 0; i < cu; i++) np[i] = 1;
    for (uint32_t p = G->I->qoff[u]; p < G->I->qoff[u + 1]; p++) {
        uint32_t c = G->I->qadj[p];
        if (G->ordered[c] || G->t->parent[c] != u) continue;
        uint64_t *sub = count_embed(G, c), npsum = 0;
        const uint32_t *to = G->T->off[u][c], *ti = G->T->idx[u][c];
        for (uint32_t iv = 0; iv < cu; iv++) {
            uint64_t s = 0;
            for (uint32_t k = to[iv]; k < to[iv + 1]; k++) s += sub[ti[k]];
            np[iv] *= s;
            npsum += s;
        }
        G->nembeds[c] = (double)npsum / cu;
        free(sub);
    }
    return np;
}

static int order_gup(const tin *I, const ttables *T, const uint32_t *cnt, uint32_t *order)
{
    const uint32_t n = I->n;
    int tree_or_cycle = 1;
    for (uint32_t u = 0; u < n; u++) if (tdeg(I->qoff, u) > 2) tree_or_cycle = 0;
    if (tree_or_cycle) return -1;
    int32_t core[T_MAXQ];
    orc_kcore(n, I->qoff, I->qadj, core);
    uint32_t nlfcnt[T_MAXQ];
    /* Filter_NLF into a fresh CandidateParam (:542-544): _NLF_Filter stops at the first
     * empty set, leaving the later counts 0 */
    for (uint32_t u = 0; u < n; u++) nlfcnt[u] = 0;
    for (uint32_t u = 0; u < n; u++) { nlfcnt[u] = nlf_vertex(I, u, NULL); if (!nlfcnt[u]) break; }
    const uint32_t root = gup_start(I, core, nlfcnt);
    tbfs t;
    bfs_tree(n, I->qoff, I->qadj, root, &t);
    if (t.cnt != n) return -1;
    uint8_t ordered[T_MAXQ] = {0}, queued[T_MAXQ] = {0};
    uint32_t len = 0;
    order[len++] = root;
    ordered[root] = 1;
    /* gupOrderCore.  bfs_order[v] below is the v-th vertex of the BFS order, compared by
     * value as the reference does (:321). */
    static double w[T_MAXQ][T_MAXQ];
    double sel[T_MAXQ];
    for (uint32_t a = 0; a < n; a++) { sel[a] = 1.0; for (uint32_t b = 0; b < n; b++) w[a][b] = 0.0; }
    for (uint32_t v = 0; v < n; v++) {
        if (core[v] < 2) continue;
        for (uint32_t p = I->qoff[v]; p < I->qoff[v + 1]; p++) {
            uint32_t x = I->qadj[p];
            if (core[x] >= 2 && t.order[v] < t.order[x]) {
                double wt = (double)T->ecnt[v][x];
                wt = wt / (double)cnt[v];
                wt = wt / (double)cnt[x];
                w[v][x] = wt;
                w[x][v] = wt;
            }
        }
    }
    typedef struct { double c; int32_t v; } pqe;
    static pqe pq[T_MAXQ * T_MAXQ * 4];
    uint32_t npq = 0;
    uint32_t v = root;
    for (;;) {
        for (uint32_t p = I->qoff[v]; p < I->qoff[v + 1]; p++) {
            uint32_t x = I->qadj[p];
            if (core[x] >= 2 && !ordered[x]) { queued[x] = 1; sel[x] *= w[v][x]; }
        }
        for (uint32_t p = I->qoff[v]; p < I->qoff[v + 1]; p++) {
            uint32_t x = I->qadj[p];
            if (core[x] >= 2 && !ordered[x]) {
                double c = sel[x] * (double)cnt[x];
                for (uint32_t q = I->qoff[x]; q < I->qoff[x + 1]; q++) {
                    uint32_t y = I->qadj[q];
                    if (queued[y] && !ordered[y]) c *= w[x][y] * sel[y];
                }
                pq[npq].c = c;
                pq[npq].v = (int32_t)x;
                npq++;
            }
        }
        int found = 0;
        while (npq) {
            uint32_t m = 0;
            for (uint32_t i = 1; i < npq; i++)
                if (pq[i].c < pq[m].c || (pq[i].c == pq[m].c && pq[i].v < pq[m].v)) m = i;
            v = (uint32_t)pq[m].v;
            pq[m] = pq[--npq];
            if (!ordered[v]) { found = 1; break; }
        }
        if (!found) break;
        order[len++] = v;
        ordered[v] = 1;
    }
    /* gupOrderTree */
    double nembeds[T_MAXQ];
    for (uint32_t u = 0; u < n; u++) nembeds[u] = 0.0;
    tgup G = {I, T, cnt, &t, ordered, nembeds};
    typedef struct { int32_t d; double e; uint32_t v; } the;
    static the hp[T_MAXQ * T_MAXQ * 4];
    uint32_t nh = 0;
    for (uint32_t u = 0; u < n; u++) {
        int has = 0;
        for (uint32_t p = I->qoff[u]; p < I->qoff[u + 1]; p++) has |= !ordered[I->qadj[p]];
        if (ordered[u] && has) {
            free(count_embed(&G, u));
            for (uint32_t p = I->qoff[u]; p < I->qoff[u + 1]; p++) {
                uint32_t x = I->qadj[p];
                if (!ordered[x]) { hp[nh].d = 0; hp[nh].e = nembeds[x]; hp[nh].v = x; nh++; }
            }
        }
    }
    while (nh) {
        uint32_t m = 0;
        for (uint32_t i = 1; i < nh; i++) {
            const the *a = &hp[i], *b = &hp[m];
            if (a->d < b->d || (a->d == b->d && (a->e < b->e || (a->e == b->e && a->v < b->v)))) m = i;
        }
        the top = hp[m];
        hp[m] = hp[--nh];
        if (ordered[top.v]) return -1; /* the reference asserts here (:434) */
        order[len++] = top.v;
        ordered[top.v] = 1;
        for (uint32_t p = I->qoff[top.v]; p < I->qoff[top.v + 1]; p++) {
            uint32_t x = I->qadj[p];
            if (!ordered[x]) { hp[nh].d = top.d - 1; hp[nh].e = nembeds[x]; hp[nh].v = x; nh++; }
        }
    }
    return len == n ? 0 : -1;
}

/* kind: 1 QSI, 2 CFL, 3 CECI, 4 GuP.  cand[u*stride ...] / cnt[u]: the filtered candidate
 * sets (the tables are built from them).  Returns 0, or -1 where the reference asserts or
 * the query is disconnected. */
int orc_order_tree(int kind, uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *dlab, uint32_t n,
                   const uint32_t *qoff, const uint32_t *qadj, const uint32_t *qlab, const uint32_t *cand,
                   const uint32_t *cnt, uint32_t stride, uint32_t *order)
{
    if (n < 2 || n > T_MAXQ || kind < 1 || kind > 4) return -1;
    tin I;
    tin_init(&I, V, off, adj, dlab, n, qoff, qadj, qlab);
    if (kind == 3) {
        /* Order_CECI (:599-607): the BFS order from CECIStartVertex */
        tbfs t;
        bfs_tree(n, qoff, qadj, ceci_daf_start(&I, 0), &t);
        if (t.cnt != n) return -1;
        memcpy(order, t.order, sizeof(uint32_t) * n);
        return 0;
    }
    ttables T;
    tables_build(&I, cand, cnt, stride, &T);
    int rc = 0;
    if (kind == 1) order_qsi(&I, &T, order);
    else if (kind == 2) rc = order_cfl(&I, &T, cnt, order);
    else rc = order_gup(&I, &T, cnt, order);
    tables_free(&T);
    return rc;
}
