/*
 * sgm_oracle_dyn.c -- CPU restatement of the reference's dynamic-order searches
 * (explore_Intersection_DynamicOrder_DAF / _VEQ, inc/explore.h:1674-2162) and of the DAF
 * weight array they rank with (computeWeightArray, inc/order.h:980-1030).
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see sgm_oracle.c).  Citations are relative to
 * the Orrinn/SubgraphMatching source tree.
 *
 * The search keeps the reference's data flow: candidate lists per query vertex from the aux
 * tables of its previous neighbours in BFS order (getCandidatesFor), a priority queue of the
 * extendable vertices per depth, copied from the depth above (vec_rank_queue), and the
 * libstdc++ heap moves of std::priority_queue under the reference's comparator, which orders
 * vertices of degree 1 last and the rest by ascending weight but leaves ties to the heap.
 *
 * Pinning (tests/test_oracle_dyn_pins.py): counts equal the plain candidate-space search's
 * and brute force (an order never changes a count); intersections, chosen orders and weights
 * equal a C++ restatement built on std::priority_queue itself (oracle/sort_probe.cpp) and a
 * Python restatement of the weight recursion.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define D_MAXQ 32

/* the reference's Intersection as sgm_oracle.c builds it (AVX2 under -mavx2, else a merge) */
uint32_t orc_isect(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out);

typedef struct {
    uint64_t count;
    uint64_t intersections;
    uint64_t nodes[D_MAXQ];  /* extensions per depth of the dynamic search (not a reference counter) */
    double enum_ms;
    int32_t timed_out;
    int32_t capped;
} dyn_stats;

/* one queue entry: (vertex, its query degree) and its weight, as the reference's
 * pair<pair<VertexID, uint32_t>, uint32_t> */
typedef struct { uint32_t u, deg, w; } dq_e;

/* extendable_vertex_compare (inc/explore.h:1675-1686): l ranks below r when l has degree 1
 * and r not, or both or neither do and l's weight is larger */
static int dq_less(const dq_e *l, const dq_e *r)
{
    if (l->deg == 1 && r->deg != 1) return 1;
    if (l->deg != 1 && r->deg == 1) return 0;
    return l->w > r->w;
}

typedef struct { dq_e a[D_MAXQ + 1]; long len; } dq_t;

/* libstdc++'s push_heap / pop_heap (the same moves as in sgm_oracle_tree.c) */
static void dq_sift_up(dq_e *f, long hole, long top, dq_e val)
{
    long parent = (hole - 1) / 2;
    while (hole > top && dq_less(&f[parent], &val)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = val;
}

static void dq_push(dq_t *h, dq_e x)
{
    h->a[h->len++] = x;
    dq_sift_up(h->a, h->len - 1, 0, h->a[h->len - 1]);
}

static void dq_pop(dq_t *h)
{
    if (h->len > 1) {
        dq_e *f = h->a;
        long len = h->len - 1, hole = 0, second = 0;
        dq_e val = f[len];
        f[len] = f[0];
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            if (dq_less(&f[second], &f[second - 1])) second--;
            f[hole] = f[second];
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            f[hole] = f[second - 1];
            hole = second - 1;
        }
        dq_sift_up(f, hole, 0, val);
    }
    h->len--;
}

/* ------------------------------------------------------------------------------------ */
typedef struct {
    uint32_t n, V;
    const uint32_t *off, *adj, *qoff, *qadj, *cand, *cnt;
    uint32_t stride;
    const uint32_t *bfs;
    uint32_t pos_of[D_MAXQ];
    uint32_t prev[D_MAXQ][D_MAXQ], np[D_MAXQ], fol[D_MAXQ][D_MAXQ], nf[D_MAXQ];
    /* aux tables of every directed query edge: row i of (a, b) lists positions in C(b) */
    uint32_t *toff[D_MAXQ][D_MAXQ], *tidx[D_MAXQ][D_MAXQ];
    const uint32_t *const *weight; /* DAF weights per vertex and candidate position, or NULL (VEQ) */
    /* search state */
    uint32_t *valid[D_MAXQ], tot[D_MAXQ], tmp_cap;
    uint32_t *tmp;
    uint32_t emb_idx[D_MAXQ], ext[D_MAXQ], mo[D_MAXQ];
    uint8_t *vis;
    dq_t q[D_MAXQ + 1];
    uint64_t cap;
    double limit_ms, t0;
    int stop;
    dyn_stats *st;
} dctx;

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_THREAD_CPUTIME_ID, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static void tables_of(dctx *c)
{
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (c->V ? c->V : 1));
    for (uint32_t a = 0; a < c->n; a++)
        for (uint32_t p = c->qoff[a]; p < c->qoff[a + 1]; p++) {
            uint32_t b = c->qadj[p];
            const uint32_t *ca = c->cand + (size_t)a * c->stride, *cb = c->cand + (size_t)b * c->stride;
            for (uint32_t v = 0; v < c->V; v++) pos[v] = -1;
            for (uint32_t k = 0; k < c->cnt[b]; k++) pos[cb[k]] = (int32_t)k;
            uint64_t e = 0;
            for (uint32_t i = 0; i < c->cnt[a]; i++)
                for (uint32_t q = c->off[ca[i]]; q < c->off[ca[i] + 1]; q++) e += pos[c->adj[q]] >= 0;
            uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * (c->cnt[a] + 1));
            uint32_t *ti = (uint32_t *)malloc(sizeof(uint32_t) * (e ? e : 1));
            e = 0;
            for (uint32_t i = 0; i < c->cnt[a]; i++) {
                to[i] = (uint32_t)e;
                for (uint32_t q = c->off[ca[i]]; q < c->off[ca[i] + 1]; q++)
                    if (pos[c->adj[q]] >= 0) ti[e++] = (uint32_t)pos[c->adj[q]];
            }
            to[c->cnt[a]] = (uint32_t)e;
            c->toff[a][b] = to;
            c->tidx[a][b] = ti;
        }
    free(pos);
}

static void tables_drop(dctx *c)
{
    for (uint32_t a = 0; a < D_MAXQ; a++)
        for (uint32_t b = 0; b < D_MAXQ; b++) { free(c->toff[a][b]); free(c->tidx[a][b]); }
}

/* the previous / following neighbours of every vertex in BFS position (_GetNeb) */
static void bfs_nebs(dctx *c)
{
    for (uint32_t i = 0; i < c->n; i++) c->pos_of[c->bfs[i]] = i;
    for (uint32_t v = 0; v < c->n; v++) {
        c->np[v] = c->nf[v] = 0;
        for (uint32_t i = 0; i < c->n; i++) {
            uint32_t w = c->bfs[i];
            int adjw = 0;
            for (uint32_t p = c->qoff[v]; p < c->qoff[v + 1]; p++) adjw |= c->qadj[p] == w;
            if (!adjw) continue;
            if (i < c->pos_of[v]) c->prev[v][c->np[v]++] = w;
            else if (i > c->pos_of[v]) c->fol[v][c->nf[v]++] = w;
        }
    }
}

/* getCandidatesFor (inc/explore.h:1691-1724): the row of the first previous neighbour's
 * table, intersected with the rows of the others, in order */
static void candidates_for(dctx *c, uint32_t u)
{
    uint32_t p0 = c->prev[u][0], r = c->emb_idx[p0];
    const uint32_t *row = c->tidx[p0][u] + c->toff[p0][u][r];
    uint32_t k = c->toff[p0][u][r + 1] - c->toff[p0][u][r];
    memcpy(c->valid[u], row, sizeof(uint32_t) * k);
    for (uint32_t i = 1; i < c->np[u]; i++) {
        uint32_t p = c->prev[u][i], rr = c->emb_idx[p];
        const uint32_t *ro = c->tidx[p][u] + c->toff[p][u][rr];
        uint32_t rk = c->toff[p][u][rr + 1] - c->toff[p][u][rr];
        c->st->intersections++;
        const uint32_t m = orc_isect(ro, rk, c->valid[u], k, c->tmp);
        memcpy(c->valid[u], c->tmp, sizeof(uint32_t) * m);
        k = m;
    }
    c->tot[u] = k;
}

/* updataExtendableV (inc/explore.h:1726-1770): a following neighbour whose last previous
 * neighbour was just matched gets its candidates and enters the queue with its weight (the
 * sum of its candidates' DAF weights, u32 wrapping; the candidate count without weights) */
static void update_ext(dctx *c, uint32_t u, dq_t *q)
{
    for (uint32_t k = 0; k < c->nf[u]; k++) {
        uint32_t f = c->fol[u][k];
        if (--c->ext[f] == 0) {
            candidates_for(c, f);
            uint32_t w = 0;
            if (c->weight)
                for (uint32_t j = 0; j < c->tot[f]; j++) w += c->weight[f][c->valid[f][j]];
            else
                w = c->tot[f];
            dq_e e = {f, c->qoff[f + 1] - c->qoff[f], w};
            dq_push(q, e);
        }
    }
}

static void restore_ext(dctx *c, uint32_t u)
{
    for (uint32_t k = 0; k < c->nf[u]; k++) c->ext[c->fol[u][k]]++;
}

/* _explore_recur_DynamicOrder_DAF / _VEQ (inc/explore.h:1779-1835, 1933-1990) */
static void rec(dctx *c, uint32_t d)
{
    uint32_t u = c->mo[d];
    if (c->tot[u] == 0) return;
    for (uint32_t p = 0; p < c->tot[u]; p++) {
        if (c->limit_ms > 0 && (p & 63) == 0 && now_ms() - c->t0 > c->limit_ms) { c->stop = 2; return; }
        uint32_t vidx = c->valid[u][p], v = c->cand[(size_t)u * c->stride + vidx];
        if (c->vis[v]) continue;
        c->st->nodes[d]++;
        if (d == c->n - 1) {
            c->st->count++;
            if (c->st->count == c->cap) { c->stop = 1; return; }
            continue;
        }
        c->emb_idx[u] = vidx;
        c->vis[v] = 1;
        c->q[d + 1] = c->q[d];
        update_ext(c, u, &c->q[d + 1]);
        uint32_t next = c->q[d + 1].a[0].u;
        dq_pop(&c->q[d + 1]);
        c->mo[d + 1] = next;
        rec(c, d + 1);
        if (c->stop) return;
        c->vis[v] = 0;
        restore_ext(c, u);
    }
}

/* explore_Intersection_DynamicOrder_DAF (weight != NULL) / _VEQ (weight == NULL),
 * inc/explore.h:1849-1931, 1993-2062.  bfs_order = the BFS order from DAFStartVertex. */
int orc_enum_dynamic(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n, const uint32_t *qoff,
                     const uint32_t *qadj, const uint32_t *cand, const uint32_t *cnt, uint32_t stride,
                     const uint32_t *bfs_order, const uint32_t *const *weight, uint64_t cap, double time_limit_ms,
                     dyn_stats *st)
{
    if (n < 2 || n > D_MAXQ) return -1;
    dctx *c = (dctx *)calloc(1, sizeof(dctx));
    c->n = n; c->V = V; c->off = off; c->adj = adj; c->qoff = qoff; c->qadj = qadj;
    c->cand = cand; c->cnt = cnt; c->stride = stride; c->bfs = bfs_order; c->weight = weight;
    c->cap = cap ? cap : UINT64_MAX;
    c->limit_ms = time_limit_ms;
    c->st = st;
    memset(st, 0, sizeof *st);
    bfs_nebs(c);
    for (uint32_t v = 1; v < n; v++) if (c->np[bfs_order[v]] == 0) { free(c); return -1; }
    tables_of(c);
    uint32_t mx = 1;
    for (uint32_t u = 0; u < n; u++) if (cnt[u] > mx) mx = cnt[u];
    for (uint32_t u = 0; u < n; u++) c->valid[u] = (uint32_t *)malloc(sizeof(uint32_t) * mx);
    c->tmp = (uint32_t *)malloc(sizeof(uint32_t) * mx);
    c->vis = (uint8_t *)calloc(V ? V : 1, 1);
    for (uint32_t u = 0; u < n; u++) c->ext[u] = c->np[u];
    const uint32_t u0 = bfs_order[0];
    c->mo[0] = u0;
    c->t0 = now_ms();
    for (uint32_t p = 0; p < cnt[u0] && !c->stop; p++) {
        uint32_t v0 = cand[(size_t)u0 * stride + p];
        c->st->nodes[0]++;
        c->emb_idx[u0] = p;
        c->vis[v0] = 1;
        c->q[1].len = 0;
        update_ext(c, u0, &c->q[1]);
        uint32_t next = c->q[1].a[0].u;
        dq_pop(&c->q[1]);
        c->mo[1] = next;
        rec(c, 1);
        if (c->stop) break;
        c->vis[v0] = 0;
        restore_ext(c, u0);
    }
    st->enum_ms = now_ms() - c->t0;
    st->capped = c->stop == 1;
    st->timed_out = c->stop == 2;
    for (uint32_t u = 0; u < n; u++) free(c->valid[u]);
    free(c->tmp);
    free(c->vis);
    tables_drop(c);
    free(c);
    return 0;
}

/* computeWeightArray (inc/order.h:980-1030) over the given BFS tree (parent[], BFS order):
 * from the last BFS vertex back, a vertex's weight per candidate is the minimum over its
 * children that have one previous neighbour of the sum of the child's weights along the
 * candidate's table row (u32 sums, wrapping), or 1 everywhere when no child qualifies;
 * weights start at UINT32_MAX (src/order_ablations.cpp:134-138).  w[u] holds cnt[u] words. */
void orc_daf_weights(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n, const uint32_t *qoff,
                     const uint32_t *qadj, const uint32_t *cand, const uint32_t *cnt, uint32_t stride,
                     const uint32_t *bfs_order, const uint32_t *parent, uint32_t *const *w)
{
    dctx *c = (dctx *)calloc(1, sizeof(dctx));
    c->n = n; c->V = V; c->off = off; c->adj = adj; c->qoff = qoff; c->qadj = qadj;
    c->cand = cand; c->cnt = cnt; c->stride = stride; c->bfs = bfs_order;
    bfs_nebs(c);
    tables_of(c);
    for (uint32_t u = 0; u < n; u++)
        for (uint32_t k = 0; k < cnt[u]; k++) w[u][k] = 0xFFFFFFFFu;
    for (int i = (int)n - 1; i >= 0; i--) {
        uint32_t x = bfs_order[i];
        int one = 1;
        /* the tree children of x in adjacency (BFS discovery) order */
        for (uint32_t p = qoff[x]; p < qoff[x + 1]; p++) {
            uint32_t ch = qadj[p];
            if (parent[ch] != x || ch == bfs_order[0] || c->pos_of[ch] < c->pos_of[x]) continue;
            if (c->np[ch] != 1) continue;
            one = 0;
            for (uint32_t k = 0; k < cnt[x]; k++) {
                uint32_t s = 0;
                for (uint32_t e = c->toff[x][ch][k]; e < c->toff[x][ch][k + 1]; e++) s += w[ch][c->tidx[x][ch][e]];
                if (s < w[x][k]) w[x][k] = s;
            }
        }
        if (one)
            for (uint32_t k = 0; k < cnt[x]; k++) w[x][k] = 1;
    }
    tables_drop(c);
    free(c);
}

uint32_t orc_dyn_stats_size(void) { return (uint32_t)sizeof(dyn_stats); }
