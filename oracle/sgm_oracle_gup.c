/*
 * sgm_oracle_gup.c -- CPU restatement of GuP's conflict detection (GUPConflictDetection,
 * inc/prune.h:488-923): reservation guards per (depth, candidate position), computed once
 * from the last depth back (Init :828-905, compute_rev :599-825), and the search-time check
 * (matchCheck :577-595) inside the aux-table search (explore_Intersection_All with
 * useGUPConflictDetection, inc/explore.h:2610-2848, PruningOpt::ForwardPrune
 * inc/prune.h:2252-2303).
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see sgm_oracle.c).  Citations are relative to
 * the Orrinn/SubgraphMatching source tree.
 *
 * A guard of candidate v of the vertex at depth d is a set of at most r data vertices (r =
 * Config::gupConflictDetectionSize = 3, inc/api.h:204): when all of them are matched at
 * depths < d, v cannot complete an embedding (every following neighbour of the vertex runs
 * out of candidates), so v is pruned.  The construction keeps the reference's data flow:
 * following neighbours from the last back, per neighbour a greedy choice of reservation
 * vertices over the pairs left uncovered, std::partition's element order (libstdc++'s
 * two-ended swap) and the last candidate of highest frequency among those that keep the
 * domain condition.  The guard kept is the one of smallest size, ties to the deeper
 * neighbour; a failed neighbour drops the guard altogether.
 *
 * Pinning (tests/test_oracle_gup_pins.py): counts with the check equal counts without it and
 * brute force (a guard may only remove candidates that complete no embedding); the guards
 * equal a restatement built on std::partition and std::unordered_set themselves
 * (oracle/sort_probe.cpp); the check prunes on the pinned inputs.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define G_MAXQ 32
#define G_MAXR 8

typedef struct {
    uint32_t n, V, r;
    const uint32_t *off, *adj, *qoff, *qadj, *cand, *cnt, *order;
    uint32_t stride;
    uint32_t fneb[G_MAXQ][G_MAXQ], nfneb[G_MAXQ];  /* following neighbours' depths, ascending */
    uint32_t former[G_MAXQ];                         /* depths below d */
    uint32_t *match;                                 /* V: depths whose candidates hold v */
    uint32_t *freq;                                  /* V: coverInfo freq */
    uint32_t *toff[G_MAXQ][G_MAXQ], *tidx[G_MAXQ][G_MAXQ]; /* tables by query vertex (a, b) */
    uint32_t *goff;                                  /* per depth: first guard record */
    uint32_t *g;                                     /* 4 words per record: size, vertices */
    /* scratch */
    uint32_t chosen[G_MAXR + 2], nchosen;
    uint32_t *gr;                                    /* pairs (vf, w) */
    uint64_t ngr, gr_cap;
    uint32_t *rev;                                   /* rev_candidates */
    uint64_t nrev, rev_cap;
} gctx;

static int in_chosen(const gctx *c, uint32_t x)
{
    for (uint32_t i = 0; i < c->nchosen; i++)
        if (c->chosen[i] == x) return 1;
    return 0;
}

static void add_chosen(gctx *c, uint32_t x)
{
    if (!in_chosen(c, x)) c->chosen[c->nchosen++] = x;
}

static uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

static void push_gr(gctx *c, uint32_t a, uint32_t b)
{
    if (c->ngr + 1 > c->gr_cap) {
        c->gr_cap = c->gr_cap ? 2 * c->gr_cap : 256;
        c->gr = (uint32_t *)realloc(c->gr, sizeof(uint32_t) * 2 * c->gr_cap);
    }
    c->gr[2 * c->ngr] = a;
    c->gr[2 * c->ngr + 1] = b;
    c->ngr++;
}

static void push_rev(gctx *c, uint32_t x)
{
    if (c->nrev + 1 > c->rev_cap) {
        c->rev_cap = c->rev_cap ? 2 * c->rev_cap : 256;
        c->rev = (uint32_t *)realloc(c->rev, sizeof(uint32_t) * c->rev_cap);
    }
    c->rev[c->nrev++] = x;
}

static const uint32_t *guard_of(const gctx *c, uint32_t dep, uint32_t idx)
{
    return c->g + 4 * (size_t)(c->goff[dep] + idx);
}

/* libstdc++'s std::partition on a random-access range (the bidirectional algorithm): from the
 * left the first element failing the predicate, from the right the last one holding it,
 * swapped, until the two meet; returns the end of the holding part */
typedef struct { const gctx *c; uint32_t ud; uint32_t dom; } ppred;

static int pred_ok(const ppred *p, uint32_t v)
{
    const gctx *c = p->c;
    return c->freq[v] > 0 && c->nchosen + 1 <= popc32((p->dom | c->match[v]) & c->former[p->ud]);
}

static uint64_t partition_rev(gctx *c, const ppred *p)
{
    uint64_t first = 0, last = c->nrev;
    while (1) {
        while (1) {
            if (first == last) return first;
            if (pred_ok(p, c->rev[first])) first++;
            else break;
        }
        last--;
        while (1) {
            if (first == last) return first;
            if (!pred_ok(p, c->rev[last])) last--;
            else break;
        }
        uint32_t t = c->rev[first]; c->rev[first] = c->rev[last]; c->rev[last] = t;
        first++;
    }
}

/* compute_rev (inc/prune.h:599-825): the guard of candidate vi of the vertex at depth ud */
static void compute_rev(gctx *c, uint32_t ud, uint32_t vi, uint32_t *out)
{
    const uint32_t u = c->order[ud];
    const uint32_t v = c->cand[(size_t)u * c->stride + vi];
    uint32_t best_sz = 0, best_dep = 0, best[G_MAXR];
    int have = 0;
    out[0] = 0;
    for (int k = (int)c->nfneb[ud] - 1; k >= 0; k--) {
        const uint32_t ufd = c->fneb[ud][k], uf = c->order[ufd];
        c->ngr = 0;
        c->nchosen = 0;
        c->nrev = 0;
        uint32_t dom = 0;
        const uint32_t *row = c->tidx[u][uf] + c->toff[u][uf][vi];
        const uint32_t rn = c->toff[u][uf][vi + 1] - c->toff[u][uf][vi];
        for (uint32_t e = 0; e < rn; e++) {
            const uint32_t fi = row[e], vf = c->cand[(size_t)uf * c->stride + fi];
            const int vf_ok = (c->match[vf] & c->former[ud]) != 0;
            const uint32_t *gg = guard_of(c, ufd, fi);
            uint32_t tmp1 = vf;
            const uint32_t *ws = gg[0] ? gg + 1 : &tmp1;  /* the trivial guard {vf} */
            const uint32_t wn = gg[0] ? gg[0] : 1;
            for (uint32_t wi = 0; wi < wn; wi++) {
                const uint32_t w = ws[wi];
                if (w == v) continue;
                if (in_chosen(c, vf)) break;
                const int w_ok = (c->match[w] & c->former[ud]) != 0;
                if (vf_ok && (vf == w || !w_ok)) {
                    add_chosen(c, vf);
                    dom |= c->match[vf];
                } else if (!vf_ok && w_ok) {
                    add_chosen(c, w);
                    dom |= c->match[w];
                }
                if ((!vf_ok && !w_ok) || c->nchosen > c->r || c->nchosen > popc32(c->former[ud] & dom)) {
                    out[0] = 0;  /* no guard at all */
                    return;
                }
            }
        }
        /* the pairs a reservation vertex must still cover */
        for (uint32_t e = 0; e < rn; e++) {
            const uint32_t fi = row[e], vf = c->cand[(size_t)uf * c->stride + fi];
            const uint32_t *gg = guard_of(c, ufd, fi);
            if (!gg[0] || in_chosen(c, vf)) continue;
            for (uint32_t wi = 0; wi < gg[0]; wi++) {
                const uint32_t w = gg[1 + wi];
                if (w == v) continue;
                if (!in_chosen(c, w)) push_gr(c, vf, w);
            }
        }
        for (uint64_t i = 0; i < c->ngr; i++) {
            const uint32_t vf = c->gr[2 * i], w = c->gr[2 * i + 1];
            if (c->freq[vf] == 0) push_rev(c, vf);
            if (c->freq[w] == 0) push_rev(c, w);
            c->freq[vf]++;
            c->freq[w]++;
        }
        while (c->nchosen < c->r) {
            ppred p = {c, ud, dom};
            const uint64_t valid = partition_rev(c, &p);
            if (valid == 0) break;
            uint64_t mi = 0;
            uint32_t mf = 0;
            for (uint64_t i = 0; i < valid; i++)
                if (c->freq[c->rev[i]] >= mf) { mi = i; mf = c->freq[c->rev[i]]; }
            const uint32_t pick = c->rev[mi];
            add_chosen(c, pick);
            /* drop the pairs it covers (order kept, as vector::erase) */
            uint64_t keep = 0;
            for (uint64_t i = 0; i < c->ngr; i++) {
                const uint32_t vf = c->gr[2 * i], w = c->gr[2 * i + 1];
                if (vf == pick || w == pick) {
                    c->freq[vf]--;
                    c->freq[w]--;
                } else {
                    c->gr[2 * keep] = vf;
                    c->gr[2 * keep + 1] = w;
                    keep++;
                }
            }
            c->ngr = keep;
        }
        for (uint64_t i = 0; i < c->nrev; i++) c->freq[c->rev[i]] = 0;
        if (c->ngr == 0 && c->nchosen) {
            /* keep the smallest guard, ties to the deeper neighbour (neighbours come deepest
             * first, so a later one of equal size never replaces the kept one) */
            if (!have || c->nchosen < best_sz) {
                have = 1;
                best_sz = c->nchosen;
                best_dep = ufd;
                memcpy(best, c->chosen, sizeof(uint32_t) * c->nchosen);
            }
        }
    }
    (void)best_dep;
    if (have) {
        /* sorted, as the ReservationGuard constructor keeps it */
        for (uint32_t a = 1; a < best_sz; a++)
            for (uint32_t b = a; b > 0 && best[b - 1] > best[b]; b--) {
                uint32_t t = best[b]; best[b] = best[b - 1]; best[b - 1] = t;
            }
        out[0] = best_sz;
        for (uint32_t i = 0; i < best_sz; i++) out[1 + i] = best[i];
    }
}

static uint32_t *table_off(uint32_t V, const uint32_t *off, const uint32_t *adj, const uint32_t *ca, uint32_t na,
                           const uint32_t *cb, uint32_t nb, uint32_t **idx_out, int32_t *pos)
{
    for (uint32_t k = 0; k < nb; k++) pos[cb[k]] = (int32_t)k;
    uint64_t e = 0;
    for (uint32_t i = 0; i < na; i++)
        for (uint32_t q = off[ca[i]]; q < off[ca[i] + 1]; q++) e += pos[adj[q]] >= 0;
    uint32_t *to = (uint32_t *)malloc(sizeof(uint32_t) * (na + 1));
    uint32_t *ti = (uint32_t *)malloc(sizeof(uint32_t) * (e ? e : 1));
    e = 0;
    for (uint32_t i = 0; i < na; i++) {
        to[i] = (uint32_t)e;
        for (uint32_t q = off[ca[i]]; q < off[ca[i] + 1]; q++)
            if (pos[adj[q]] >= 0) ti[e++] = (uint32_t)pos[adj[q]];
    }
    to[na] = (uint32_t)e;
    for (uint32_t k = 0; k < nb; k++) pos[cb[k]] = -1;
    *idx_out = ti;
    (void)V;
    return to;
}

/* Init (inc/prune.h:828-905): guards of every (depth, candidate position) for `order`, from
 * the last depth back.  goff[d] = first record of depth d (records in depth order, cnt of
 * order[d] each; goff has n + 1 entries, filled here); out = 4 words per record: guard size
 * (0 = none), then its vertices ascending.  r <= G_MAXR - 2. */
int orc_gup_guards(uint32_t V, const uint32_t *off, const uint32_t *adj, uint32_t n, const uint32_t *qoff,
                   const uint32_t *qadj, const uint32_t *cand, const uint32_t *cnt, uint32_t stride,
                   const uint32_t *order, uint32_t r, uint32_t *goff, uint32_t *out)
{
    if (n < 1 || n > G_MAXQ || r < 1 || r > 3) return -1;
    gctx *c = (gctx *)calloc(1, sizeof(gctx));
    c->n = n; c->V = V; c->r = r; c->off = off; c->adj = adj; c->qoff = qoff; c->qadj = qadj;
    c->cand = cand; c->cnt = cnt; c->order = order; c->stride = stride;
    uint32_t pos_of[G_MAXQ];
    for (uint32_t i = 0; i < n; i++) pos_of[order[i]] = i;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t u = order[i];
        c->nfneb[i] = 0;
        for (uint32_t p = qoff[u]; p < qoff[u + 1]; p++)
            if (pos_of[qadj[p]] > i) c->fneb[i][c->nfneb[i]++] = pos_of[qadj[p]];
        for (uint32_t a = 1; a < c->nfneb[i]; a++)
            for (uint32_t b = a; b > 0 && c->fneb[i][b - 1] > c->fneb[i][b]; b--) {
                uint32_t t = c->fneb[i][b]; c->fneb[i][b] = c->fneb[i][b - 1]; c->fneb[i][b - 1] = t;
            }
        c->former[i] = i ? (c->former[i - 1] | (1u << (i - 1))) : 0u;
    }
    c->match = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    c->freq = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t u = order[i];
        for (uint32_t j = 0; j < cnt[u]; j++) c->match[cand[(size_t)u * stride + j]] |= 1u << i;
    }
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (V ? V : 1));
    for (uint32_t v = 0; v < V; v++) pos[v] = -1;
    for (uint32_t a = 0; a < n; a++)
        for (uint32_t p = qoff[a]; p < qoff[a + 1]; p++) {
            const uint32_t b = qadj[p];
            c->toff[a][b] = table_off(V, off, adj, cand + (size_t)a * stride, cnt[a], cand + (size_t)b * stride,
                                      cnt[b], &c->tidx[a][b], pos);
        }
    goff[0] = 0;
    for (uint32_t i = 0; i < n; i++) goff[i + 1] = goff[i] + cnt[order[i]];
    c->goff = goff;
    c->g = out;
    memset(out, 0, sizeof(uint32_t) * 4 * (size_t)goff[n]);
    for (int i = (int)n - 1; i >= 0; i--)
        for (uint32_t j = 0; j < cnt[order[i]]; j++) compute_rev(c, (uint32_t)i, j, out + 4 * (size_t)(goff[i] + j));
    for (uint32_t a = 0; a < G_MAXQ; a++)
        for (uint32_t b = 0; b < G_MAXQ; b++) { free(c->toff[a][b]); free(c->tidx[a][b]); }
    free(pos); free(c->match); free(c->freq); free(c->gr); free(c->rev);
    free(c);
    return 0;
}
