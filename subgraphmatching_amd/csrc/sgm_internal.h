// Internal layouts shared by the host driver and the gfx950 kernels.
//
// Device data layout (HBM):
//   data graph    off[V+1] u32, adj[2E] u32 (each list sorted ascending), by_deg[V] u32
//                 (vertex ids by degree, descending, ties by id), ge_cnt[33] u32 (number of
//                 vertices with degree >= d, d = 0..32)
//   per query     vmask[V] u32: bit u set iff data vertex v is in C(u) -- the reference's
//                 candidates_exsist[u][v] (inc/api.h:15) transposed so that one load answers
//                 "which query vertices may take v"; bm[n][W] u64 per-vertex bitmaps with
//                 rank[n][W] u32 exclusive popcount prefixes (W = ceil(V/64)); cand lists
//                 (ascending) with cand_off[n+1]; aux tables (CSR per directed query edge).
#pragma once
#include <stdint.h>

#define SGM_MAXQ 32
#define SGM_MAXOPS 512
#define SGM_INVALID 0xFFFFFFFFu
#define SGM_NOPARENT 0xFFFFu

// Intersection-cache plan on the device (the reference's PlanIR, inc/code_gen.h:682-847).
// Ops are sorted by depth, so the ops computed at depth d are the contiguous id range
// [op_first[d], op_first[d+1]) -- the reference's setOps[d].
struct sgm_plan_dev {
    uint32_t n;
    uint32_t n_ops;
    uint32_t order[SGM_MAXQ];      // query vertex at each depth
    uint32_t iter_set[SGM_MAXQ];   // op id iterated at depth d (SGM_INVALID at depth 0)
    uint32_t op_first[SGM_MAXQ + 1];
    uint64_t anc[SGM_MAXQ];        // DAF ancestor masks over depths (inc/prune.h:100-104)
    uint16_t op_parent[SGM_MAXOPS];
    uint8_t op_vdepth[SGM_MAXOPS];
    uint8_t op_depth[SGM_MAXOPS];
};

// Aux-table plan (explore_Intersection_All, inc/explore.h:2610-2848).  For depth d the
// backward neighbours (in order position, GetPreviousNeb) are entries
// [prev_first[d], prev_first[d+1]); entry e names the neighbour's depth and the device
// table (offsets over C(prev), candidate positions in C(order[d])) of that directed edge.
#define SGM_MAXPREV 512
struct sgm_auxplan_dev {
    uint32_t n;
    uint32_t order[SGM_MAXQ];
    uint32_t prev_first[SGM_MAXQ + 1];
    uint32_t prev_depth[SGM_MAXPREV];
    // table of backward edge e: the row of data vertex v is tab_idx[e][tab_off[e][v] ..
    // tab_off[e][v + 1]) -- offsets by vertex id, shared by every edge into the same vertex
    const uint32_t* tab_off[SGM_MAXPREV];
    const uint32_t* tab_idx[SGM_MAXPREV];
    const uint32_t* cand[SGM_MAXQ];   // candidate list of the query vertex at each depth
    const uint32_t* root_order;       // positions in cand[0] in degree order (the items' roots)
    uint64_t anc[SGM_MAXQ];           // DAF ancestor masks over depths (failing sets, inc/prune.h:100-104)
    // GuP conflict detection (inc/prune.h:488-923), or null: per depth, 4 words per candidate
    // position -- the guard's size (0 = none), then its data vertices
    const uint32_t* guard[SGM_MAXQ];
};

// Plan of the dynamic-order searches (explore_Intersection_DynamicOrder_DAF / _VEQ,
// inc/explore.h:1674-2162), indexed by query vertex, not depth: the previous neighbours of u
// in the start order (GetPreviousNeb, order-position order) are edges [prev_first[u],
// prev_first[u+1]) with their tables (p -> u) and p itself; the following neighbours
// (GetFollowingNeb) are fol[fol_first[u] .. fol_first[u+1]).  weight[u] holds the DAF weight
// of each position of C(u) (computeWeightArray), or is null when the ranking is by candidate
// count (the VEQ variant).  deg1: bit u set when u has query degree 1.
struct sgm_dynplan_dev {
    uint32_t n;
    uint32_t root;                    // order[0]
    uint32_t deg1;
    uint32_t prev_first[SGM_MAXQ + 1];
    uint32_t fol_first[SGM_MAXQ + 1];
    uint8_t prev_vtx[SGM_MAXPREV];
    uint8_t fol[SGM_MAXPREV];
    const uint32_t* tab_off[SGM_MAXPREV];  // rows by data vertex id, as in sgm_auxplan_dev
    const uint32_t* tab_idx[SGM_MAXPREV];
    const uint32_t* cand[SGM_MAXQ];
    const uint32_t* weight[SGM_MAXQ];
    const uint32_t* root_order;  // positions in cand[root] in degree order (the items' roots)
};

// The directed query edges (a, b_j), j < k, out of one query vertex a in the fill pass of an
// aux-table build (inc/table.h:7-137): rows = C(a); edge j's entries (positions in C(b_j),
// from b_j's bitmap and ranks) go to eidx[j] at its scanned row offsets rowoff[j]; vmask =
// the query's membership words (descriptors of many queries share one launch).  The
// reference's _edge_v (the candidate at each position) is not stored: C(b_j)[position].

// One row pass of computeWeightArray (inc/order.h:997-1029) in a batched launch: rows of the
// table (x -> child) given by toff / tidx, child weights `in`, x's weights `out`; in == null
// fills out[0, rows) with `fill` (the start value, or 1 for a vertex without a one-parent
// child).
struct sgm_wdesc {
    const uint32_t* toff;  // the table's row offsets by data vertex id
    const uint32_t* tidx;
    const uint32_t* rowv;  // C(x): the data vertex of each row
    const uint32_t* in;
    uint32_t* out;
    uint32_t rows, fill;
};

// One exclusive scan of the aux-table build: in[0, n) into out[0, n] (out[n] = total), the
// total to *tot and the maximum to *maxv (one block per scan, many scans per launch).
struct sgm_scandesc {
    const uint32_t* in;
    uint32_t* out;
    uint32_t* tot;
    uint32_t* maxv;
    uint32_t n, pad;
};

// One query of an aux-table build (inc/table.h:7-137): its row pool holds, for every query
// vertex b and data vertex v in U_b (v a candidate of some query neighbour of b), the row
// N(v) n C(b) as positions in C(b).  poff: n blocks of V + 1 words (row lengths, then the
// pool offsets of b); rrec: V records of rstride words (a multiple of 4 >= n), the position
// of w in every C(b); nbr[u]: the query neighbours of u as a bit mask; pool + pbase[b]: b's
// rows.
struct sgm_pooldesc {
    const uint32_t* vmask;
    uint32_t* poff;
    uint32_t* rrec;
    uint32_t* pool;
    const uint64_t* bm;    // n * W candidate bitmaps
    const uint32_t* rank;  // n * W bitmap word ranks
    uint32_t n, rstride;
    uint32_t nbr[SGM_MAXQ];
    unsigned long long pbase[SGM_MAXQ];
};

// One directed query edge (a, b) of an aux-table build: rows = C(a), b's pool offsets; its
// entry total and longest row go to slot `slot` of the totals and maxima.
// The depth-0 candidates of one query in degree order (descending, ties by id), as positions
// in C(u): the root walk of the aux-table and dynamic-order searches.
struct sgm_rootdesc {
    const uint32_t* vmask;
    const uint64_t* bm;    // C(u)'s bitmap words
    const uint32_t* rank;  // their ranks
    uint32_t* out;         // |C(u)| positions
    uint32_t u, pad;
};

struct sgm_edgedesc {
    const uint32_t* rows;
    const uint32_t* voff;
    uint32_t nrows, slot;
};

struct sgm_stats_dev {
    unsigned long long count;
    unsigned long long intersections;
    unsigned long long nodes[SGM_MAXQ];
    unsigned long long fs_prunes;
    unsigned long long elems_read;      // list elements fed to intersections (|A|+|B|)
    unsigned long long elems_out;       // intersection results (|A n B|), incl. the fused last level
    unsigned long long elems_scanned;   // iteration-set elements classified (filter + injectivity)
    unsigned long long overflow;        // per-wave stack overflows (result invalid if > 0)
    unsigned long long items_done;
    unsigned long long busy_ticks;      // wall-clock ticks its waves spent on the query's items
    unsigned long long timed_out;       // the waves' time passed the query's budget (count partial)
    unsigned long long probe_words;     // words the intersections read under the strategy used: keys
                                        // plus one bitmap word or ceil(log2 n) + 1 search words each
};

// Depth-1 splitting of the heaviest roots.  The first SGM_SPLIT_ROOTS candidate roots of a
// query in degree order (among the first SGM_SPLIT_SCAN positions) are cut into at most
// SGM_SPLIT_CHUNKS chunks of their depth-1 loop (whole 64-element batches), handed out
// before the remaining roots.  Depth-1 siblings interact only through the failing-set
// cut-off (a cut at depth 1 skips the rest of the loop; depth 0 never prunes,
// inc/prune.h:119-137), so each chunk keeps its counters in a record and k_split_reduce
// adds the chunks up to and including a root's first cutting chunk -- the counters of the
// sequential search.  Per-query split block, u32 words:
//   [0] n_split  [1] root_first (the first position not split)
//   [2, 2+R) the split roots (vertex ids)  [2+R, 2+2R) cutmin (smallest start of a
//   cutting chunk, ~0 = none)  [2+2R, 2+2R+IC) done flags  then IC triples (root, lo, hi)
// Records, u64, field-major (field f of item s at rec[f * IC + s]): nodes[32],
// intersections, fs_prunes, elems_read, elems_out, elems_scanned, probe_words.
#define SGM_SPLIT_ROOTS 64u
#define SGM_SPLIT_CHUNKS 32u  // 16 / 32 / 64: labelled Chung-Lu 10^5 launch 217 / 182 / 172 ms
#define SGM_SPLIT_ITEMS (SGM_SPLIT_ROOTS * SGM_SPLIT_CHUNKS)
#define SGM_SPLIT_SCAN 4096u
#define SGM_SPLIT_WORDS (2u + 2u * SGM_SPLIT_ROOTS + 4u * SGM_SPLIT_ITEMS)
#define SGM_SPLIT_FIELDS 38u
// queries with at most this many candidate roots are split by default (a launch's 4096-odd
// waves would otherwise sit idle behind their heaviest roots); more roots keep every wave
// busy without it, and the split variant of the search costs a few per cent
#define SGM_SPLIT_MAX_ROOTS 4096u

// One query of a batched cache + filter + failing-set enumeration.  Its depth-0 items are
// the prefix of length n_items of the degree-sorted vertex list: the split chunks (when
// `split` is set) first, handed out by `snext`, then the roots from root_first on, by `next`.
struct sgm_equery {
    const sgm_plan_dev* plan;
    const uint32_t* vmask;
    uint32_t* lastbits;             // ceil(V/32) words: bit v iff v in C(u_L), u_L = the plan's last
                                    // vertex (k_last_bits, before the search)
    uint32_t* stop;                 // raised when the cap is reached (or the run overflowed)
    sgm_stats_dev* st;
    uint32_t* next;                 // next depth-0 item of this query
    uint32_t* split;                // split block (SGM_SPLIT_WORDS words) or null: no splitting
    uint32_t* snext;                // next split chunk
    unsigned long long* rec;        // SGM_SPLIT_FIELDS * SGM_SPLIT_ITEMS chunk records
    unsigned long long cap;
    unsigned long long budget;      // wall-clock ticks of wave time the query may use (0 = no limit)
    uint32_t n_items;
    uint32_t root_u;
    uint32_t root_lo, root_hi;      // only depth-0 candidates v with root_lo <= v < root_hi are searched
};

// The host's stop request lives in pinned host memory and is written by a plain CPU store
// (a queued copy could not start while a persistent grid holds every CU).  Waves poll the
// device stop word; each wave copies the host word into it at most once per
// SGM_RELAY_TICKS of the device clock (100 MHz: 2.6 ms), so a full grid of 4096 waves reads
// host memory about 1.6 million times per second.  Between items a wave looks at the clock
// for this only every 64 item attempts (most attempts skip a non-candidate at once).
#define SGM_RELAY_TICKS (1ull << 18)
// A wave adds its time on a query to the query's total (the budget check) after at most
// SGM_CHARGE_TICKS of unpublished time (100 MHz: 82 us).
#define SGM_CHARGE_TICKS (1ull << 13)

// Refinement program of a BFS-tree filter (CFL / CECI / DAF), one per query: word 0 = step
// count, then per step (bit | source << 8, pivot mask) -- see k_refine.  At most 3 (n - 1)
// steps for n <= 32 query vertices.
#define SGM_RPROG_WORDS 256u
#define SGM_RPROG_MAXSTEPS ((SGM_RPROG_WORDS - 1u) / 2u)

// Per-query description used by the batched filter kernels.
//   flags[0 .. 2*SGM_MAXQ)   removals per GQL phase (round*n + u)
//   flags[SGM_GQL_TOUCH_ANY] bit u: some touch[] word has bit u (round-1 checks of u need it)
//   flags[SGM_GQL_DIRTY_ANY] bit u: some dirty[] word has bit u (round 2 must revisit u)
//   flags[SGM_GQL_WORK ..+3]  checks decided by the degree table, by a lane matcher, by the
//                             wave matcher, and vertices whose removal marked neighbours
#define SGM_GQL_TOUCH_ANY (2 * SGM_MAXQ)
#define SGM_GQL_DIRTY_ANY (2 * SGM_MAXQ + 1)
#define SGM_GQL_WORK (2 * SGM_MAXQ + 2)
#define SGM_GQL_FLAG_WORDS (2 * SGM_MAXQ + 6)
// GQL scratch per query, in words, by rank.  Summary words OR their group's words: level 1
// over SGM_GQL_GROUP ranks (G1 = V/G words), level 2 over 64 level-1 words (G2 = G1/64):
//   vmr[V] | touch[V] | dirty[V] | tsum[G1] | dsum[G1] | tsum2[G2] | dsum2[G2] | f0[V] | fsum[G1] | fsum2[G2]
// (touch .. dsum2 start clear; the rest is written before it is read)
#define SGM_GQL_GROUP 4u  // 2, 8, 16, 32 measured slower on power-law graphs
#define SGM_GQL_GROUP2 64u
inline size_t sgm_gql_groups(uint32_t V) { return ((size_t)V + SGM_GQL_GROUP - 1) / SGM_GQL_GROUP; }
inline size_t sgm_gql_groups2(uint32_t V) { return (sgm_gql_groups(V) + SGM_GQL_GROUP2 - 1) / SGM_GQL_GROUP2; }
inline size_t sgm_gql_scratch_words(uint32_t V) { return 4 * (size_t)V + 3 * sgm_gql_groups(V) + 3 * sgm_gql_groups2(V); }
inline size_t sgm_gql_clear_words(uint32_t V) { return 2 * (size_t)V + 2 * sgm_gql_groups(V) + 2 * sgm_gql_groups2(V); }
// GQL runs in degree-rank space (rank r = position in the degree-sorted vertex list), where
// the candidates of every NLF threshold are a prefix; `vmr` holds the rank-space membership
// words and is copied back to vmask (vertex ids) after the last phase.
struct sgm_qdesc {
    uint64_t req[SGM_MAXQ][4];       // byte j of req[u]: (j+1)-th largest degree in N_q(u), 0 past deg_q(u)
    uint32_t n;
    uint32_t qdeg[SGM_MAXQ];
    uint32_t nbr[SGM_MAXQ];          // neighbour bitmask of each query vertex
    uint32_t* vmask;                 // V words, by vertex id
    uint32_t* removed;               // SGM_GQL_FLAG_WORDS words (see above)
    uint32_t* vmr;                   // V words of GQL scratch: membership by degree rank
    uint32_t* touch;                 // V words of GQL scratch (by rank): a neighbour lost a bit of N_q(u)
    uint32_t* dirty;                 // V words of GQL scratch (by rank): recheck (round 2, u)
    uint32_t* f0;                    // V words (by rank): bit u = fails the NLF table test of u
    uint32_t* fsum;                  // per group of SGM_GQL_GROUP ranks: OR of f0
    uint32_t* tsum;                  // per group: OR of touch
    uint32_t* dsum;                  // per group: OR of dirty
    uint32_t* fsum2;                 // per SGM_GQL_GROUP2 groups: OR of fsum
    uint32_t* tsum2;                 // ... of tsum
    uint32_t* dsum2;                 // ... of dsum
    // Labelled NLF / GQL (lab_on = the data graph or the query carries a label other than 0).
    // lreq: first[n+1] | label[P] | count[P] (P = first[n]): u's neighbour labels with their
    // multiplicities, ascending by label, in entries [first[u], first[u+1]).
    uint32_t lab_on;
    uint32_t qlab[SGM_MAXQ];
    const uint32_t* lreq;
};

// Words of a query's lreq block (see sgm_qdesc) for n vertices and adjacency length A.
inline size_t sgm_lreq_words(uint32_t n, size_t A) { return (size_t)n + 1 + 2 * A; }

// Points d's GQL scratch at `base` (sgm_gql_scratch_words(V) words, layout above).
inline void sgm_gql_layout(sgm_qdesc& d, uint32_t* base, uint32_t V)
{
    const size_t G = sgm_gql_groups(V), G2 = sgm_gql_groups2(V);
    d.vmr = base;
    d.touch = d.vmr + V;
    d.dirty = d.touch + V;
    d.tsum = d.dirty + V;
    d.dsum = d.tsum + G;
    d.tsum2 = d.dsum + G;
    d.dsum2 = d.tsum2 + G2;
    d.f0 = d.dsum2 + G2;
    d.fsum = d.f0 + V;
    d.fsum2 = d.fsum + G;
}
