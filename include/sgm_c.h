/*
 * sgm_c.h -- C ABI of the MI355X (gfx950) filter -> order -> enumerate engine.
 *
 * Plain C types only: host buffers are caller-owned with explicit sizes; device objects
 * (context, resident data graph, query with its candidate space) are opaque,
 * library-owned handles.  Every entry point returns 0 (SGM_OK) or a negative SGM_E*
 * code and never exits the process; sgm_ctx_last_error() holds a message.  One context
 * per host thread; calls on one context are not thread-safe except sgm_stop().
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * Orrinn/SubgraphMatching source tree).  Graphs and queries may carry vertex labels (all 0
 * when none are set); queries have 2..32 vertices.
 */
#ifndef SGM_C_H
#define SGM_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGM_ABI_VERSION 4

enum {
    SGM_OK = 0,
    SGM_EINVAL = -1,    /* bad argument or malformed graph */
    SGM_EHIP = -2,      /* HIP runtime error */
    SGM_ENOMEM = -3,    /* device allocation failed */
    SGM_EEMPTY = -4,    /* some candidate set is empty (reference filters return false) */
    SGM_EORDER = -5,    /* order is not a connected permutation (reference PlanIR exits) */
    SGM_ESTATE = -6,    /* call out of sequence (e.g. enumerate before filter) */
    SGM_ESTOPPED = -7,  /* sgm_stop() interrupted the search: counts are partial */
    SGM_EDEVICE = -8,   /* a device-side capacity check tripped */
    SGM_ENODEV = -9     /* no usable GPU */
};

enum {
    SGM_FILTER_NLF = 1,   /* Filter_NLF (inc/filter.h:1049) */
    SGM_FILTER_GQL = 2,   /* Filter_GQL (inc/filter.h:1057) */
    SGM_FILTER_CFL = 3,   /* Filter_CFL (inc/filter.h:1065, _CFL_Filter :481) */
    SGM_FILTER_CECI = 4,  /* Filter_CECI (inc/filter.h:1073, _CECI_Filter :602); where the reference
                             ends the process on an empty set (:700-703) this returns SGM_EEMPTY */
    SGM_FILTER_DAF = 5    /* Filter_DAF (inc/filter.h:1081, _DAF_Filter :942) */
};

enum {
    SGM_ENUM_CACHE_FILTER_FS = 0, /* explore_Intersection_Cache_DAF_Filter (inc/explore.h:1073) */
    SGM_ENUM_CACHE_FILTER = 1,    /* same search without failing-set pruning */
    SGM_ENUM_AUX_TABLES = 2,      /* explore_Intersection_All<false,false,false,false,false> (inc/explore.h:2757) */
    SGM_ENUM_AUX_TABLES_FS = 3,   /* explore_Intersection_All<true,false,false,false,false>: the same with DAF
                                     failing sets (inc/prune.h:2099-2260, 82-171) */
    SGM_ENUM_DYN_DAF = 4,         /* explore_Intersection_DynamicOrder_DAF (inc/explore.h:1849-1931): the order
                                     argument is the start (BFS) order; the next vertex is chosen per search
                                     node by the DAF weights of sgm_daf_weights / sgm_set_daf_weights for that
                                     order (computed in the call when none are resident for it) */
    SGM_ENUM_DYN_VEQ = 5,         /* explore_Intersection_DynamicOrder_VEQ (inc/explore.h:1993-2062): the same,
                                     ranked by candidate counts */
    SGM_ENUM_AUX_TABLES_GUPCD = 6,    /* explore_Intersection_All<false,false,true,false,false>: the aux-table
                                         search with GuP's conflict detection (inc/prune.h:488-923): a candidate
                                         whose reservation guard is wholly matched is pruned; the guards (up to
                                         3 vertices, gupConflictDetectionSize) are computed on the host for the
                                         order in the call and kept; prunes are counted in sgm_stats.fs_prunes */
    SGM_ENUM_AUX_TABLES_FS_GUPCD = 7  /* explore_Intersection_All<true,false,true,false,false>: both; a pruned
                                         candidate gives a full failing set (inc/prune.h:2296-2300);
                                         fs_prunes counts failing-set cut-offs and guard prunes together */
};

typedef struct sgm_ctx sgm_ctx;
typedef struct sgm_graph sgm_graph;
typedef struct sgm_query sgm_query;

typedef struct {
    uint64_t cap;          /* output limit: count = min(#embeddings, cap); UINT64_MAX = none */
    int32_t method;        /* SGM_ENUM_* */
    uint32_t blocks;       /* persistent blocks of 4 waves; 0 = automatic */
    uint32_t stack_words;  /* per-wave set-op stack (u32 words); 0 = automatic */
    uint32_t flags;        /* SGM_ENUM_F_* */
    uint32_t hub_min_degree; /* intersect against adjacency bitmaps only for lists of at least
                                this many vertices (0 = 64, the smallest with a bitmap) */
    uint32_t time_limit_ms;  /* per-query work budget (0 = none): a query's search stops once the
                                waves working on it have spent this many milliseconds of device
                                time on it, summed over the waves; its result is then
                                sgm_stats.timed_out and its count is not a result (the per-query
                                timeout of src/overall_performance.cpp:290-312, counted in device
                                worker time as the reference counts thread CPU time,
                                inc/explore.h:1093,1126-1127) */
    uint32_t root_begin;     /* search only the depth-0 candidates v with root_begin <= v < root_end */
    uint32_t root_end;       /* (vertex ids; root_end == 0: every candidate).  The depth-0 loop
                                never prunes and each subtree's failing-set slots are reset
                                before they are read (inc/prune.h:119-137, inc/explore.h:
                                1030-1066), so counts and statistics over disjoint ranges add up
                                to those of the whole search: a query splits over launches or
                                GPUs by ranges (no reference counterpart) */
    uint32_t split_chunks;   /* depth-1 splitting of the heaviest roots (cache + filter methods):
                                at most this many chunks of a root's depth-1 loop, 1..32 (0 = 32);
                                counts and statistics do not depend on it */
    uint32_t reserved[1];
} sgm_enum_opts;

/* sgm_enum_opts.flags: intersect against hub adjacency lists by binary search instead of
 * their bitmaps (same results; for measurement) */
#define SGM_ENUM_F_NO_BITMAPS 1u
/* sgm_enum_opts.flags: evaluate the last level one extension at a time instead of a chunk of
 * extensions per pass (same results and statistics; for measurement) */
#define SGM_ENUM_F_SERIAL_LAST 2u
/* sgm_enum_opts.flags: by default, a query with at most 4096 candidate roots has the depth-1
 * loops of its 64 highest-degree roots split into chunks for other waves.  NO_SPLIT searches
 * every root in one wave; SPLIT_ALL splits every query (same results and statistics either
 * way; for measurement) */
#define SGM_ENUM_F_NO_SPLIT 4u
#define SGM_ENUM_F_SPLIT_ALL 8u

typedef struct {
    uint64_t count;          /* min(#embeddings, cap) */
    uint64_t intersections;  /* set intersections (Profiler::total_intersection semantics) */
    uint64_t nodes[32];      /* extensions per depth */
    uint64_t fs_prunes;      /* failing-set sibling cut-offs */
    uint64_t elems_read;     /* list elements fed to intersections, sum of |A|+|B| */
    uint64_t elems_out;      /* intersection results, sum of |A n B| */
    uint64_t elems_scanned;  /* iteration-set elements tested against filter and injectivity */
    double kernel_ms;        /* enumeration kernel time (HIP events) */
    int32_t capped;          /* the cap was reached */
    int32_t stopped;         /* sgm_stop() interrupted: count is partial */
    uint32_t retries;        /* re-runs with a larger set-op stack */
    int32_t timed_out;       /* the query passed sgm_enum_opts.time_limit_ms: count is partial */
    uint64_t busy_ns;        /* cache methods: wave time spent on this query, summed over the
                                waves (a work measure of the order; no reference counterpart) */
    uint64_t probe_words;    /* words the intersections read under the strategy used: the keys of
                                the shorter list plus, per key, one bitmap word or the ceil(log2 n)
                                + 1 words of a binary search in a list of n */
} sgm_stats;

/* ---- context ------------------------------------------------------------------------ */
const char* sgm_version(void);
int sgm_device_count(int* count);
int sgm_ctx_create(int device, sgm_ctx** out);
/* Also destroys the context's batches and cost models that are still alive (their handles
 * become invalid).  Graphs and queries are destroyed by their own calls. */
void sgm_ctx_destroy(sgm_ctx* ctx);
const char* sgm_ctx_last_error(const sgm_ctx* ctx);
/* Device time (HIP events) of the last filter call on this context: the NLF kernel, the
 * refinement (the GQL phases, or the CFL / CECI / DAF steps with their NLF counts) and the
 * candidate bitmap/count kernels (no reference
 * counterpart: the reference times filters with a host clock,
 * src/optimizations_ablation.cpp:64-82). */
int sgm_ctx_filter_ms(const sgm_ctx* ctx, double* nlf_ms, double* gql_ms, double* cand_ms);
/* Cooperative stop of a running enumeration (replaces the std::atomic<bool>& stop_flag
 * polled by the reference, inc/explore.h:1057, src/overall_performance.cpp:302-305).  The
 * request is a plain store into a word of pinned host memory that the running kernels poll
 * (no GPU command is queued, so it reaches a grid that occupies the whole GPU); it takes
 * effect within a few milliseconds.  Safe to call from any host thread. */
int sgm_stop(sgm_ctx* ctx);

/* ---- data graph (replaces Graph::Load's CSR, inc/graph.h:190-266) ----------------- */
/* offsets[V+1], adj[adj_len]: undirected simple graph, every list strictly ascending.  The
 * graph stays resident in HBM until sgm_graph_destroy. */
int sgm_graph_upload(sgm_ctx* ctx, const uint32_t* offsets, const uint32_t* adj, uint32_t num_vertices,
                     uint64_t adj_len, sgm_graph** out);
void sgm_graph_destroy(sgm_graph* g);
uint32_t sgm_graph_max_degree(const sgm_graph* g);

/* ---- query and candidate space (replaces Query::Load + CandidateParam, inc/api.h:9-46) */
int sgm_query_create(sgm_graph* g, const uint32_t* qoff, const uint32_t* qadj, uint32_t n, sgm_query** out);
void sgm_query_destroy(sgm_query* q);

/* ---- vertex labels (replaces the labels of Graph::Load / Query, inc/graph.h:190-266, and
 * the NLF maps and label index built from them, inc/graph.h:76-112) --------------------
 * labels[num_vertices]: u32 values.  NULL or all zero makes the graph unlabeled again.  The
 * filters (NLF, GQL) and the plan honour labels when the graph or the query has any other
 * than 0; enumeration follows from the candidate sets, as in the reference's cache path
 * (inc/explore.h:200-235 intersects whole neighbour lists).  Set labels before filtering:
 * a query's labels reset its filter state; batches take query labels through
 * sgm_batch_filter_labeled. */
int sgm_graph_set_labels(sgm_graph* g, const uint32_t* labels, uint32_t num_vertices);
int sgm_query_set_labels(sgm_query* q, const uint32_t* labels, uint32_t n);

/* Filter_NLF / Filter_GQL / Filter_CFL / Filter_CECI / Filter_DAF (inc/filter.h:1049-1087;
 * _NLF_Filter :80-93, _GQL_Filter :233-313, _CFL_Filter :481-583, _CECI_Filter :602-920,
 * _DAF_Filter :942-1020).
 * Returns SGM_OK, or SGM_EEMPTY when some C(u) is empty (the reference returns false). */
int sgm_filter(sgm_query* q, int kind);
/* Same for many queries of one graph in one set of launches; status[i] gets each result. */
int sgm_filter_batch(sgm_query* const* qs, uint32_t nq, int kind, int32_t* status);
/* Work of the last GQL filter of q (no reference counterpart): work[0] checks decided by
 * the degree table, work[1] by a lane-level matcher, work[2] by the wave matcher, work[3]
 * removals. */
int sgm_gql_work(const sgm_query* q, uint64_t* work);
/* candidates_count[] (inc/api.h:12) */
int sgm_cand_counts(const sgm_query* q, uint32_t* counts);
/* candidates[u] in ascending vertex id (inc/api.h:11); out_cap >= counts[u] */
int sgm_cand_download(sgm_query* q, uint32_t u, uint32_t* out, uint32_t out_cap);
/* candidates_exsist (inc/api.h:17) as one word per data vertex: bit u set iff v in C(u) */
int sgm_cand_exist_download(sgm_query* q, uint32_t* words, uint32_t num_words);

/* buildTables (inc/table.h:7-137) for every directed query edge, on the device */
int sgm_build_tables(sgm_query* q);
/* edge_matrix[a][b] (inc/graph.h:13-49): e_cnt and _max_degree */
int sgm_table_info(sgm_query* q, uint32_t a, uint32_t b, uint64_t* e_cnt, uint32_t* max_degree);
/* edge_matrix[a][b]: _offset[|C(a)|+1], _edge[e_cnt], _edge_v[e_cnt] (the device keeps one row
 * pool per query vertex b, read by every edge into b; this call gathers the edge's layout) */
int sgm_table_download(sgm_query* q, uint32_t a, uint32_t b, uint32_t* offsets, uint32_t* idx, uint32_t* vtx);

/* ---- orders and plan (host; replaces Order_RI inc/order.h:1138, Order_GQL :563, PlanIR
 * inc/code_gen.h:699) ------------------------------------------------------------------ */
int sgm_order_ri(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint32_t* order);
int sgm_order_gql(const sgm_query* q, uint32_t* order);
/* Order_QSI (inc/order.h:610-668): from the candidate space of a filtered query; builds the
 * query's aux tables (sgm_build_tables) if they are missing. */
int sgm_order_qsi(sgm_query* q, uint32_t* order);
/* Order_CFL (inc/order.h:781-976): core paths, tree paths and leaves by estimated embeddings
 * over the aux tables (built if missing; the estimates run on the device). */
int sgm_order_cfl(sgm_query* q, uint32_t* order);
/* Order_CECI (inc/order.h:599-607): the BFS order from CECIStartVertex, which reads the NLF
 * counts of the query's vertices only, so q need not be filtered. */
int sgm_order_ceci(sgm_query* q, uint32_t* order);
/* Order_GuP (inc/order.h:537-561): 2-core by edge weights, then the tree part by path counts
 * over the aux tables (built if missing; count_embed runs on the device).  SGM_EINVAL for a
 * query whose vertices all have degree <= 2, where the reference asserts (:484-492). */
int sgm_order_gup(sgm_query* q, uint32_t* order);
/* Query::KCore core numbers (inc/graph.h:387-458) and a BFS tree (inc/common_ops.h:318-350):
 * host helpers of the orders, exported for inspection; the BFS returns the vertex count. */
int sgm_query_kcore(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, int32_t* core);
int sgm_query_bfs(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint32_t start, uint32_t* order,
                  uint32_t* parent, uint32_t* level);
/* Plan export for inspection: per op depth, parent (UINT32_MAX = none), v_depth, bitmask
 * of backward-neighbour depths; iter_set[n].  Arrays hold up to 512 ops. */
int sgm_plan_export(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, const uint32_t* order, uint32_t* n_ops,
                    uint32_t* depth, uint32_t* parent, uint32_t* vdepth, uint32_t* emask, uint32_t* iter_set);
/* The same for a labelled query (labels[n], NULL = all 0): ops are keyed by the follow
 * vertex's label (inc/code_gen.h:640-654, 777-808). */
int sgm_plan_export_labeled(const uint32_t* qoff, const uint32_t* qadj, const uint32_t* labels, uint32_t n,
                            const uint32_t* order, uint32_t* n_ops, uint32_t* depth, uint32_t* parent,
                            uint32_t* vdepth, uint32_t* emask, uint32_t* iter_set);

/* ---- dynamic-order searches (explore_Intersection_DynamicOrder_DAF / _VEQ,
 * inc/explore.h:1674-2162; driver src/order_ablations.cpp:131-172) ---------------------
 * DAFStartVertex (inc/filter.h:922-940) of a query: the start of the BFS order both searches
 * take (sgm_query_bfs gives the order). */
int sgm_daf_start_vertex(sgm_query* q, uint32_t* start);
/* computeWeightArray (inc/order.h:978-1031) on the device: bfs_order[n] receives the BFS order
 * from DAFStartVertex; the weights stay resident with the query for SGM_ENUM_DYN_DAF over
 * that order, and weights[u] (counts[u] words each; weights == NULL: none) receives them.
 * q must be filtered; its aux tables are built if missing.  SGM_EEMPTY when some candidate
 * set is empty. */
int sgm_daf_weights(sgm_query* q, uint32_t* bfs_order, uint32_t* const* weights);
/* The caller's weight_array for start order `order` (counts[u] words per vertex), made the
 * resident weights of SGM_ENUM_DYN_DAF (the drop-in passes the weight_array it was given). */
int sgm_set_daf_weights(sgm_query* q, const uint32_t* order, const uint32_t* const* weights);
/* GUPConflictDetection::Init (inc/prune.h:828-905) for `order` (q filtered; its aux tables are
 * built if missing): goff[n+1] = the first record of each depth (records in depth order, one
 * per candidate position of the vertex at that depth); words (words_cap >= 4 * goff[n]) = 4
 * per record: the guard's size (0 = none), then its data vertices ascending.  The guards
 * stay resident for SGM_ENUM_AUX_TABLES_GUPCD over that order; goff / words may be NULL. */
int sgm_gup_guards(sgm_query* q, const uint32_t* order, uint32_t* goff, uint32_t* words, uint64_t words_cap);

/* ---- enumeration (replaces explore_Intersection_Cache_DAF_Filter inc/explore.h:1073,
 * explore_Intersection_All<false x5> :2757 and the dynamic-order searches :1849, :1993) -- */
int sgm_enumerate(sgm_query* q, const uint32_t* order, const sgm_enum_opts* opts, sgm_stats* out);
/* sgm_enumerate without allocating: no reference counterpart.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The cache search methods only (SGM_ENUM_CACHE_FILTER_FS /
 * SGM_ENUM_CACHE_FILTER, hub_min_degree <= 64).  Every per-call device buffer (plan,
 * descriptor, statistics, stop and item words, split blocks, the last-level bitmap and the
 * set-op stack) comes from one static device array of 256 MiB per GPU, held until the
 * process ends; the grid is sized so that the waves' stacks fit in it (a stack overflow
 * re-runs with deeper stacks on fewer waves).  A graph whose fixed part (about
 * V/8 + 1 MiB) or a stack_words request that does not fit is rejected with SGM_EINVAL.
 * All calls on one GPU share its array: calls that overlap on one GPU race. */
int sgm_enumerate_noalloc(sgm_query* q, const uint32_t* order, const sgm_enum_opts* opts, sgm_stats* out);
/* Many queries of one data graph in one persistent launch (every method: cache + filter
 * with or without failing sets, and the aux-table search, whose missing tables are built in
 * the call; no reference counterpart: the reference runs one query per call).  Each query
 * gets its own cap and its own sgm_stats; out[i].kernel_ms is the time of the whole launch.
 * orders == NULL or orders[i] == NULL: the RI order (Order_RI), computed in the call; for
 * the dynamic-order methods, the BFS order from DAFStartVertex. */
int sgm_enumerate_batch(sgm_query* const* qs, const uint32_t* const* orders, uint32_t nq, const sgm_enum_opts* opts,
                        sgm_stats* out);

/* ---- batched runner (no reference counterpart: the reference runs one query per call) --
 * Up to `slots` queries of one data graph per batch.  Per-slot device buffers are
 * allocated once, so a batch allocates nothing (except a rare stack-overflow rerun); the
 * filter and the enumeration of a batch each replay a hipGraph captured on first use.
 * Orders are chosen on the host between the two calls: RI (orders == NULL or
 * orders[i] == NULL), the GQL order from the batch's candidate counts, or any connected
 * order of the caller's (a learned one).  Calls on one batch are not thread-safe. */
typedef struct sgm_batch sgm_batch;
int sgm_batch_create(sgm_graph* g, uint32_t slots, sgm_batch** out);
void sgm_batch_destroy(sgm_batch* b);
/* Load nq <= slots queries (CSR each, 2..32 vertices) and run NLF / NLF + GQL on them
 * (Filter_NLF / Filter_GQL, inc/filter.h:1049-1063).  status[i] = SGM_OK or SGM_EEMPTY;
 * optional cand_counts[i*32 + u] = |C(u)| and cand_degsum[i*32 + u] = sum of the data
 * degrees of C(u).  Candidate lists and tables are not built on this path. */
int sgm_batch_filter(sgm_batch* b, uint32_t nq, const uint32_t* const* qoffs, const uint32_t* const* qadjs,
                     const uint32_t* ns, int kind, int32_t* status, uint32_t* cand_counts, uint64_t* cand_degsum);
/* The same with query vertex labels (qlabs NULL, or qlabs[i] NULL: all 0); a labelled
 * graph (sgm_graph_set_labels) or labelled queries run the labelled NLF / GQL. */
int sgm_batch_filter_labeled(sgm_batch* b, uint32_t nq, const uint32_t* const* qoffs, const uint32_t* const* qadjs,
                             const uint32_t* const* qlabs, const uint32_t* ns, int kind, int32_t* status,
                             uint32_t* cand_counts, uint64_t* cand_degsum);
/* Order_GQL (inc/order.h:563-597) of loaded query i from the batch's candidate counts. */
int sgm_batch_order_gql(const sgm_batch* b, uint32_t i, uint32_t* order);
/* Enumerate the loaded batch (cache + filter methods) in one persistent launch; stats[i]
 * per query, stats[i].kernel_ms = the batch's enumeration time. */
int sgm_batch_enumerate(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* opts, sgm_stats* stats);
/* The same in two halves: _launch queues the enumeration on the batch's own stream and
 * returns at once; _wait blocks until it is done and fills stats.  Between the two the
 * host may filter and launch another batch of the same context, whose grid then runs on
 * the CUs this batch's waves leave as they run out of work.  The batch itself accepts no
 * other call in between (SGM_ESTATE).  A stop request (sgm_stop) ends every launch in
 * flight. */
int sgm_batch_enumerate_launch(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* opts);
int sgm_batch_enumerate_wait(sgm_batch* b, sgm_stats* stats);
/* Device time of the last sgm_batch_filter / sgm_batch_enumerate (HIP events). */
int sgm_batch_last_ms(const sgm_batch* b, double* filter_ms, double* enum_ms);
/* Autotuner: enumerates the loaded batch with each persistent grid size (2, 4, 8 blocks per
 * CU) and intersection strategy (adjacency bitmaps for lists of >= 64, 256, 1024 vertices,
 * or binary search only); then, from the fastest of those, with 8 and 16 work chunks per
 * heavy root (split_chunks; the default is 32) and with half and twice the automatic
 * per-wave stack (stack_words).  Checks that every setting gives the same counts and returns
 * the fastest in *best (cap and method from *base). */
int sgm_batch_autotune(sgm_batch* b, const uint32_t* const* orders, const sgm_enum_opts* base, sgm_enum_opts* best,
                       double* best_ms);

/* ---- learned-order cost model (no reference counterpart; the reference samples orders,
 * inc/order.h:1883-1959, but has no learned ranking) -----------------------------------
 * A 3-layer message-passing network over the query graph (hidden 64, BF16 operands on
 * MFMA with FP32 accumulation) that predicts log2(1 + search-tree nodes) of a (query,
 * order) pair from per-vertex features: query degree, |C(u)|, mean candidate degree and the
 * order position with its backward neighbours.  It only ranks orders; counts never depend
 * on it. */
typedef struct sgm_costmodel sgm_costmodel;
typedef struct {
    uint32_t n;                   /* query vertices (1..32) */
    const uint32_t* qoff;         /* query CSR */
    const uint32_t* qadj;
    const uint32_t* order;        /* a permutation of 0..n-1 */
    const uint32_t* cand_count;   /* n: |C(u)| */
    const uint64_t* cand_degsum;  /* n: sum of the data degrees over C(u) */
} sgm_cm_sample;
int sgm_costmodel_create(sgm_ctx* ctx, uint64_t seed, sgm_costmodel** out);
void sgm_costmodel_destroy(sgm_costmodel* m);
uint32_t sgm_costmodel_param_count(void);
/* Copy the fp32 parameters out (get) and/or in (set); each sgm_costmodel_param_count() floats. */
int sgm_costmodel_params(sgm_costmodel* m, float* get, const float* set);
/* The model input of one sample: feats[32*32] (row = order position), backward / forward
 * neighbour masks per row. */
int sgm_costmodel_features(const sgm_cm_sample* s, float* feats, uint32_t* back, uint32_t* fwd);
/* Predicted log2(1 + nodes) of each sample. */
int sgm_costmodel_predict(sgm_costmodel* m, const sgm_cm_sample* s, uint32_t count, float* out);
/* Mean squared error and its gradient over the samples (no update). */
int sgm_costmodel_grad(sgm_costmodel* m, const sgm_cm_sample* s, const float* target, uint32_t count, float* grad,
                       float* loss);
/* Adam training (targets: log2(1 + nodes)); epoch_loss[epochs] gets each epoch's mean loss. */
int sgm_costmodel_train(sgm_costmodel* m, const sgm_cm_sample* s, const float* target, uint32_t count,
                        uint32_t epochs, uint32_t minibatch, float lr, uint64_t seed, float* epoch_loss);
/* Accumulated device time (ms) and MFMA flops of training and prediction. */
int sgm_costmodel_timing(const sgm_costmodel* m, double* train_ms, double* train_flops, double* infer_ms,
                         double* infer_flops);
/* Learned orders for the loaded batch (RI fallback): per query, RI, the GQL order and
 * n_samples seeded sampled orders are scored by the model on the GPU; the best-scored order
 * is taken when its predicted log2 node count is below RI's by more than `margin`, else RI.
 * orders[i*32 ..] receives query i's order; optional pred_best / pred_ri (predictions of
 * the chosen and the RI order) and learned[i] (1 when the learned order was taken). */
int sgm_batch_rank_orders(sgm_batch* b, sgm_costmodel* m, uint32_t n_samples, uint64_t seed, float margin,
                          uint32_t* orders, float* pred_best, float* pred_ri, int32_t* learned);
/* A seeded random connected order of a query (the sampled orders the model ranks): start
 * at a uniformly chosen vertex, then repeatedly a uniformly chosen unvisited vertex adjacent
 * to the visited ones. */
int sgm_order_sample(const uint32_t* qoff, const uint32_t* qadj, uint32_t n, uint64_t seed, uint32_t* order);

#ifdef __cplusplus
}
#endif
#endif /* SGM_C_H */
