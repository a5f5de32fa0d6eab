// Launch wrappers of the gfx950 kernels (filters and tables in sgm_kernels.hip,
// enumeration in sgm_enum.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sgm_internal.h"

// Per-query candidate-space descriptor for the bitmap / rank / compaction kernels.
struct sgm_cdesc {
    uint32_t n;
    uint32_t* vmask;     // V
    uint64_t* bm;        // n * W
    uint32_t* rank;      // n * W
    uint32_t* cnt;       // n (written by k_rank_scan)
    uint32_t* cand_off;  // n+1 (host-provided prefix of cnt)
    uint32_t* list;      // sum of cnt
};

namespace sgm {
void launch_nlf(hipStream_t s, const uint32_t* off, uint32_t V, const sgm_qdesc* qs, uint32_t nq);
// GQL in degree-rank space (see k_gql_phase): the pre-pass once per filter, then one
// launch per phase (round * n + u) for the whole batch
void launch_gql_prepass(hipStream_t s, const uint64_t* dtab, uint32_t V, const uint32_t* ge_cnt, const sgm_qdesc* qs,
                        uint32_t nq);
void launch_gql_phase(hipStream_t s, const uint32_t* roff, const uint32_t* radj, uint32_t V, const uint32_t* ge_cnt,
                      const sgm_qdesc* qs, uint32_t nq, uint32_t phase);
// p[rep * stride + i] = val for rep < reps, i < len
void launch_fill_strided(hipStream_t s, uint32_t* p, size_t len, size_t stride, uint32_t reps, uint32_t val);
void launch_nlf_rank(hipStream_t s, const uint32_t* ge_cnt, uint32_t V, const sgm_qdesc* qs, uint32_t nq);
// labelled NLF (queries with lab_on; the unlabelled kernels skip those): by vertex id into
// vmask, or by degree rank into vmr with every member touched (the start of a labelled GQL)
void launch_nlf_lab(hipStream_t s, const uint32_t* off, const uint32_t* lab, const uint32_t* run_off,
                    const uint32_t* run_lab, const uint32_t* run_cnt, uint32_t V, const sgm_qdesc* qs, uint32_t nq);
void launch_nlf_rank_lab(hipStream_t s, const uint32_t* by_deg, const uint32_t* off, const uint32_t* lab,
                         const uint32_t* run_off, const uint32_t* run_lab, const uint32_t* run_cnt, uint32_t V,
                         const sgm_qdesc* qs, uint32_t nq);
void launch_unrank(hipStream_t s, const uint32_t* rank, uint32_t V, const sgm_qdesc* qs, uint32_t nq);
void launch_rank_adj(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* by_deg,
                     const uint32_t* rank, const uint32_t* roff, uint32_t V, uint32_t* radj);
// degree table for the GQL shortcut, by rank: 4 planes of V words, 32 bytes per vertex
void launch_nbr_deg(hipStream_t s, const uint32_t* roff, const uint32_t* radj, uint32_t V, uint64_t* dtab);
// batch candidate statistics: cnt[q*32+u] = |C(u)|, degsum[q*32+u] = sum of degrees (zeroed first)
void launch_cand_stats(hipStream_t s, const uint32_t* off, uint32_t V, const sgm_qdesc* qs, uint32_t nq,
                       uint32_t* cnt, unsigned long long* degsum);
// one launch each: clear every query's flag words; gather the candidate counts (32 per query)
void launch_clear_flags(hipStream_t s, const sgm_qdesc* qs, uint32_t nq);
void launch_gather_counts(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t* out);
void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t val);
void launch_bitmaps(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V);
void launch_rank_scan(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V, uint32_t max_n);
void launch_compact(hipStream_t s, const sgm_cdesc* cs, uint32_t nq, uint32_t V);
// aux tables of many queries at once: the row pools' lengths (one block row per query), one
// scan per segment (pool offsets per (query, b)), the edges' entry totals and longest rows,
// and the pools' entries
void launch_scan_segments(hipStream_t s, const sgm_scandesc* ds, uint32_t m);
// each query's depth-0 candidates in degree order, as positions in C(u) (one block per query)
void launch_root_order(hipStream_t s, const uint32_t* by_deg, uint32_t V, const sgm_rootdesc* ds, uint32_t nq);
void launch_pool_count(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_pooldesc* ds,
                       uint32_t nq, uint32_t kmax);
void launch_pool_fill(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_pooldesc* ds,
                      uint32_t nq);
void launch_edge_stats(hipStream_t s, const sgm_edgedesc* ds, uint32_t m, uint32_t max_rows, uint32_t* tot,
                       uint32_t* maxv);
// BFS-tree filters (CFL / CECI / DAF): step `step` of every query's refinement program
// (SGM_RPROG_WORDS words per query, see k_refine), and the copy of the NLF words into vmr
void launch_refine(hipStream_t s, const uint32_t* off, const uint32_t* adj, uint32_t V, const sgm_qdesc* qs,
                   uint32_t nq, const uint32_t* prog, uint32_t step);
void launch_copy_vmask(hipStream_t s, uint32_t V, const sgm_qdesc* qs, uint32_t nq);
// order estimates over an aux table: out[j] (= or *=) the row sum of in[idx] over the row of
// data vertex rowv[j] (in == null: row lengths), *total += all row sums, all mod 2^64
void launch_row_gather(hipStream_t s, const uint32_t* toff, const uint32_t* tidx, const uint32_t* rowv, uint32_t rows,
                       const unsigned long long* in, unsigned long long* out, int mul, unsigned long long* total);
void launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long val);
// DAF weights (computeWeightArray): one reverse-BFS step of many queries, m descriptors
// (out[j] = min(out[j], u32 sum of in[idx] over row j), or a fill), max_rows = the largest
void launch_weight_step(hipStream_t s, const sgm_wdesc* ds, uint32_t m, uint32_t max_rows);
// sgm_enum.hip
size_t enum_smem_bytes(uint32_t max_ops);
// dev_stop: the device stop word every wave polls; host_stop: the host's request (pinned host
// memory), copied into dev_stop by the polling waves; budget: wave-time ticks per query (0 = none);
// only depth-0 candidates v with root_lo <= v < root_hi are searched
void launch_enum_aux(hipStream_t s, uint32_t blocks, const sgm_auxplan_dev* plans, const uint32_t* item_first,
                     uint32_t nq, uint32_t* stack, uint32_t stack_words, uint32_t* next, sgm_stats_dev* st,
                     uint32_t* stops, uint32_t* dev_stop, const uint32_t* host_stop, unsigned long long cap,
                     unsigned long long budget, uint32_t root_lo, uint32_t root_hi, int use_fs,
                     const uint32_t* hub_slot, const uint32_t* hub_bits, uint32_t hub_words);
// the dynamic-order searches (DAF / VEQ ranking) over sgm_dynplan_dev plans, items as above
void launch_enum_dyn(hipStream_t s, uint32_t blocks, const sgm_dynplan_dev* plans, const uint32_t* item_first,
                     uint32_t nq, uint32_t* stack, uint32_t stack_words, uint32_t* next, sgm_stats_dev* st,
                     uint32_t* stops, uint32_t* dev_stop, const uint32_t* host_stop, unsigned long long cap,
                     unsigned long long budget, uint32_t root_lo, uint32_t root_hi, const uint32_t* hub_slot,
                     const uint32_t* hub_bits, uint32_t hub_words);
// bitmap of C(u_L) per query (Q.lastbits) from its candidate words; before launch_enum_cache
void launch_last_bits(hipStream_t s, const sgm_equery* qs, uint32_t nq, uint32_t V);
// depth-1 split tables (queries with Q.split set), before launch_enum_cache; and the chunk
// records added to the query statistics, after it
void launch_split_roots(hipStream_t s, const uint32_t* off, const uint32_t* items, const sgm_equery* qs, uint32_t nq,
                        uint32_t V, uint32_t max_chunks);
void launch_split_reduce(hipStream_t s, const sgm_equery* qs, uint32_t nq);
void launch_enum_cache(hipStream_t s, uint32_t blocks, const uint32_t* off, const uint32_t* adj, const uint32_t* items,
                       const sgm_equery* qs, uint32_t nq, uint32_t max_ops, uint32_t* stack, uint32_t stack_words,
                       uint32_t* dev_stop, const uint32_t* host_stop, int use_fs, const uint32_t* hub_slot,
                       const uint32_t* hub_bits, uint32_t hub_words, int serial_last, bool split);
void launch_hub_bits(hipStream_t s, const uint32_t* off, const uint32_t* adj, const uint32_t* hubs, uint32_t n_hubs,
                     uint32_t* bits, uint32_t words);
}  // namespace sgm
