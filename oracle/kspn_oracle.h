/*
 * kspn_oracle.h -- CPU oracle for the SCC path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product (libkspn) never links, loads or calls it.
 *
 * Every function restates the reference algorithm it cites (paths are
 * relative to the rawbby/KaSpan source tree).  Types follow the reference
 * built with KASPAN_INDEX64: vertex ids uint32, CSR offsets uint64
 * (kaspan/include/kaspan/graph/base.hpp:17-28).
 */
#ifndef KSPN_ORACLE_H
#define KSPN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_UNDECIDED 0xFFFFFFFFu

/* ---- counter-based random numbers (SplitMix64 output i of seed) ---------- */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i);
uint32_t orc_perm_pow2(uint64_t seed, uint32_t bits, uint32_t x);

/* ---- generators: edge lists, self-loops dropped, duplicates kept --------- */
uint64_t orc_gen_gnm(uint64_t seed, uint32_t n, uint64_t draws, uint32_t* src, uint32_t* dst);
uint64_t orc_gen_rmat(uint64_t seed, uint32_t scale, uint64_t draws, double a, double b, double c,
                      int permute, uint32_t* src, uint32_t* dst);
uint64_t orc_gen_lattice(uint64_t seed, uint32_t side, double p_bidir, uint32_t* src, uint32_t* dst);
uint64_t orc_lattice_max_edges(uint32_t side);
/* planted-SCC fuzzy graph in the shape of test/fuzzy.hpp:81-218; labels = planted min ids */
uint64_t orc_gen_fuzzy(uint64_t seed, uint32_t n, double d, uint32_t* src, uint32_t* dst,
                       uint64_t cap, uint32_t* labels);
/* power-law planted SCC graph (component sizes, Hamiltonian cycles, chords,
 * forward-only inter-component edges); labels = planted canonical labels */
/* R-MAT on any n >= 1: ceil(log2 n) levels, an id x >= n folds to x - n,
 * then the cycle-walking bijection of [0, n) scatters the ids */
uint64_t orc_gen_rmat_n(uint64_t seed, uint32_t n, uint64_t draws, double a, double b, double c, uint32_t* src,
                        uint32_t* dst);
/* batch of independent graphs: per graph its kind (0 G(n,m), 1 R-MAT), n,
 * draws and seed; n and the average out-degree are log-uniform */
void orc_batch_plan(uint64_t seed, uint32_t count, uint32_t n_min, uint32_t n_max, double deg_min, double deg_max,
                    double p_rmat, uint32_t* kind, uint32_t* n, uint64_t* draws, uint64_t* gseed);
uint64_t orc_gen_planted(uint64_t seed, uint32_t n, double expo, uint32_t max_size, double intra_deg,
                         double inter_deg, uint32_t* src, uint32_t* dst, uint64_t cap,
                         uint32_t* labels);
/* labels of orc_gen_planted without generating its edges */
void orc_planted_labels(uint64_t seed, uint32_t n, double expo, uint32_t max_size, uint32_t* labels);
uint64_t orc_planted_sizes(uint64_t seed, uint32_t n, double expo, uint32_t max_size, uint32_t* sizes,
                           uint64_t cap);

/* ---- CSR ------------------------------------------------------------------ */
/* sort by (src,dst), drop self-loops and duplicates; returns m; fw_off has n+1 entries */
uint64_t orc_csr_from_edges(uint32_t n, uint64_t count, const uint32_t* src, const uint32_t* dst,
                            uint64_t* off, uint32_t* adj);
/* backward_complement_graph, scc/backward_complement.hpp:22-68 */
void orc_transpose(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, uint64_t* bw_off,
                   uint32_t* bw_adj);
/* induced subgraph on mask, allgather_sub_graph.hpp:26-126 (single rank); returns sub_m */
uint64_t orc_subgraph(uint32_t n, const uint64_t* off, const uint32_t* adj, const uint8_t* mask,
                      uint32_t* super_ids, uint32_t* sub_n, uint64_t* sub_off, uint32_t* sub_adj);

/* ---- SCC ------------------------------------------------------------------ */
/* iterative Tarjan, scc/tarjan.hpp:20-122; labels = minimum member id */
uint32_t orc_tarjan(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t* labels);
/* orc_tarjan for each graph of a batch union (vtx_off[count+1]); local labels; on `threads` threads */
void orc_tarjan_batch(uint32_t count, const uint32_t* vtx_off, const uint64_t* off, const uint32_t* adj,
                      uint32_t* labels, uint32_t threads);
/* normalise_scc_id, test/normalise_scc.hpp:10-25 */
void orc_canonicalize(uint32_t n, uint32_t* labels);
/* mutual reachability by BFS from every vertex (n <= 4096) */
int orc_brute_force(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t* labels);

typedef struct orc_scc_stats {
  uint32_t pivot;
  uint32_t decided_trim_first;
  uint32_t decided_fwbw;
  uint32_t decided_trim_normal;
  uint32_t decided_color;
  uint32_t color_rounds;
  uint32_t decided_residual;
  uint32_t residual_n;
  uint64_t residual_m;
} orc_scc_stats;

/* trim_1_first, scc/trim_1_local.hpp:12-46; live[k]=1 for undecided; returns pivot (n if none) */
uint32_t orc_trim_1_first(uint32_t n, const uint64_t* fw_off, const uint64_t* bw_off, uint8_t* live,
                          uint32_t* scc_id);
/* forward_backward_search, scc/pivot_search.hpp:15-85; returns number decided */
uint32_t orc_fwbw(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                  const uint32_t* bw_adj, uint8_t* live, uint32_t pivot, uint32_t* scc_id);
/* one cache_rot_label_search_pull round, scc/cache_multi_pivot_search_pull.hpp:40-266 */
uint32_t orc_color_round(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                         const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                         uint32_t rotation, uint32_t* color, uint32_t* scc_id);
/* kaspan::scc, scc/scc.hpp:21-87 (single rank); raw (non-canonical) ids.
 * faithful != 0 also allocates and clears the ghost caches the reference
 * allocates at one rank (scc.hpp:53,77,82). */
int orc_kaspan_scc(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                   const uint32_t* bw_adj, uint32_t* scc_id, int faithful, orc_scc_stats* st);
/* scc_trim_ex_residual, scc/variant/scc_trim_ex_residual.hpp:17-128 (single rank) */
int orc_scc_trim_ex_residual(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                             const uint64_t* bw_off, const uint32_t* bw_adj, uint32_t* scc_id,
                             orc_scc_stats* st);
/* scc_trim_ex, scc/variant/scc_trim_ex.hpp:16-100 (single rank, push colouring) */
int orc_scc_trim_ex(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                    const uint32_t* bw_adj, uint32_t* scc_id, orc_scc_stats* st);

/* scc_hpc_like (2), scc_ispan_like (3), scc_hpc_trim_ex (4),
 * scc_trim_ex_light_residual (6) at one rank; raw ids; -1 for another number */
int orc_scc_variant(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                    const uint32_t* bw_adj, int variant, uint32_t* scc_id, orc_scc_stats* st);
/* fixpoint of trim-1 (no live in- or no live out-neighbour) from live[]; returns the count removed */
uint32_t orc_trim_fixpoint(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                           const uint32_t* bw_adj, uint8_t* live);

/* ---- learned pivot scorer: FP64 reference of the MLP ---------------------- */
void orc_pivot_features(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                        const uint64_t* bw_off, const uint32_t* bw_adj, float* feat);
void orc_mlp_forward(uint32_t rows, const float* x, const double* w1, const double* b1,
                     const double* w2, const double* b2, const double* w3, double b3, double* out);
double orc_mlp_loss_grad(uint32_t rows, const float* x, const float* target, const double* w1,
                         const double* b1, const double* w2, const double* b2, const double* w3,
                         double b3, double* gw1, double* gb1, double* gw2, double* gb2, double* gw3,
                         double* gb3);

#ifdef __cplusplus
}
#endif
#endif
