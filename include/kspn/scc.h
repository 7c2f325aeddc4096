/*
 * kspn/scc.h -- C ABI of libkspn, a single-GPU (MI355X, gfx950) strongly
 * connected components solver with the staged method of rawbby/KaSpan.
 *
 * Plain C: fixed-width integers, plain pointers and sizes. `stream` is a
 * hipStream_t passed as void* (NULL = the default stream). Every call
 * returns 0 on success or a negative KSPN_E_* code; nothing aborts and no
 * C++ exception crosses this boundary.
 *
 * Graph contract (kaspan/include/kaspan/graph/bidi_graph_part.hpp:12-27,
 * graph/base.hpp:17-28 with KASPAN_INDEX64): n vertices with uint32 ids,
 * m directed edges; fw_off[n+1] / fw_adj[m] is the out-edge CSR and
 * bw_off[n+1] / bw_adj[m] its exact transpose (kspn_transpose builds it);
 * offsets are uint64 with off[0] = 0 and off[n] = m. Neighbours inside a
 * row need not be sorted; self-loops and duplicate edges are tolerated.
 *
 * Output convention: labels[v] is the canonical component id, the minimum
 * vertex id of v's SCC -- the reference's ids after normalise_scc_id
 * (kaspan/include/kaspan/test/normalise_scc.hpp:10-25).
 */
#ifndef KSPN_SCC_H
#define KSPN_SCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSPN_OK 0
#define KSPN_E_INVALID (-1)       /* bad argument or malformed CSR */
#define KSPN_E_OOM (-2)           /* device or host allocation failed */
#define KSPN_E_HIP (-3)           /* a HIP runtime call failed */
#define KSPN_E_ITER_CAP (-4)      /* max_rounds reached before every vertex was decided */
#define KSPN_E_UNIMPLEMENTED (-5) /* option combination not built */
#define KSPN_E_IO (-6)            /* a file could not be opened or read, or has the wrong size */

#define KSPN_UNDECIDED 0xFFFFFFFFu

/* opaque per-device context: scratch memory, HIP events, pinned counters */
typedef struct kspn_ctx kspn_ctx;

typedef struct kspn_graph {
  uint32_t n;
  uint64_t m;
  const uint64_t* fw_off; /* device, n+1 */
  const uint32_t* fw_adj; /* device, m */
  const uint64_t* bw_off; /* device, n+1 */
  const uint32_t* bw_adj; /* device, m */
} kspn_graph;

/* pipeline shapes, one per variant of the reference's driver
 * (tool/bench_kaspan/src/main.cpp:63-71), at one rank; all give identical
 * canonical labels. "residual" stages decide everything left. */
#define KSPN_METHOD_KASPAN 0  /* scc.hpp:21-87 (also async/scc.hpp): trim-1 first pass + pivot, FW-BW, trim-1 normal pass, colour rounds */
#define KSPN_METHOD_TRIM_EX 1 /* variant/scc_trim_ex.hpp: exhaustive trim-1 (+trim-2), FW-BW rounds, colour rounds with trims */
#define KSPN_METHOD_HPC_LIKE 2               /* variant/scc_hpc_like.hpp: trim-1 first pass, FW-BW, colour rounds */
#define KSPN_METHOD_ISPAN_LIKE 3             /* variant/scc_ispan_like.hpp: trim-1 first pass, FW-BW, two trim-1 normal passes, residual */
#define KSPN_METHOD_HPC_TRIM_EX 4            /* variant/scc_hpc_trim_ex.hpp: exhaustive trim-1, pivot, FW-BW, colour rounds */
#define KSPN_METHOD_TRIM_EX_RESIDUAL 5       /* variant/scc_trim_ex_residual.hpp: exhaustive trim-1, residual */
#define KSPN_METHOD_TRIM_EX_LIGHT_RESIDUAL 6 /* variant/scc_trim_ex_light_residual.hpp: exhaustive trim-1, trim after decisions, residual */

#define KSPN_PIVOT_PLAIN 0   /* max in*out degree product, smallest id on ties (trim_1_local.hpp:32-45) */
/* multi-pivot rounds: per partition, the vertices in the top half-octave of in*out are the candidates, and
 * the MLP score (kspn_mlp_*) picks among them; single pivot: the MLP picks among the top in*out vertices */
#define KSPN_PIVOT_LEARNED 1
#define KSPN_PIVOT_RANDOM 2  /* a seeded uniform choice among the same candidates (for collecting pivot outcomes) */

#define KSPN_FLAG_PROFILE 1u  /* time every kernel launch with HIP events (slower; fills kspn_stats.kern) */
#define KSPN_FLAG_NO_TRIM2 2u /* KSPN_METHOD_TRIM_EX without the trim-2 pass */
#define KSPN_FLAG_NO_TAIL 4u  /* never hand the rest to the single-workgroup tail solver */
#define KSPN_FLAG_NO_REFINE 8u /* colour rounds over the whole live graph, without partition refinement */
#define KSPN_FLAG_NO_COOP 16u  /* every FW/BW level as separate launches with a host round trip each (no chained or persistent levels) */
#define KSPN_FLAG_STATS_MEMORY 32u /* fill kspn_phase.memory (one hipMemGetInfo per scope that carries it) */
/* multi-pivot FW-BW rounds search backward over the whole partition and split it in three (forward-only,
 * backward-only, unreached; Fleischer, Hendrickson and Pinar). Default: backward only within the forward
 * set, which is the pivot's SCC plus what it reaches, and a two-way split (forward-only, the rest). */
#define KSPN_FLAG_FULL_BW 64u
/* The device-driven solve: a fixed sequence of kernels (trim-1 first pass, one FW-BW from the plain pivot of
 * every graph, colour rounds to completion) whose persistent kernels loop over levels and rounds on the
 * device, with no host round trip until the labels are complete. It is captured into a HIP graph on the
 * first call and replayed while the graph, labels, batch offsets, alpha, beta and max_rounds stay the same.
 * Plain pivot and the KASPAN or HPC_LIKE method only (KSPN_E_UNIMPLEMENTED otherwise); kspn_stats then
 * holds ms_total and the decided counts per scope, not per-scope times or kernel profiles. */
#define KSPN_FLAG_GRAPH 128u
/* Every FW / BW reachability of the host-driven solve as one asynchronous launch: no levels; wavefronts
 * take the claimed vertices' row slices from a queue as they appear (coop_workgroups workgroups of 256
 * threads, 0 -> 512). The reached sets, and so the labels, are the same as level by level. */
#define KSPN_FLAG_ASYNC 256u
/* Push levels on the device (frontier degree sum <= coop_edges, largest degree <= coop_degree) run as chains
 * of chain_levels launches, one level each, enqueued back to back with a single host round trip per chain: no
 * grid barrier, and the coop_workgroups workgroups need not be resident together. A level that ends the run
 * (empty frontier, pull switch, frontier too large) makes the rest of its chain return at once. This is the
 * default; KSPN_FLAG_CHAIN asks for it explicitly (it wins over KSPN_FLAG_PERSISTENT). */
#define KSPN_FLAG_CHAIN 512u
/* Those push levels in one persistent launch instead, with a grid barrier (agent-scope release and acquire)
 * between levels; coop_workgroups must be resident together (capped at what the device holds). Same labels. */
#define KSPN_FLAG_PERSISTENT 1024u
/* Push levels whose frontier has at most 2^19 out-edges, none of its vertices more than 64, run on the device
 * as chains of chain_levels launches (0 -> 8) with no degree scan: eight lanes read one row, each lane claims
 * eight edges together. A larger frontier, a vertex of higher degree, the pull switch or an empty frontier
 * ends a chain. Taken before the chains of KSPN_FLAG_CHAIN / KSPN_FLAG_PERSISTENT. Same labels. */
#define KSPN_FLAG_NARROW 2048u
/* An edge that a forward push level of the multi-pivot rounds, a colour pull sweep or a colour push finds joining
 * two partitions (or reaching a decided vertex) is marked in a bitmap of one bit per adjacency entry, and later
 * levels and sweeps of the solve skip it before reading its other end's state (partitions only split and decided
 * vertices stay decided, so such an edge never again joins two vertices of one partition). The context keeps
 * m / 8 bytes per direction for it, cleared at the start of each solve. Same labels. */
#define KSPN_FLAG_PRUNE 4096u
/* Push levels of a reach whose frontier has no vertex of out-degree above 64 (known from the previous level),
 * and every push level of the colour rounds, run without a degree scan: each workgroup takes 256 consecutive
 * frontier entries, scans their degrees in LDS and processes their edges 2048 at a time. Taken after the narrow
 * chains of KSPN_FLAG_NARROW. Suits graphs without high-degree vertices (a chunk holding a hub is one
 * workgroup's work). Same labels. */
#define KSPN_FLAG_CHUNKED 8192u

typedef struct kspn_opts {
  uint32_t method;             /* KSPN_METHOD_* */
  uint32_t pivot;              /* KSPN_PIVOT_* */
  uint32_t fwbw_rounds;        /* FW-BW rounds before colouring (TRIM_EX); 0 -> 1 */
  uint32_t residual_threshold; /* hand the rest to residual Tarjan once live <= this (0: never) */
  uint32_t max_rounds;         /* colour-round cap (0 -> unlimited); exceeding it -> KSPN_E_ITER_CAP */
  uint32_t tail_threshold;     /* colour the rest in one workgroup once live <= this (0 -> 4096) */
  uint32_t flags;              /* KSPN_FLAG_* */
  float alpha;                 /* push->pull switch: edges(frontier) * alpha > edges(unvisited); 0 -> 14 */
  float beta;                  /* pull->push switch: frontier * beta < live; 0 -> 24 */
  const float* mlp_params;     /* device, KSPN_MLP_PARAMS floats, for KSPN_PIVOT_LEARNED */
  uint32_t* pivot_log;         /* device, KSPN_LOG_WORDS u32 per record, or NULL: outcome of every multi-pivot FW-BW pivot */
  uint32_t pivot_log_cap;      /* records that fit in pivot_log */
  uint32_t seed;               /* KSPN_PIVOT_RANDOM */
  uint32_t coop_edges;         /* persistent multi-level kernel takes frontiers with at most this many out-edges (0 -> 2^20) */
  uint32_t coop_degree;        /* ... and no vertex of out-degree above this (0 -> 2048) */
  uint32_t coop_workgroups;    /* its workgroups, at most one per CU (0 -> 128) */
  /* launch shapes (the autotuner's; 0 -> the default) */
  uint32_t pull_ilp;           /* pull sweep: vertices in flight per thread, 1, 2 or 4 (0 -> 2) */
  uint32_t expand_ept;         /* push expansion: edges per thread of a workgroup tile, 2, 4, 8 or 16 (0 -> per level) */
  uint32_t sweep_wgs;          /* pull sweep: workgroups per CU (0 -> 8, 16 above 2^25 vertices) */
  uint32_t list_wgs;           /* pull list levels: workgroups per CU at most (0 -> 4) */
  uint32_t chain_levels;       /* KSPN_FLAG_CHAIN: push levels enqueued per host round trip (0 -> 16, at most 256) */
  uint32_t learned_rounds;     /* KSPN_PIVOT_LEARNED, multi-pivot rounds: the first learned_rounds rounds take the
                                  learned pivot, the later ones the plain rule (0 -> every round) */
} kspn_opts;

/* pivot outcome record: pivot, round, partition size, SCC size (decided),
 * forward-only, backward-only and unreached vertices of the partition, 0 */
#define KSPN_LOG_WORDS 8

/* per-kernel profile line, filled with KSPN_FLAG_PROFILE */
#define KSPN_KERN_MAX 32
typedef struct kspn_kern_stat {
  char name[32];
  uint32_t launches;
  double ms;    /* summed launch durations (HIP events around each launch) */
  double bytes; /* summed algorithmic bytes (each array element touched once, stored width) */
} kspn_kern_stat;

/* one statistics scope of the reference (debug/statistic.hpp:124-205) as the
 * variant's code names it (scc.hpp:33-86, variant/scc_*.hpp): "trim_1_first",
 * "trim_ex_first", "trim_2", "select_pivot", "forward_backward_search",
 * "trim_1_normal", "trim_ex", "color", "color_trim_ex", "residual". A scope
 * entered more than once (trims between FW-BW rounds) accumulates. */
#define KSPN_PHASE_MAX 12
typedef struct kspn_phase {
  char name[32];
  double ms;           /* GPU time (HIP events on the solve's stream), summed over entries */
  uint64_t decided;    /* vertices the scope decided ("decided_count") */
  uint64_t memory;     /* device bytes in use when the scope ended (hipMemGetInfo), if has_memory */
  uint32_t entries;    /* times the scope was entered */
  uint32_t has_memory; /* the reference records "memory" for this scope */
  double bytes;        /* KSPN_FLAG_PROFILE: algorithmic bytes of the scope's kernel launches */
  double kernel_ms;    /* KSPN_FLAG_PROFILE: summed durations of those launches */
} kspn_phase;

typedef struct kspn_stats {
  double ms_total, ms_trim, ms_fwbw, ms_color, ms_residual, ms_canon;
  uint32_t decided_trim, decided_trim2, decided_fwbw, decided_color, decided_residual;
  uint32_t pivot;          /* first FW-BW pivot */
  uint32_t fwbw_levels_push, fwbw_levels_pull;
  uint32_t color_rounds, color_iters;
  uint32_t residual_n;
  uint32_t fwbw_rounds_done;  /* multi-pivot FW-BW rounds run */
  uint32_t partitions_max;    /* largest partition count of those rounds */
  uint32_t pivot_log_count;   /* records written to opts.pivot_log */
  uint32_t fwbw_levels_coop;  /* FW/BW push levels run inside the persistent multi-level kernel */
  uint64_t residual_m;
  uint64_t edges_scanned;  /* adjacency entries read, all phases */
  /* KSPN_FLAG_PROFILE: algorithmic bytes, summed over the scopes as ms_* are */
  double bytes_trim, bytes_fwbw, bytes_color, bytes_residual, bytes_canon;
  uint32_t n_kern;
  kspn_kern_stat kern[KSPN_KERN_MAX];
  uint32_t n_phase;                   /* scopes in order of first entry */
  kspn_phase phase[KSPN_PHASE_MAX];
} kspn_stats;

/* defaults: KASPAN method, plain pivot, 1 FW-BW round, no residual, no cap */
void kspn_opts_default(kspn_opts* o);
const char* kspn_strerror(int code);
/* the calling thread's last failing HIP call inside the library (call, HIP
 * error name, source location); "" if none */
const char* kspn_last_error(void);

/* ---- context ------------------------------------------------------------ */
/* Creates a context on `device`; scratch grows on demand to the largest graph seen. */
int kspn_ctx_create(kspn_ctx** out, int device);
int kspn_ctx_destroy(kspn_ctx* ctx);
/* pre-size scratch for graphs up to n vertices and m edges */
int kspn_ctx_reserve(kspn_ctx* ctx, uint32_t n, uint64_t m);

/* ---- the hot path ------------------------------------------------------- */
/* kaspan::scc(bidi_graph_part_view<single_part_view>, vertex_t* scc_id)
 * (kaspan/include/kaspan/scc/scc.hpp:21-25): fills labels[n] (device) with
 * canonical ids. Stream-ordered; returns after the labels are complete
 * (the host reads phase counters). `stats` may be NULL. */
int kspn_scc(kspn_ctx* ctx, const kspn_graph* g, uint32_t* labels, const kspn_opts* opts,
             kspn_stats* stats, void* stream);

/* the same with host-resident CSR arrays and host labels (copies in and out) */
int kspn_scc_host(kspn_ctx* ctx, uint32_t n, uint64_t m, const uint64_t* fw_off, const uint32_t* fw_adj,
                  const uint64_t* bw_off, const uint32_t* bw_adj, uint32_t* labels, const kspn_opts* opts,
                  kspn_stats* stats);

/* ---- batches of independent graphs ------------------------------------- */
/* A batch is the union of `count` graphs with batch-global ids: graph i owns
 * vertices [vtx_off[i], vtx_off[i+1]) and the edge slots [edge_off[i],
 * edge_off[i+1]) of fw_adj and bw_adj; its rows point only inside its own
 * vertex range (fw_off[vtx_off[i]] = edge_off[i]). Offsets are uint64, so
 * the union may hold more than 2^32 edges. There is no batch entry point in
 * the reference: it solves one graph per scc() call (scc/scc.hpp:21-25);
 * this is that call for every graph of the batch at once. */
typedef struct kspn_batch {
  uint32_t count;
  uint32_t n;               /* vtx_off[count] */
  uint64_t m;               /* edge_off[count] */
  const uint32_t* vtx_off;  /* device, count+1 */
  const uint64_t* edge_off; /* device, count+1 */
  const uint64_t* fw_off;   /* device, n+1 */
  const uint32_t* fw_adj;   /* device, m */
  const uint64_t* bw_off;   /* device, n+1 */
  const uint32_t* bw_adj;   /* device, m */
} kspn_batch;

/* kaspan::scc for every graph of the batch: labels[vtx_off[i] + v] is the
 * canonical id of v inside graph i (its SCC's smallest local id), the value
 * kspn_scc gives for graph i alone. Graph i is partition i of the
 * multi-pivot FW-BW rounds (one pivot per graph and round). Requires
 * count <= n. Stream-ordered; returns after the labels are complete. */
int kspn_scc_batch(kspn_ctx* ctx, const kspn_batch* b, uint32_t* labels, const kspn_opts* opts, kspn_stats* stats,
                   void* stream);

/* ---- stages callable on their own ---------------------------------------- */
/* backward_complement_graph (kaspan/include/kaspan/scc/backward_complement.hpp:22-68):
 * reverse CSR with ascending sources inside each row. Device pointers. */
int kspn_transpose(kspn_ctx* ctx, uint32_t n, uint64_t m, const uint64_t* fw_off, const uint32_t* fw_adj,
                   uint64_t* bw_off, uint32_t* bw_adj, void* stream);

/* normalise_scc_id (kaspan/include/kaspan/test/normalise_scc.hpp:10-25) for
 * ids that are vertex ids: labels[v] <- min { u : labels[u] == labels[v] }. */
int kspn_canonicalize(kspn_ctx* ctx, uint32_t n, uint32_t* labels, void* stream);

/* allgather_fw_sub_graph at one rank (kaspan/include/kaspan/scc/allgather_sub_graph.hpp:253-301):
 * induced subgraph on mask[v] != 0; sub ids ascend with global ids; an edge
 * is kept iff both ends are in the subgraph, in CSR order. super_ids[n],
 * sub_off[n+1], sub_adj[m] are caller buffers (upper bounds); *sub_n and
 * *sub_m are written on the host. */
int kspn_subgraph(kspn_ctx* ctx, uint32_t n, const uint64_t* off, const uint32_t* adj, const uint8_t* mask,
                  uint32_t* super_ids, uint64_t* sub_off, uint32_t* sub_adj, uint32_t* sub_n,
                  uint64_t* sub_m, void* stream);

/* tarjan (kaspan/include/kaspan/scc/tarjan.hpp:20-136) on the GPU, one
 * wavefront, labels = minimum member id. For small graphs (the residual). */
int kspn_tarjan(kspn_ctx* ctx, uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* adj,
                uint32_t* labels, void* stream);

/* kspn_tarjan without any allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). The work area
 * (6 words per vertex) is a fixed device array of the library,
 * 24 MiB + 256 B on every GPU it runs on, held until the process ends, so
 * n may not exceed KSPN_TARJAN_NOALLOC_MAX_N (KSPN_E_INVALID otherwise).
 * Runs on the calling thread's current device; stream-ordered, it returns
 * after the launch without synchronising. Calls that overlap on one GPU
 * share the work area and race: order them on one stream. */
#define KSPN_TARJAN_NOALLOC_MAX_N (1u << 20)
int kspn_tarjan_noalloc(uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* adj, uint32_t* labels,
                        void* stream);

/* ---- learned pivot scorer (no reference counterpart; see KSPN_PIVOT_LEARNED) ----
 * MLP 6 -> 32 -> 32 -> 1 (ReLU). Parameter vector of KSPN_MLP_PARAMS floats:
 * W1[32][6], b1[32], W2[32][32], b2[32], w3[32], b3. Features per vertex:
 * log1p(in), log1p(out), log1p(in)*log1p(out)/8, log1p(sum of out-degrees of
 * out-neighbours), log1p(sum of in-degrees of in-neighbours), log1p(min(in,out)). */
#define KSPN_MLP_PARAMS 1313

/* feat[count][6] for vertices[0..count) (vertices NULL: 0..count-1); device pointers */
int kspn_pivot_features(kspn_ctx* ctx, const kspn_graph* g, const uint32_t* vertices, uint32_t count, float* feat,
                        void* stream);
/* scores out[rows] = MLP(x[rows][6]); weights rounded to BF16, FP32 accumulation (MFMA) */
int kspn_mlp_forward(kspn_ctx* ctx, const float* params, const float* x, uint32_t rows, float* out, void* stream);
/* gradient of 0.5*mean((MLP(x) - target)^2): grad[KSPN_MLP_PARAMS], *loss = 0.5*sum of squared errors
 * (device). Matrix products on MFMA with BF16 operands (X, W1, relu(A1), W2 and the pre-activation
 * gradients) and FP32 accumulation; biases, w3 and the loss in FP32. */
int kspn_mlp_grad(kspn_ctx* ctx, const float* params, const float* x, const float* target, uint32_t rows, float* grad,
                  float* loss, void* stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8); state[2*KSPN_MLP_PARAMS] starts at 0; step >= 1 */
int kspn_mlp_adam(kspn_ctx* ctx, float* params, const float* grad, float* state, uint32_t step, float lr, void* stream);
/* pivot outcome log: sizes[v] = size of v's SCC, from canonical labels (device) */
int kspn_scc_sizes(kspn_ctx* ctx, uint32_t n, const uint32_t* labels, uint32_t* sizes, void* stream);

/* ---- seeded graph construction (inputs for tests and benchmarks) --------- */
#define KSPN_GEN_GNM 0     /* p0 = n, p1 = draws */
#define KSPN_GEN_RMAT 1    /* p0 = scale, p1 = draws, a,b,c, permute */
#define KSPN_GEN_LATTICE 2 /* p0 = side, a = p_bidirectional */
#define KSPN_GEN_PLANTED 3 /* p0 = n, p1 = max component size, a = exponent, b = intra out-degree, c = inter out-degree */

typedef struct kspn_gen_params {
  uint32_t kind;
  uint64_t seed;
  uint64_t p0, p1;
  double a, b, c;
  uint32_t permute;
} kspn_gen_params;

/* Generates the edge list on the device, removes self-loops and duplicates
 * and builds both CSRs. The library allocates the four arrays with
 * hipMalloc; free them with kspn_free. */
int kspn_generate(kspn_ctx* ctx, const kspn_gen_params* p, uint32_t* n, uint64_t* m, uint64_t** fw_off,
                  uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj, void* stream);
/* CSR + reverse CSR from a device edge list (src[i] -> dst[i]); drops self-loops and duplicates */
int kspn_build_csr(kspn_ctx* ctx, uint32_t n, uint64_t count, const uint32_t* src, const uint32_t* dst,
                   uint64_t* m, uint64_t** fw_off, uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj,
                   void* stream);
int kspn_free(void* device_ptr);

/* seeded batch of independent graphs: graph i is G(n,m) or R-MAT with
 * probability p_rmat, n log-uniform in [n_min, n_max], draws = n * degree
 * with the degree log-uniform in [deg_min, deg_max] */
typedef struct kspn_batch_gen {
  uint64_t seed;
  uint32_t count;
  uint32_t n_min, n_max;
  double deg_min, deg_max;
  double p_rmat;
  double a, b, c; /* R-MAT quadrant probabilities */
} kspn_batch_gen;

/* per-graph plan (host arrays of count entries): kind (KSPN_GEN_GNM or
 * KSPN_GEN_RMAT), vertices, edge draws before deduplication, seed */
int kspn_batch_plan(const kspn_batch_gen* p, uint32_t* kind, uint32_t* n, uint64_t* draws, uint64_t* seed);
/* generates and packs every graph of the plan (self-loops and duplicates
 * removed, reverse CSR built per graph). The library allocates the six
 * arrays of *out with hipMalloc; free each with kspn_free. fw_adj and
 * bw_adj hold room for the draws of the plan (at least out->m). */
int kspn_generate_batch(kspn_ctx* ctx, const kspn_batch_gen* p, kspn_batch* out, void* stream);
/* canonical labels of a KSPN_GEN_PLANTED graph, known by construction (device labels[n]) */
int kspn_planted_labels(kspn_ctx* ctx, const kspn_gen_params* p, uint32_t* labels, void* stream);

/* ---- graphs stored as a manifest (scc/adapter/manifest.hpp) -------------
 * A text file of "key value" lines ('%' starts a comment line) naming four
 * binary files of dense unsigned integers: fw/bw row offsets (n+1 elements
 * of graph.head.bytes each) and fw/bw adjacency (m elements of
 * graph.csr.bytes each), in graph.endian byte order. Written by the
 * reference's convert_graph (scc/adapter/external_edgelist.hpp:348-367). */
#define KSPN_MANIFEST_TEXT 256
#define KSPN_MANIFEST_PATH 4096
typedef struct kspn_manifest {
  uint32_t schema_version;           /* schema.version */
  uint32_t big_endian;               /* graph.endian: 0 little, 1 big */
  uint64_t node_count;               /* graph.node_count */
  uint64_t edge_count;               /* graph.edge_count */
  uint32_t contains_self_loops;      /* graph.contains_self_loops */
  uint32_t contains_duplicate_edges; /* graph.contains_duplicate_edges */
  uint32_t head_bytes;               /* graph.head.bytes */
  uint32_t csr_bytes;                /* graph.csr.bytes */
  char code[KSPN_MANIFEST_TEXT];     /* graph.code */
  char name[KSPN_MANIFEST_TEXT];     /* graph.name */
  /* {fw,bw}.{head,csr}.path resolved against the manifest's directory */
  char fw_head_path[KSPN_MANIFEST_PATH];
  char fw_csr_path[KSPN_MANIFEST_PATH];
  char bw_head_path[KSPN_MANIFEST_PATH];
  char bw_csr_path[KSPN_MANIFEST_PATH];
} kspn_manifest;

/* manifest::load (scc/adapter/manifest.hpp:136-197): every key present,
 * numbers as whole unsigned decimals, booleans "true"/"false", endian
 * "little"/"big", each path a regular file. KSPN_E_INVALID for a malformed
 * manifest, KSPN_E_IO for a missing file; kspn_last_error names the key. */
int kspn_manifest_parse(const char* path, kspn_manifest* out);
/* load_graph_from_manifest (scc/adapter/manifest.hpp:200-250) into caller
 * device buffers: fw_off/bw_off hold node_count+1 entries, fw_adj/bw_adj
 * edge_count. Each file must hold exactly its elements' bytes (KSPN_E_IO
 * otherwise). Every offset must be <= edge_count and every id <
 * node_count; offsets must start at 0, end at edge_count and not decrease
 * (KSPN_E_INVALID otherwise). Synchronises the stream before returning. */
int kspn_manifest_load(kspn_ctx* ctx, const kspn_manifest* mf, uint64_t* fw_off, uint32_t* fw_adj,
                       uint64_t* bw_off, uint32_t* bw_adj, void* stream);
/* kspn_manifest_load into arrays the library allocates with hipMalloc
 * (free each with kspn_free); *n and *m receive the graph's sizes */
int kspn_manifest_graph(kspn_ctx* ctx, const kspn_manifest* mf, uint32_t* n, uint64_t* m, uint64_t** fw_off,
                        uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj, void* stream);
/* dense_unsigned_accessor::get over a device array
 * (memory/dense_unsigned_accessor.hpp:15-110, dense_unsigned_ops.hpp:14-44):
 * out[i] (out_bytes 4 or 8) = element i of `packed` (elem_bytes 1..8, byte
 * order big_endian ? big : little). KSPN_E_INVALID if any value exceeds
 * `limit`. Synchronises the stream before returning. */
int kspn_unpack_dense(kspn_ctx* ctx, const void* packed, uint64_t count, uint32_t elem_bytes, uint32_t big_endian,
                      uint64_t limit, void* out, uint32_t out_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
