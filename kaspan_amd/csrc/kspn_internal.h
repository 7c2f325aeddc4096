// kspn_internal.h -- device helpers and the context layout shared by the
// libkspn translation units. gfx950 only: wave64, 256-thread workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/kspn/scc.h"

namespace kspn {

constexpr int WG = 256;                 // threads per workgroup (4 waves)
constexpr int WAVE = 64;                // CDNA wavefront
constexpr uint32_t UNDEC = KSPN_UNDECIDED;

#define KSPN_TRY(expr)                 \
  do {                                 \
    int _rc = (expr);                  \
    if (_rc != KSPN_OK) return _rc;    \
  } while (0)

// records the failing call for kspn_last_error (thread-local)
void note_error(const char* what, const char* detail, const char* file, int line);
void clear_error();

#define KSPN_HIP(expr)                                                         \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::kspn::note_error(#expr, hipGetErrorName(_e), __FILE__, __LINE__);      \
      return KSPN_E_HIP;                                                       \
    }                                                                          \
  } while (0)

// ---------------------------------------------------------------- bitmaps
// One bit per vertex in uint32 words; arrays are padded to whole 64-bit
// groups so a wavefront can write its 64 ballot bits as two words.
__host__ __device__ inline uint64_t bm_words(uint64_t n) { return ((n + 63) / 64) * 2; }

__device__ inline bool bm_get(const uint32_t* bm, uint32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

// set a bit; true for the one thread that turned it on
__device__ inline bool bm_claim(uint32_t* bm, uint32_t v) {
  uint32_t mask = 1u << (v & 31);
  if (bm[v >> 5] & mask) return false;
  return (atomicOr(&bm[v >> 5], mask) & mask) == 0;
}

// clear a bit; true for the one thread that turned it off
__device__ inline bool bm_unclaim(uint32_t* bm, uint32_t v) {
  uint32_t mask = 1u << (v & 31);
  if (!(bm[v >> 5] & mask)) return false;
  return (atomicAnd(&bm[v >> 5], ~mask) & mask) != 0;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Wavefront-aggregated append: every active lane calls it; lanes with
// pred append val at consecutive slots after one atomicAdd per wave.
__device__ inline void wave_append(bool pred, uint32_t val, uint32_t* queue, uint32_t* count) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  uint32_t lane = lane_id();
  uint32_t leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (pred) {
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    queue[base + __popcll(below)] = val;
  }
}

// wave-aggregated counter add
__device__ inline void wave_count(bool pred, uint32_t* count) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  uint32_t lane = lane_id();
  uint32_t leader = __ffsll((unsigned long long)mask) - 1;
  if (lane == leader) atomicAdd(count, (uint32_t)__popcll(mask));
}

// ---------------------------------------------------------------- reductions
__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  return v;  // valid in lane 0
}
__device__ inline uint32_t wave_min(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off));
  return v;
}
__device__ inline uint64_t wave_sum64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  return v;
}
// inclusive prefix sum over the wavefront
__device__ inline uint64_t wave_incl_scan64(uint64_t v) {
  const uint32_t lane = lane_id();
  for (uint32_t o = 1; o < WAVE; o <<= 1) {
    const uint64_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// block-wide 64-bit sum; every thread calls; result in thread 0
__device__ inline uint64_t block_sum64(uint64_t v, uint64_t* s_red) {
  v = wave_sum64(v);
  if (lane_id() == 0) s_red[threadIdx.x / WAVE] = v;
  __syncthreads();
  uint64_t r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < WG / WAVE; ++w) r += s_red[w];
  __syncthreads();
  return r;
}

// block-wide sum / min; every thread of the block calls; result in thread 0
__device__ inline uint32_t block_sum(uint32_t v, uint32_t* s_red) {
  v = wave_sum(v);
  if (lane_id() == 0) s_red[threadIdx.x / WAVE] = v;
  __syncthreads();
  uint32_t r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < WG / WAVE; ++w) r += s_red[w];
  __syncthreads();
  return r;
}
__device__ inline uint32_t block_min(uint32_t v, uint32_t* s_red) {
  v = wave_min(v);
  if (lane_id() == 0) s_red[threadIdx.x / WAVE] = v;
  __syncthreads();
  uint32_t r = 0xFFFFFFFFu;
  if (threadIdx.x == 0)
    for (int w = 0; w < WG / WAVE; ++w) r = min(r, s_red[w]);
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- LDS-staged queue
// Appends are collected in LDS (one LDS atomic per wavefront) and written
// out with one global atomicAdd per flush, so a launch issues a few
// thousand global atomics instead of one per wavefront.
struct BlockQueue {
  uint32_t* buf;  // __shared__, cap entries
  uint32_t* n;    // __shared__
  uint32_t* base; // __shared__
  uint32_t cap;
};

__device__ inline void bq_init(BlockQueue& q) {
  if (threadIdx.x == 0) *q.n = 0;
  __syncthreads();
}

__device__ inline void bq_push(BlockQueue& q, bool pred, uint32_t val) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  uint32_t lane = lane_id();
  uint32_t leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t b = 0;
  if (lane == leader) b = atomicAdd(q.n, (uint32_t)__popcll(mask));
  b = __shfl(b, leader);
  if (pred) {
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    q.buf[b + __popcll(below)] = val;
  }
}

// every thread of the block calls; writes the staged entries to queue
__device__ inline void bq_flush(BlockQueue& q, uint32_t* queue, uint32_t* count) {
  __syncthreads();
  uint32_t cnt = *q.n;
  if (cnt) {
    if (threadIdx.x == 0) *q.base = atomicAdd(count, cnt);
    __syncthreads();
    uint32_t b = *q.base;
    for (uint32_t i = threadIdx.x; i < cnt; i += WG) queue[b + i] = q.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *q.n = 0;
  }
  __syncthreads();
}

// bq_flush for a queue of `cap` entries: the count always grows, the
// entries are written only while they fit
__device__ inline void bq_flush_capped(BlockQueue& q, uint32_t* queue, uint32_t* count, uint32_t cap) {
  __syncthreads();
  uint32_t cnt = *q.n;
  if (cnt) {
    if (threadIdx.x == 0) *q.base = atomicAdd(count, cnt);
    __syncthreads();
    uint32_t b = *q.base;
    if (b + cnt <= cap)
      for (uint32_t i = threadIdx.x; i < cnt; i += WG) queue[b + i] = q.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *q.n = 0;
  }
  __syncthreads();
}

// flush when fewer than `room` free slots remain (block-uniform decision)
__device__ inline void bq_maybe_flush(BlockQueue& q, uint32_t* queue, uint32_t* count, uint32_t room) {
  __syncthreads();
  if (*q.n + room > q.cap) bq_flush(q, queue, count);
}

__device__ inline void bq_maybe_flush_capped(BlockQueue& q, uint32_t* queue, uint32_t* count, uint32_t cap,
                                             uint32_t room) {
  __syncthreads();
  if (*q.n + room > q.cap) bq_flush_capped(q, queue, count, cap);
}

__device__ inline uint32_t rotr32(uint32_t x, uint32_t r) { r &= 31; return r ? (x >> r) | (x << (32 - r)) : x; }
__device__ inline uint32_t rotl32(uint32_t x, uint32_t r) { r &= 31; return r ? (x << r) | (x >> (32 - r)) : x; }

// ---------------------------------------------------------------- counters
// Device-side counters, read back through a pinned mirror.
constexpr uint32_t MAP_BANK = 64;  // words per bank of the host-mapped publication page (sequence word at 63)

enum Counter : int {
  C_Q0 = 0,   // size of queue written by a launch (ping)
  C_Q1,       // pong
  C_DECIDED,  // decisions made by a launch
  C_AUX,
  C_ROOM,    // the tail's sub-CSR did not fit its buffer
  C_RES_N,
  C_FLAG,
  C_MIN,   // minimum vertex of a decided set
  C_FSEG,  // size of the current frontier segment
  C_DTOT,  // running length of the decided list
  C_ERR,   // set by a kernel that met a broken internal invariant
  C_CURSOR,  // write cursor of an unordered bitmap compaction
  C_SURV0,   // sizes of the candidate lists of pull list levels (four, in rotation)
  C_SURV1,
  C_SURV2,
  C_SURV3,
  C_ISO,     // live vertices without a live neighbour, counted by the tail preparation
  C_LREC,    // claims of the last pull list level of a batch
  C_LIVE,    // vertices left live after the last FW-BW, listed by k_fwbw_finish for the tail solver
  C_SPEC,    // k_spec_gate: a speculatively launched pull sweep is to run
  C_BADID,   // a tail kernel met a listed vertex id >= n and skipped it
  C_AERR,    // the asynchronous reach gave up (1: no progress, 2: queue overflow)
  C_SERR,    // the look-back degree scan met a count or ticket it could not use
  C_DLIST,   // length of the list of vertices a multi-pivot round decided (k_part_split)
  C_COUNT_ = 28  // at most 32: the publication map holds the persistent kernel's results from word 32 on
};

struct KernTimer {
  const char* name;
  uint32_t launches = 0;
  double ms = 0, bytes = 0;
};

struct EventPair {
  hipEvent_t a, b;
  int kern;
  int phase;  // statistics scope open at the launch (-1: none)
  double bytes;
};

// ---------------------------------------------------------------- context
// The device-driven solve's buffers and its captured HIP graph
// (kspn_graphsolve.hip). The graph is kept while the inputs, outputs,
// options and buffers it was captured with stay the same.
struct GraphMode {
  void* ctl = nullptr;                     // control block (gs::GCtl)
  unsigned long long* pc = nullptr;        // per vertex: partition << 32 | colour
  unsigned long long* q0 = nullptr;        // frontier row slices (two queues)
  unsigned long long* q1 = nullptr;
  uint64_t qcap = 0;                       // entries per queue
  unsigned long long* pkey = nullptr;      // per graph: pivot key
  uint32_t* gmin = nullptr;                // per graph: smallest member of the pivot's SCC
  uint32_t cap_n = 0, cap_count = 0;
  uint64_t cap_m = 0;
  uint32_t grid = 0;                       // persistent workgroups, all resident at once (0: not checked yet)
  hipStream_t cap_stream = nullptr;        // capture stream (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t key[16] = {};                   // what the captured graph was built for
  void* h_status = nullptr;                // pinned copy of the control block
  hipEvent_t ev[2] = {nullptr, nullptr};
  uint32_t replays = 0;                    // launches of the current captured graph
};

struct Ctx {
  int device = 0;
  uint32_t cap_n = 0;
  uint64_t cap_m = 0;
  // bitmaps (bm_words(n) words each)
  uint32_t* live = nullptr;  // 1 = undecided
  uint32_t* fwv = nullptr;   // forward-reached from the pivot
  uint32_t* bwv = nullptr;   // backward-reached (inside fwv)
  uint32_t* inq = nullptr;   // queue membership
  uint32_t* inq2 = nullptr;  // chunked colour levels: the other queue bitmap (levels alternate between the two)
  uint32_t* cchain = nullptr;  // chunked colour levels: control words of a device chain of push levels
  uint32_t* fnew = nullptr;  // vertices reached by the last pull level
  uint32_t* sbm = nullptr;   // candidates the last pull sweep left unreached
  // vertex arrays
  uint32_t* qa = nullptr;
  uint32_t* qb = nullptr;
  uint32_t* qc = nullptr;
  uint32_t* ulist = nullptr;  // live vertices, ascending
  uint32_t* col = nullptr;    // colours; reused as min-id table
  uint64_t* deg = nullptr;    // frontier degrees (n+1)
  uint64_t* scan = nullptr;   // exclusive scan of deg (n+1)
  uint64_t* part = nullptr;   // per-workgroup partials
  uint32_t* cnt = nullptr;    // device counters
  unsigned long long* cnt64 = nullptr;  // device 64-bit counters (edges scanned)
  uint32_t* h_cnt = nullptr;  // host copy of the published counters
  volatile uint32_t* h_map = nullptr;  // pinned, device-mapped: counters published by k_publish (two banks of MAP_BANK)
  uint32_t* d_map = nullptr;  // device address of h_map
  uint64_t* h_u64 = nullptr;  // pinned scratch for 64-bit reads
  void* cub_tmp = nullptr;
  size_t cub_bytes = 0;
  // residual
  uint64_t* sub_off = nullptr;
  uint32_t* sub_adj = nullptr;
  uint64_t cap_sub_m = 0;
  uint32_t* tj = nullptr;  // Tarjan work arrays (6 * n words)
  uint32_t* mlp_buf = nullptr;  // learned pivot: candidates, histogram, features, scores
  // partitioned FW-BW
  uint32_t* pid = nullptr;     // partition of every vertex
  uint64_t* pbest = nullptr;   // per sparse partition id: best pivot key (3n slots)
  uint32_t* pmin = nullptr;    // per partition: smallest member of the pivot's SCC
  uint32_t* pflags = nullptr;  // per sparse partition id: used (pbest != 0)
  uint32_t* pnew = nullptr;    // per sparse partition id: compact index (scan of pflags)
  uint32_t* pcnt = nullptr;    // per partition: outcome counters for the pivot log
  float* vscore = nullptr;     // per vertex: MLP score
  uint32_t* vprio = nullptr;   // per vertex: pivot priority of the multi-pivot rounds (fixed within a solve)
  unsigned long long* cpw = nullptr;    // colour rounds: (partition << 32 | colour) per vertex
  uint8_t* deg8 = nullptr;              // features: min(degree, 255) per vertex, out then in
  uint32_t deg8_cap = 0;
  unsigned long long* ds_st = nullptr;  // look-back scan: status word per tile
  uint32_t* ds_ticket = nullptr;        // its ticket counter
  uint32_t ds_epoch = 0;                // epoch of its last launch
  uint32_t* tstart = nullptr;           // per expansion tile of the last look-back scan: its first frontier entry
  uint64_t tstart_cap = 0;              // entries tstart holds
  uint32_t* prune_fw = nullptr;         // KSPN_FLAG_PRUNE: one bit per fw_adj entry (edges joining two partitions)
  uint32_t* prune_bw = nullptr;         // ... and per bw_adj entry
  uint64_t prune_words = 0;             // words each holds
  float* vfeat = nullptr;      // features of live vertices (6 per vertex) + scores
  // profiling
  bool profile = false;
  std::vector<KernTimer> kt;
  std::vector<EventPair> pending;
  std::vector<hipEvent_t> free_events;
  int cur_phase = -1;                     // statistics scope open now (kspn_stats.phase index), -1: none
  double phase_bytes[KSPN_PHASE_MAX] = {};  // profile mode: algorithmic bytes of the launches in each scope
  double phase_kms[KSPN_PHASE_MAX] = {};    // profile mode: summed launch durations in each scope
  hipStream_t stream = nullptr;
  int num_cu = 256;
  uint32_t pub_seq = 0;  // sequence number of the last counter publication (h_map[63])
  // persistent multi-level reachability (cooperative launch)
  GraphMode gm;                    // KSPN_FLAG_GRAPH
  unsigned long long* aq = nullptr;  // asynchronous reach: slot queue (all empty between reaches)
  uint64_t aq_cap = 0;
  void* actl = nullptr;              // its control block
  struct CoopCtl* coop = nullptr;  // device control block
  uint32_t coop_grid = 0;          // co-resident workgroups (0: cooperative launch unavailable)
};

// Control block of the persistent reachability kernel: per-level values
// live in rings of three indexed by level % 3 (frontier being expanded,
// next frontier being filled, the one after is cleared for reuse).
struct CoopCtl {
  uint32_t cnt[4];              // frontier sizes
  uint32_t work[4];             // 64-vertex chunks handed out
  unsigned long long nsum[4];   // out-degree sums of the frontiers
  uint32_t maxd[4];             // largest out-degree in the frontiers
  unsigned long long expanded;  // edges expanded by all levels of the launch
  unsigned long long verts;     // frontier vertices expanded by all levels of the launch
  uint32_t bar_count, bar_pad;  // grid barrier: arrivals so far in the launch
  uint32_t levels, reason, err, pad;
  unsigned long long lexp[4];   // edges expanded by the wavefronts' local rounds of a level
  unsigned long long lvert[4];  // vertices expanded by them
  double mrest[4];              // chained launches: edges not yet expanded at the level of each slot
};

// host helpers implemented in kspn_scc.hip
int ctx_reserve(Ctx* c, uint32_t n, uint64_t m);
int read_counters(Ctx* c, hipStream_t s, int count);
int exclusive_scan_u64(Ctx* c, const uint64_t* in, uint64_t* out, uint64_t items, hipStream_t s);
int exclusive_scan_u32(Ctx* c, const uint32_t* in, uint32_t* out, uint64_t items, hipStream_t s);
int build_live_list(Ctx* c, uint32_t n, const uint32_t* bitmap, uint32_t* list, uint32_t* count_out, hipStream_t s);

inline uint32_t grid_for(uint64_t items, uint32_t per_block = WG, uint32_t cap = 1u << 20) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

// the device-driven solve (KSPN_FLAG_GRAPH): captured once per shape, replayed
int graph_solve(Ctx* c, const kspn_graph* g, uint32_t* labels, const kspn_opts& o, const uint32_t* vtx_off,
                uint32_t count, kspn_stats* st, hipStream_t s);
void graph_release(Ctx* c);

int select_pivot_learned_impl(Ctx* c, const kspn_graph* g, const float* params, hipStream_t s, uint32_t* pivot);
int mlp_features(Ctx* c, const kspn_graph* g, const uint32_t* list, uint32_t count, float* feat, hipStream_t s);
int mlp_forward(Ctx* c, const float* params, const float* x, uint32_t rows, float* out, hipStream_t s);
int mlp_score_top(Ctx* c, const kspn_graph* g, const float* params, uint32_t live_n, uint32_t div, uint32_t* list,
                  uint32_t* count_dev, float* feat, float* sc, float* vscore, hipStream_t s, uint32_t* scored);

}  // namespace kspn

// the opaque handle of the C ABI
struct kspn_ctx {
  kspn::Ctx c;
};
