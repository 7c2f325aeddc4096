// kspn_scc.hip -- the SCC pipeline on one MI355X (gfx950, wave64).
//
// Stage map (paths are in the rawbby/KaSpan tree):
//   k_trim_first + k_pivot_reduce  trim_1_first, scc/trim_1_local.hpp:12-46
//   fwbw()                         forward_backward_search, scc/pivot_search.hpp:15-85
//                                  (push levels: k_expand; pull levels: k_pull_*)
//   k_trim_sweep (NORMAL)          trim_1_normal, scc/trim_1_local.hpp:48-73
//   k_trim_sweep (TRIM1) loop      trim_1_exhaustive[_first], scc/trim_1_exhaustive.hpp:14-131
//   k_trim2_*                      trim-2 (no reference counterpart)
//   color_round()                  cache_rot_label_search_pull, scc/cache_multi_pivot_search_pull.hpp:40-266
//   residual()                     allgather_fw_sub_graph + tarjan, scc/variant/scc_trim_ex_residual.hpp:110-124
//   canonicalize()                 normalise_scc_id, test/normalise_scc.hpp:10-25
//
// Device state lives in bitmaps (1 bit per vertex, L2-resident at 2 MiB per
// 16.7M vertices) and uint32 vertex queues; frontier expansion is load
// balanced over edges (exclusive scan of frontier degrees, 2048-edge tiles,
// LDS-staged frontier window) so hub vertices split across workgroups.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <mutex>
#include <new>

#include "kspn_internal.h"

namespace kspn {

// =====================================================================
// generic kernels
// =====================================================================

__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < count; i += (uint64_t)gridDim.x * WG) p[i] = v;
}

// *p = v (small device writes without pageable-memory copies)
__global__ void k_set_u32(uint32_t* p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}
__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src) {
  if (threadIdx.x == 0) *dst = *src;
}

// counters -> host-mapped pinned memory: C_COUNT_ u32 counters at 0, u64
// slots (*extra, cnt64[0..2]) at 48; the sequence number goes last (map[63]) after a
// system-scope fence, so a host that sees it sees the counters
__device__ inline void publish_seq(uint32_t* map, uint32_t seq) {
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(&map[63], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Optionally first reduces per-workgroup partials (claimed count -> *pcount,
// edges scanned -> cnt64[0], next-level degree sum -> cnt64[1]); after the
// copy it can zero *reset and cnt64[0..2] for the next level, so a level
// needs no separate reset or reduction launch. With `gate` (SpecGate),
// it does nothing unless *gate is set; with sg.gate it evaluates SpecGate
// first. One workgroup of WG threads.
// The host's direction rule for the level after a push level, evaluated by
// that level's publication before it resets the counters, so that a pull
// sweep can be launched before the host has read them: pull when (degree
// sum of the claimed vertices, cnt64[1]) * alpha > m_rest - *expanded (the
// edges the push level expanded) and more than 64 vertices were claimed,
// with the host's operands and double arithmetic. *gate = 1 for pull;
// otherwise the four list counters `lcnt` are zeroed (a sweep would have
// zeroed them), so the list levels launched behind the gated sweep find no
// candidates.
struct SpecGate {
  uint32_t* gate;
  const uint32_t* claimed;
  const uint64_t* expanded;
  double m_rest;
  double alpha;
  uint32_t* lcnt;
};

__global__ __launch_bounds__(WG) void k_publish(uint32_t* cnt, unsigned long long* cnt64, const uint64_t* extra,
                                                uint32_t* map, uint32_t seq, const uint64_t* __restrict__ part,
                                                uint32_t nparts, uint32_t* pcount, uint32_t* reset,
                                                uint32_t* reset2, const uint32_t* gate, SpecGate sg) {
  __shared__ uint64_t s_red[WG / WAVE];
  if (gate && *gate == 0) return;  // a speculative launch that lost: nothing is published
  if (sg.gate && threadIdx.x == 0) {
    const double rest = sg.m_rest - (double)*sg.expanded;
    const bool pull = (double)cnt64[1] * sg.alpha > rest && *sg.claimed > 64;
    *sg.gate = pull ? 1u : 0u;
    if (!pull) sg.lcnt[0] = sg.lcnt[1] = sg.lcnt[2] = sg.lcnt[3] = 0;
  }
  if (part) {
    uint64_t a = 0, b = 0, d = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += WG) {
      a += part[3 * i];
      b += part[3 * i + 1];
      d += part[3 * i + 2];
    }
    a = block_sum64(a, s_red);
    b = block_sum64(b, s_red);
    d = block_sum64(d, s_red);
    if (threadIdx.x == 0) {
      *pcount = (uint32_t)a;
      cnt64[0] = b;
      cnt64[1] = d;
    }
    __syncthreads();
  }
  uint32_t t = threadIdx.x;
  if (t < C_COUNT_) map[t] = cnt[t];
  if (t == 32) {
    uint64_t x = extra ? *extra : 0;
    map[48] = (uint32_t)x;
    map[49] = (uint32_t)(x >> 32);
  }
  if (t == 33) {
    uint64_t y = cnt64[0];
    map[50] = (uint32_t)y;
    map[51] = (uint32_t)(y >> 32);
  }
  if (t == 34) {
    uint64_t y = cnt64[1];
    map[52] = (uint32_t)y;
    map[53] = (uint32_t)(y >> 32);
  }
  if (t == 35) {
    uint64_t y = cnt64[2];
    map[54] = (uint32_t)y;
    map[55] = (uint32_t)(y >> 32);
  }
  publish_seq(map, seq);
  if (reset) {
    __syncthreads();  // every copy above is done
    if (threadIdx.x == 0) {
      *reset = 0;
      if (reset2) *reset2 = 0;
      cnt64[0] = 0;
      cnt64[1] = 0;
      cnt64[2] = 0;
    }
  }
}

__global__ void k_zero3(uint32_t* a, uint32_t* b, uint32_t* c) {
  if (threadIdx.x == 0) *a = *b = *c = 0;
}

__global__ void k_zero2(uint32_t* a, uint32_t* b) {
  if (threadIdx.x == 0) {
    *a = 0;
    *b = 0;
  }
}

// FW-BW seed: bm = {v} (the word holding v gets its bit, all others 0),
// queue = [v], *count = 1; one launch for the four steps
// bitmap = {v}, queue = [v]; with vdev, v is read there and 0xFFFFFFFF
// (no pivot) seeds nothing
__global__ void k_seed(uint32_t* bm, uint64_t words, uint32_t v, const uint32_t* vdev, uint32_t* queue,
                       uint32_t* count) {
  if (vdev) v = *vdev;
  const bool none = v == 0xFFFFFFFFu;
  for (uint64_t w = blockIdx.x * (uint64_t)WG + threadIdx.x; w < words; w += (uint64_t)gridDim.x * WG)
    bm[w] = (!none && w == (v >> 5)) ? (1u << (v & 31)) : 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    queue[0] = v;
    *count = none ? 0u : 1u;
  }
}

constexpr uint32_t B2Q_STEPS = 4;  // 4 beat 8 and 2 on R-MAT 24's survivor lists
constexpr uint64_t B2Q_CHUNK = (uint64_t)WG * 4 * B2Q_STEPS;

// inclusive prefix sum over the wavefront
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t o = 1; o < WAVE; o <<= 1) {
    const uint32_t x = __shfl_up(v, o);
    if (lane >= o) v += x;
  }
  return v;
}

// Appends the set bits of a workgroup's chunk of STEPS * 4 * WG words
// (w[t][k] is word base + t*4*WG + 4*threadIdx.x + k) to list, reserving the chunk's
// range with one atomicAdd on *cursor; each thread writes its bits to a run
// of its own. Every thread of the workgroup calls.
template <uint32_t STEPS>
__device__ inline void chunk_to_list(const uint32_t (&w)[STEPS][4], uint64_t base, uint32_t* __restrict__ list,
                                     uint32_t* cursor, uint32_t* s_wave, uint32_t* s_base) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / WAVE;
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t t = 0; t < STEPS; ++t) cnt += __popc(w[t][0]) + __popc(w[t][1]) + __popc(w[t][2]) + __popc(w[t][3]);
  const uint32_t inc = wave_incl_scan(cnt);
  if (lane == WAVE - 1) s_wave[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < WG / WAVE; ++k) {
      const uint32_t t = s_wave[k];
      s_wave[k] = tot;
      tot += t;
    }
    *s_base = tot ? atomicAdd(cursor, tot) : 0u;
  }
  __syncthreads();
  uint32_t pos = *s_base + s_wave[wave] + inc - cnt;
#pragma unroll
  for (uint32_t t = 0; t < STEPS; ++t) {
    const uint64_t i0 = base + (uint64_t)t * WG * 4 + (uint64_t)threadIdx.x * 4;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      uint32_t x = w[t][k];
      while (x) {
        const uint32_t b = __ffs(x) - 1;
        list[pos++] = (uint32_t)((i0 + k) * 32 + b);
        x &= x - 1;
      }
    }
  }
  __syncthreads();
}

// Unordered bitmap -> list. Each workgroup takes chunks of B2Q_CHUNK words
// (B2Q_STEPS 16-byte loads per thread, all in flight together) and appends
// their set bits with chunk_to_list: few, large chunks keep the atomics on
// *cursor (zero before the launch) few. For lists whose order does not
// matter. With `guard`, nothing is written when *guard exceeds guard_max
// (read after the words, so the two loads overlap). Launch b2q_grid(words)
// workgroups.
__global__ __launch_bounds__(WG) void k_bitmap_to_queue(const uint32_t* __restrict__ bm, uint64_t words,
                                                        uint32_t* __restrict__ list, uint32_t* cursor,
                                                        const unsigned long long* guard = nullptr,
                                                        unsigned long long guard_max = 0) {
  __shared__ uint32_t s_wave[WG / WAVE];
  __shared__ uint32_t s_base;
  bool first = true;
  for (uint64_t base = (uint64_t)blockIdx.x * B2Q_CHUNK; base < words; base += (uint64_t)gridDim.x * B2Q_CHUNK) {
    uint32_t w[B2Q_STEPS][4];
#pragma unroll
    for (uint32_t t = 0; t < B2Q_STEPS; ++t) {
      const uint64_t i0 = base + (uint64_t)t * WG * 4 + (uint64_t)threadIdx.x * 4;
      if (i0 + 4 <= words) {
        const uint4 q = *reinterpret_cast<const uint4*>(bm + i0);  // i0 is a multiple of 4 words
        w[t][0] = q.x;
        w[t][1] = q.y;
        w[t][2] = q.z;
        w[t][3] = q.w;
      } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) w[t][k] = i0 + k < words ? bm[i0 + k] : 0u;
      }
    }
    if (first) {
      first = false;
      if (guard && *guard > guard_max) return;  // the set bits are more than the caller wants listed
    }
    chunk_to_list(w, base, list, cursor, s_wave, &s_base);
  }
}

// set bit v of bm
__global__ void k_bm_set_one(uint32_t* bm, uint32_t v) {
  if (threadIdx.x == 0) atomicOr(&bm[v >> 5], 1u << (v & 31));
}

// clear queue members' bits
__global__ void k_bm_clear_list(uint32_t* bm, const uint32_t* list, const uint32_t* count_ptr) {
  uint32_t cnt = *count_ptr;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < cnt; i += gridDim.x * WG) {
    uint32_t v = list[i];
    atomicAnd(&bm[v >> 5], ~(1u << (v & 31)));
  }
}

// popcount per 32-bit word; entry `words` is 0 so the scan's last entry is the total
__global__ void k_bm_popc(const uint32_t* bm, uint64_t words, uint32_t* pc) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i <= words; i += (uint64_t)gridDim.x * WG)
    pc[i] = (i < words) ? __popc(bm[i]) : 0u;
}

// write set-bit positions ascending, word i starting at pos[i]
__global__ void k_bm_scatter(const uint32_t* bm, uint64_t words, const uint32_t* pos, uint32_t* list) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < words; i += (uint64_t)gridDim.x * WG) {
    uint32_t w = bm[i];
    uint32_t p = pos[i];
    while (w) {
      uint32_t b = __ffs(w) - 1;
      list[p++] = (uint32_t)(i * 32 + b);
      w &= w - 1;
    }
  }
}

// degree of each frontier vertex; deg[f] = 0 so the scan yields the total
__global__ void k_frontier_deg(const uint32_t* __restrict__ F, const uint32_t* fcount, const uint64_t* __restrict__ off,
                               uint64_t* __restrict__ deg) {
  uint32_t f = *fcount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i <= f; i += gridDim.x * WG) {
    if (i < f) {
      uint32_t v = F[i];
      deg[i] = off[v + 1] - off[v];
    } else {
      deg[i] = 0;
    }
  }
}

// =====================================================================
// edge-balanced frontier expansion
// =====================================================================

enum ExpandOp : int {
  OP_FW = 0,        // live target not yet reached forward: claim its fwv bit
  OP_BW = 1,        // forward-reached target not yet reached backward: claim its bwv bit
  OP_NEIGH = 2,     // collect live targets once (trim candidates)
  OP_COLOR = 3,     // push colour max into live targets
  OP_COLOR_BW = 4,  // backward closure: decide live in-neighbours of the same colour
  OP_PBW = 5,       // live target in the source's partition not yet reached backward: claim its bwv bit
};

struct OpArgs {
  uint32_t* live;
  uint32_t* fwv;  // forward-reached
  uint32_t* bwv;  // backward-reached (within fwv)
  uint32_t* inq;
  uint32_t* col;
  uint32_t* labels;
  uint32_t value;                // pivot or rotation
  const uint64_t* noff;          // OP_FW/OP_BW: offsets whose degree the next level expands
  unsigned long long* nsum;      // OP_FW/OP_BW: sum of those degrees over claimed targets; nsum[1]: their maximum
  const uint32_t* part;          // OP_FW/OP_PBW: partition of every vertex (nullptr: one partition)
  const uint32_t* pcand;         // OP_PBW: the vertices it may claim (live, or fwv for a search within FW sets)
  uint32_t dead_pid;             // part holds DEAD_PID for every vertex that is not live, so a partition match
                                 // implies live and the live bitmap need not be read
  unsigned long long* cp;        // OP_COLOR/OP_COLOR_BW: per vertex (partition << 32 | colour)
  uint32_t* prune;               // OP_FW/OP_PBW with part: one bit per entry of the adjacency being expanded, set
                                 // once the edge is seen to join two partitions (or reach a decided vertex);
                                 // partitions only split and decided vertices stay decided, so a set edge is
                                 // of no use to any later level of the solve (nullptr: not kept)
  uint32_t* inq_clear;           // k_expand_chunk<OP_COLOR>: the queue bitmap holding the frontier's bits, cleared
                                 // as the frontier is read (the level queues into inq, the other bitmap)
};

// partition id of a vertex that is not live during the multi-pivot rounds
// (above every id they use: 3 * partitions + 2 < 2^32 - 1 for n <= DEAD_PID_MAX_N)
constexpr uint32_t DEAD_PID = 0xFFFFFFFFu;
constexpr uint64_t DEAD_PID_MAX_N = 0x55555554ull;

// reachability claim of edge u -> w: true for the one thread that claims w
template <int OP>
__device__ inline bool reach_claim(uint32_t u, uint32_t w, const OpArgs& a) {
  if constexpr (OP == OP_FW) {
    if (a.part && a.dead_pid) return a.part[w] == a.part[u] && bm_claim(a.fwv, w);
    return bm_get(a.live, w) && (!a.part || a.part[w] == a.part[u]) && bm_claim(a.fwv, w);
  } else if constexpr (OP == OP_BW) {
    return bm_get(a.fwv, w) && bm_claim(a.bwv, w);
  } else {
    if (a.dead_pid && a.pcand == a.live) return a.part[w] == a.part[u] && bm_claim(a.bwv, w);
    return bm_get(a.pcand, w) && a.part[w] == a.part[u] && bm_claim(a.bwv, w);
  }
}

template <int OP>
__device__ inline void expand_apply(uint32_t u, uint32_t w, const OpArgs& a, BlockQueue& q, uint64_t& nsum,
                                    uint32_t& nmax) {
  bool push = false;
  if constexpr (OP == OP_FW || OP == OP_BW || OP == OP_PBW) {
    push = reach_claim<OP>(u, w, a);
  } else if constexpr (OP == OP_NEIGH) {
    push = bm_get(a.live, w) && bm_claim(a.inq, w);
  } else if constexpr (OP == OP_COLOR) {
    const unsigned long long xu = a.cp[u], xw = a.cp[w];
    if ((uint32_t)xw != 0u && (uint32_t)xw < (uint32_t)xu && (xw >> 32) == (xu >> 32)) {
      const unsigned long long old = atomicMax(&a.cp[w], xu);
      push = ((uint32_t)old < (uint32_t)xu) && bm_claim(a.inq, w);
    }
  } else if constexpr (OP == OP_COLOR_BW) {
    const unsigned long long xu = a.cp[u];
    if (a.cp[w] == xu && bm_unclaim(a.live, w)) {
      a.labels[w] = rotl32(~(uint32_t)xu, a.value);
      push = true;
    }
  }
  if constexpr (OP == OP_FW || OP == OP_BW || OP == OP_PBW) {
    if (push) {
      const uint64_t d = a.noff[w + 1] - a.noff[w];
      nsum += d;
      nmax = max(nmax, (uint32_t)min(d, (uint64_t)0xFFFFFFFFu));
    }
  }
  bq_push(q, push, w);
}

// A wavefront-private LDS queue of claimed vertices: entries are staged
// without block barriers, so a wavefront may loop on its own trip count.
struct WaveQueue {
  uint32_t* buf;  // __shared__, this wavefront's entries
  uint32_t n;     // wave-uniform
};

__device__ inline void bq_push(WaveQueue& q, bool pred, uint32_t val) {
  const uint64_t mask = __ballot(pred);
  if (pred) {
    const uint32_t lane = lane_id();
    const uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    q.buf[q.n + __popcll(below)] = val;
  }
  q.n += (uint32_t)__popcll(mask);
}

// the staged entries to queue[*count ...], one atomic per flush
__device__ inline void wq_flush(WaveQueue& q, uint32_t* queue, uint32_t* count) {
  if (q.n == 0) return;
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(count, q.n);
  base = __shfl(base, 0);
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k = lane; k < q.n; k += WAVE) queue[base + k] = q.buf[k];
  __builtin_amdgcn_wave_barrier();
  q.n = 0;
}

// The reach claims of a thread's K edges u[r] -> x[r] (ok[r]: the edge
// exists), with the loads of each step issued for all K together: candidate
// and visited words, partitions, the claiming atomics, the claimed
// vertices' degrees. Same claims as expand_apply on each edge in turn.
// Claimed vertices go to q (a BlockQueue or a WaveQueue). p[r] is the
// edge's position in the adjacency: with a.prune, an edge whose bit is set
// is skipped before its target is read, and an edge found to join two
// partitions gets its bit set.
template <int OP, int K, class Q>
__device__ inline void reach_batch(const uint32_t (&u)[K], const uint32_t (&x)[K], const bool (&ok_in)[K],
                                   const uint64_t (&p)[K], const OpArgs& a, Q& q, uint64_t& nsum, uint32_t& nmax) {
  const uint32_t* cand = OP == OP_BW ? a.fwv : (OP == OP_PBW ? a.pcand : a.live);
  uint32_t* vis = OP == OP_FW ? a.fwv : a.bwv;
  const bool part = OP == OP_PBW || (OP == OP_FW && a.part);
  // candidates = live and partition ids mark the dead: the id test alone decides
  const bool any_live = part && a.dead_pid && cand == a.live;
  bool ok[K];
#pragma unroll
  for (int r = 0; r < K; ++r) ok[r] = ok_in[r];
  if (part && a.prune) {
    uint32_t pw[K];
#pragma unroll
    for (int r = 0; r < K; ++r) pw[r] = ok[r] ? a.prune[p[r] >> 5] : 0u;
#pragma unroll
    for (int r = 0; r < K; ++r) ok[r] = ok[r] && !((pw[r] >> (p[r] & 31)) & 1u);
  }
  uint32_t cw[K], vw[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    cw[r] = ok[r] ? (any_live ? 0xFFFFFFFFu : cand[x[r] >> 5]) : 0u;
    vw[r] = ok[r] ? vis[x[r] >> 5] : 0xFFFFFFFFu;
  }
  bool c[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const uint32_t bit = 1u << (x[r] & 31);
    c[r] = (cw[r] & bit) && !(vw[r] & bit);
  }
  if (part) {
    uint32_t pu[K], px[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      pu[r] = c[r] ? a.part[u[r]] : 0u;
      px[r] = c[r] ? a.part[x[r]] : 0u;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
      if (a.prune && c[r] && pu[r] != px[r]) atomicOr(&a.prune[p[r] >> 5], 1u << (p[r] & 31));
      c[r] = c[r] && pu[r] == px[r];
    }
  }
  uint32_t old[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const uint32_t bit = 1u << (x[r] & 31);
    old[r] = c[r] ? atomicOr(&vis[x[r] >> 5], bit) : bit;
  }
  bool cl[K];
  uint64_t d0[K], d1[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    cl[r] = !(old[r] & (1u << (x[r] & 31)));
    d0[r] = cl[r] ? a.noff[x[r]] : 0;
    d1[r] = cl[r] ? a.noff[x[r] + 1] : 0;
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    if (cl[r]) {
      const uint64_t d = d1[r] - d0[r];
      nsum += d;
      nmax = max(nmax, (uint32_t)min(d, (uint64_t)0xFFFFFFFFu));
    }
    bq_push(q, cl[r], x[r]);
  }
}

// The colour ops of a thread's K edges u[r] -> x[r], loads of each step
// issued for all K together. OP_COLOR: a live target of smaller colour in
// the same partition takes the source's colour by atomicMax on its
// (partition, colour) word and is queued once (inq). OP_COLOR_BW: a live
// in-neighbour with the same word is decided (removed from live, labelled
// with the colour's root) and queued.
// Same effect as expand_apply on each edge in turn. With a.prune (p[r]:
// the edge's position), a marked edge is skipped, and an edge whose ends
// are in different partitions (a decided end holds partition 0) is marked.
template <int OP, int K>
__device__ inline void color_batch(const uint32_t (&u)[K], const uint32_t (&x)[K], const bool (&ok_in)[K],
                                   const uint64_t (&p)[K], const OpArgs& a, BlockQueue& q) {
  bool ok[K];
#pragma unroll
  for (int r = 0; r < K; ++r) ok[r] = ok_in[r];
  if (a.prune) {
    uint32_t pw[K];
#pragma unroll
    for (int r = 0; r < K; ++r) pw[r] = ok[r] ? a.prune[p[r] >> 5] : 0u;
#pragma unroll
    for (int r = 0; r < K; ++r) ok[r] = ok[r] && !((pw[r] >> (p[r] & 31)) & 1u);
  }
  unsigned long long xu[K], xw[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    xu[r] = ok[r] ? a.cp[u[r]] : 0ull;
    xw[r] = ok[r] ? a.cp[x[r]] : 0ull;
  }
  if (a.prune) {
#pragma unroll
    for (int r = 0; r < K; ++r)
      if (ok[r] && (xw[r] >> 32) != (xu[r] >> 32)) atomicOr(&a.prune[p[r] >> 5], 1u << (p[r] & 31));
  }
  bool c[K];
  if constexpr (OP == OP_COLOR) {
#pragma unroll
    for (int r = 0; r < K; ++r)
      c[r] = ok[r] && (uint32_t)xw[r] != 0u && (uint32_t)xw[r] < (uint32_t)xu[r] && (xw[r] >> 32) == (xu[r] >> 32);
    unsigned long long old[K];
#pragma unroll
    for (int r = 0; r < K; ++r) old[r] = c[r] ? atomicMax(&a.cp[x[r]], xu[r]) : ~0ull;
    uint32_t qw[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      c[r] = c[r] && (uint32_t)old[r] < (uint32_t)xu[r];
      qw[r] = c[r] ? atomicOr(&a.inq[x[r] >> 5], 1u << (x[r] & 31)) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) bq_push(q, c[r] && !((qw[r] >> (x[r] & 31)) & 1u), x[r]);
  } else {
#pragma unroll
    for (int r = 0; r < K; ++r) c[r] = ok[r] && xw[r] == xu[r];
    uint32_t lw[K];
#pragma unroll
    for (int r = 0; r < K; ++r) lw[r] = c[r] ? a.live[x[r] >> 5] : 0u;
    uint32_t old[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const uint32_t bit = 1u << (x[r] & 31);
      c[r] = c[r] && (lw[r] & bit);
      old[r] = c[r] ? atomicAnd(&a.live[x[r] >> 5], ~bit) : 0u;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const bool push = c[r] && ((old[r] >> (x[r] & 31)) & 1u);
      if (push) a.labels[x[r]] = rotl32(~(uint32_t)xu[r], a.value);
      bq_push(q, push, x[r]);
    }
  }
}

constexpr int EXP_EPT = 8;        // edges per thread of a tile (default; 4 and 16 are also built)
constexpr int EXP_WIN = 1024;     // frontier entries staged in LDS per tile

// Each workgroup takes 2048-edge tiles of the concatenated adjacency of the
// frontier. Thread 0 finds the frontier window covering the tile (two
// binary searches over the scanned degrees), the window (vertex, row start,
// scan value) is staged in LDS, and each thread maps its edges to their
// source by a forward-only LDS search. Consecutive threads read
// consecutive adjacency entries.
template <int OP, int EPT_ = EXP_EPT>
__global__ __launch_bounds__(WG) void k_expand(const uint32_t* __restrict__ F, const uint32_t* fcount,
                                               const uint64_t* __restrict__ S, const uint64_t* __restrict__ off,
                                               const uint32_t* __restrict__ adj, OpArgs a, uint32_t* __restrict__ next,
                                               uint32_t* next_cnt, const uint32_t* __restrict__ tstart,
                                               uint64_t tcap) {
  // the window's entries: scan value relative to the tile start (0 for the
  // first, whose edges may begin before it), row start minus scan value (so
  // edge e of entry j is adj[s_base[j] + e]) and vertex; 16 bytes per entry
  // and a one-tile queue keep the workgroup's LDS at 24 KB with EPT 8 (six
  // workgroups per CU)
  __shared__ int32_t s_S[EXP_WIN];
  __shared__ uint64_t s_base[EXP_WIN];
  __shared__ uint32_t s_v[EXP_WIN];
  __shared__ uint32_t s_i0, s_w;
  constexpr int EXP_TILE = WG * EPT_;  // edges per workgroup tile
  __shared__ uint32_t s_out[EXP_TILE];  // a tile claims at most one vertex per edge; flushed after every tile
  __shared__ uint32_t s_on, s_ob;
  __shared__ unsigned long long s_nsum;
  __shared__ uint32_t s_nmax;
  BlockQueue q{s_out, &s_on, &s_ob, EXP_TILE};
  if (threadIdx.x == 0) {
    s_nsum = 0;
    s_nmax = 0;
  }
  bq_init(q);
  uint64_t nsum = 0;
  uint32_t nmax = 0;
  const uint32_t f = *fcount;
  if (f == 0) return;
  const uint64_t total = S[f];
  const uint64_t ntiles = (total + EXP_TILE - 1) / EXP_TILE;
  // a grid sized before the edge total was known leaves blocks without a tile
  if (blockIdx.x >= ntiles) return;
  if (ntiles > tcap) tstart = nullptr;  // not from the scan of this frontier
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t e0 = t * EXP_TILE;
    const uint64_t e1 = min(e0 + (uint64_t)EXP_TILE, total);
    uint32_t i0, w;
    if (tstart) {
      // the scan recorded the entry holding each tile's first edge; the
      // window ends with the next tile's (whose edges may all lie beyond
      // this tile: the LDS search never selects it then)
      i0 = tstart[t];
      w = (t + 1 < ntiles ? tstart[t + 1] : f - 1) - i0 + 1;
    } else {
      // the last frontier entries whose edges start at or before e0 and at
      // or before e1 - 1: two 128-way searches side by side (threads 0..127
      // and 128..255). Each round probes 128 evenly spaced scan values at
      // once, so a tile waits for about log_128(f) loads instead of
      // 2 log_2(f) dependent ones. Invariant: S[lo] <= e, and hi == f or
      // S[hi] > e; the scan is non-decreasing, so the probes that pass form
      // a prefix.
      const uint32_t half = threadIdx.x >> 7, t = threadIdx.x & 127;
      const uint64_t e = half ? e1 - 1 : e0;
      uint32_t lo = 0, hi = f;
      while (__syncthreads_or(hi - lo > 1)) {
        const uint32_t step = (hi - lo + 127) / 128;
        const uint64_t p = (uint64_t)lo + (uint64_t)t * step;
        const bool pred = hi - lo > 1 && p < hi && S[p] <= e;
        const uint32_t c0 = __syncthreads_count(half == 0 && pred);
        const uint32_t c1 = __syncthreads_count(half == 1 && pred);
        const uint32_t c = half ? c1 : c0;  // >= 1 while the half is still searching (its probe 0 is lo)
        if (hi - lo > 1) {
          hi = (uint32_t)min((uint64_t)lo + (uint64_t)c * step, (uint64_t)hi);
          lo += (c - 1) * step;
        }
      }
      if (threadIdx.x == 0) s_i0 = lo;
      if (threadIdx.x == 128) s_w = lo;
      __syncthreads();
      i0 = s_i0;
      w = s_w - i0 + 1;
    }
    if (w <= (uint32_t)EXP_WIN) {
      for (uint32_t j = threadIdx.x; j < w; j += WG) {
        uint32_t v = F[i0 + j];
        s_v[j] = v;
        const uint64_t sj = S[i0 + j];
        // entries after the first start inside the tile: 0 < sj - e0 < EXP_TILE
        s_S[j] = j ? (int32_t)(sj - e0) : 0;
        s_base[j] = off[v] - sj;
      }
      __syncthreads();
      uint32_t lo = 0;
      if constexpr (OP == OP_FW || OP == OP_BW || OP == OP_PBW || OP == OP_COLOR || OP == OP_COLOR_BW) {
        // the thread's edges of the tile, their claims in flight together
        constexpr int EPT = EXP_TILE / WG;
        uint32_t uu[EPT], xx[EPT];
        uint64_t pp[EPT];
        bool ok[EPT];
#pragma unroll
        for (int r = 0; r < EPT; ++r) {
          const uint64_t e = e0 + threadIdx.x + (uint64_t)r * WG;
          ok[r] = e < e1;
          uu[r] = xx[r] = 0;
          pp[r] = 0;
          if (ok[r]) {
            const int32_t el = (int32_t)(e - e0);
            uint32_t hi = w;
            while (hi - lo > 1) {
              uint32_t mid = (lo + hi) >> 1;
              if (s_S[mid] <= el) lo = mid; else hi = mid;
            }
            uu[r] = s_v[lo];
            pp[r] = s_base[lo] + e;
            xx[r] = adj[pp[r]];
          }
        }
        if constexpr (OP == OP_COLOR || OP == OP_COLOR_BW) color_batch<OP, EPT>(uu, xx, ok, pp, a, q);
        else reach_batch<OP, EPT, BlockQueue>(uu, xx, ok, pp, a, q, nsum, nmax);
      } else {
        for (uint64_t e = e0 + threadIdx.x; e < e1; e += WG) {
          const int32_t el = (int32_t)(e - e0);
          uint32_t hi = w;
          while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (s_S[mid] <= el) lo = mid; else hi = mid;
          }
          uint32_t u = s_v[lo];
          uint32_t x = adj[s_base[lo] + e];
          expand_apply<OP>(u, x, a, q, nsum, nmax);
        }
      }
    } else {
      // window larger than LDS (many zero-degree frontier entries): search the scan in global memory
      uint32_t lo = i0;
      for (uint64_t e = e0 + threadIdx.x; e < e1; e += WG) {
        uint32_t hi = i0 + w;
        while (hi - lo > 1) {
          uint32_t mid = (lo + hi) >> 1;
          if (S[mid] <= e) lo = mid; else hi = mid;
        }
        uint32_t u = F[lo];
        uint32_t x = adj[off[u] + (e - S[lo])];
        expand_apply<OP>(u, x, a, q, nsum, nmax);
      }
    }
    bq_maybe_flush(q, next, next_cnt, EXP_TILE);  // a tile appends at most EXP_TILE entries
  }
  bq_flush(q, next, next_cnt);
  if constexpr (OP == OP_FW || OP == OP_BW || OP == OP_PBW) {
    if (nsum) atomicAdd(&s_nsum, (unsigned long long)nsum);
    if (nmax) atomicMax(&s_nmax, nmax);
    __syncthreads();
    if (threadIdx.x == 0 && s_nsum) atomicAdd(a.nsum, s_nsum);
    if (threadIdx.x == 0 && s_nmax) atomicMax(&a.nsum[1], (unsigned long long)s_nmax);
  }
}

// Push levels over frontiers of small out-degrees, without a degree scan:
// a workgroup takes CHUNK_N consecutive frontier entries, scans their
// out-degrees in LDS and applies OP to the chunk's edges WG * CHUNK_EPT at a time
// (lane l of step r takes edge base + r * WG + l of the chunk, so
// consecutive lanes read consecutive entries of a row) with reach_batch or
// color_batch. The reach levels the host sends here have no vertex of
// out-degree above CHUNK_MAXD, so the chunks of a planted-SCC or lattice
// frontier carry about the same edge count (the colour levels, whose
// degrees the host does not know, take the same path when the options ask
// for it). Claimed vertices, and for the reach ops the next frontier's
// degree sum and maximum, go out as in k_expand. For the reach ops the
// frontier's edge count, which the host knows, is stored to *stot, where a
// degree scan would have left its total; with esum, every chunk adds its
// edge count there (statistics of the colour levels).
constexpr uint32_t CHUNK_PER = 1;                 // frontier entries per thread (1 beat 2 and 4 on planted 5e7)
constexpr uint32_t CHUNK_N = CHUNK_PER * WG;      // frontier entries per chunk
constexpr int CHUNK_EPT = 8;                      // edges per thread of a step
constexpr uint32_t CHUNK_MAXD = 64;               // largest out-degree of a reach frontier the host sends here
constexpr uint32_t CHUNK_QCAP = 2 * WG * CHUNK_EPT;

// The chunks blockIdx.x, blockIdx.x + gridDim.x, ... of frontier F (f
// entries); every thread of the workgroup calls. LDS: s_v, s_b, s_S
// (CHUNK_N, CHUNK_N, CHUNK_N + 1 entries) and s_w (a word per wavefront).
template <int OP>
__device__ inline void chunk_body(const uint32_t* __restrict__ F, uint32_t f, const uint64_t* __restrict__ off,
                                  const uint32_t* __restrict__ adj, const OpArgs& a, BlockQueue& q, uint32_t* next,
                                  uint32_t* next_cnt, uint32_t* err, unsigned long long* esum, uint32_t* s_v,
                                  uint64_t* s_b, uint32_t* s_S, uint64_t* s_w, uint64_t& nsum, uint32_t& nmax) {
  const uint32_t nchunks = (uint32_t)(((uint64_t)f + CHUNK_N - 1) / CHUNK_N);
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const uint64_t i0 = (uint64_t)ch * CHUNK_N;
    uint32_t d[CHUNK_PER];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < CHUNK_PER; ++k) {
      const uint32_t j = CHUNK_PER * threadIdx.x + k;
      uint32_t v = 0;
      uint64_t b = 0;
      d[k] = 0;
      if (i0 + j < f) {
        v = F[i0 + j];
        b = off[v];
        d[k] = (uint32_t)min(off[v + 1] - b, (uint64_t)0xFFFFFFFFu);
        if constexpr (OP == OP_COLOR) {
          if (a.inq_clear) atomicAnd(&a.inq_clear[v >> 5], ~(1u << (v & 31)));
        }
      }
      s_v[j] = v;
      s_b[j] = b;
      sum += d[k];
    }
    const uint64_t incl = wave_incl_scan64(sum);
    if (lane == WAVE - 1) s_w[wid] = incl;
    __syncthreads();
    uint64_t woff = 0, T = 0;
    for (uint32_t w = 0; w < WG / WAVE; ++w) {
      woff += w < wid ? s_w[w] : 0;
      T += s_w[w];
    }
    uint64_t ex = woff + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < CHUNK_PER; ++k) {
      s_S[CHUNK_PER * threadIdx.x + k] = (uint32_t)ex;
      ex += d[k];
    }
    if (threadIdx.x == 0) s_S[CHUNK_N] = (uint32_t)T;
    __syncthreads();
    // the in-chunk scan is 32-bit: a chunk of more edges is reported, not expanded past them
    if (T > 0xFFFFFFFFull && threadIdx.x == 0) *err = 1u;
    const uint32_t total = (uint32_t)min(T, 0xFFFFFFFFull);
    if (esum && threadIdx.x == 0 && total) atomicAdd(esum, (unsigned long long)total);
    uint32_t lo = 0;  // each thread's edges come in increasing order: the search only moves forward
    for (uint32_t base = 0; base < total; base += WG * CHUNK_EPT) {
      uint32_t uu[CHUNK_EPT], xx[CHUNK_EPT];
      uint64_t pp[CHUNK_EPT];
      bool ok[CHUNK_EPT];
#pragma unroll
      for (int r = 0; r < CHUNK_EPT; ++r) {
        const uint32_t e = base + threadIdx.x + (uint32_t)r * WG;
        ok[r] = e < total;
        uu[r] = xx[r] = 0;
        pp[r] = 0;
        if (ok[r]) {
          uint32_t hi = CHUNK_N;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_S[mid] <= e) lo = mid; else hi = mid;
          }
          uu[r] = s_v[lo];
          pp[r] = s_b[lo] + (e - s_S[lo]);
          xx[r] = adj[pp[r]];
        }
      }
      if constexpr (OP == OP_COLOR || OP == OP_COLOR_BW) color_batch<OP, CHUNK_EPT>(uu, xx, ok, pp, a, q);
      else reach_batch<OP, CHUNK_EPT, BlockQueue>(uu, xx, ok, pp, a, q, nsum, nmax);
      bq_maybe_flush(q, next, next_cnt, WG * CHUNK_EPT);  // a step appends at most WG * CHUNK_EPT entries
    }
    __syncthreads();  // the next chunk rewrites the LDS tables
  }
}

template <int OP>
__global__ __launch_bounds__(WG) void k_expand_chunk(const uint32_t* __restrict__ F, const uint32_t* fcount,
                                                     const uint64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ adj, OpArgs a,
                                                     uint32_t* __restrict__ next, uint32_t* next_cnt, uint32_t* err,
                                                     uint64_t* stot, uint64_t etot, unsigned long long* esum) {
  __shared__ uint32_t s_v[CHUNK_N];
  __shared__ uint64_t s_b[CHUNK_N];       // row starts
  __shared__ uint32_t s_S[CHUNK_N + 1];   // exclusive prefix of the degrees within the chunk
  __shared__ uint64_t s_w[WG / WAVE];
  __shared__ uint32_t s_out[CHUNK_QCAP];
  __shared__ uint32_t s_on, s_ob;
  __shared__ unsigned long long s_nsum;
  __shared__ uint32_t s_nmax;
  BlockQueue q{s_out, &s_on, &s_ob, CHUNK_QCAP};
  if (threadIdx.x == 0) {
    s_nsum = 0;
    s_nmax = 0;
  }
  bq_init(q);
  if (stot && blockIdx.x == 0 && threadIdx.x == 0) *stot = etot;
  uint64_t nsum = 0;
  uint32_t nmax = 0;
  chunk_body<OP>(F, *fcount, off, adj, a, q, next, next_cnt, err, esum, s_v, s_b, s_S, s_w, nsum, nmax);
  bq_flush(q, next, next_cnt);
  if constexpr (OP == OP_FW || OP == OP_BW || OP == OP_PBW) {
    if (nsum) atomicAdd(&s_nsum, (unsigned long long)nsum);
    if (nmax) atomicMax(&s_nmax, nmax);
    __syncthreads();
    if (threadIdx.x == 0 && s_nsum) atomicAdd(a.nsum, s_nsum);
    if (threadIdx.x == 0 && s_nmax) atomicMax(&a.nsum[1], (unsigned long long)s_nmax);
  }
}

// Colour push levels chained on the device (KSPN_FLAG_CHUNKED): launch L
// of a chain expands the frontier of level L (count ring[L % 3], buffer
// F0 for even L, F1 for odd) into the other buffer (count ring[(L + 1) % 3])
// with the chunk body, queueing into inq0 (even L) or inq1 (odd L) and
// clearing its frontier's bits in the other (for L = 0 only with
// first_clear), and zeroes ring[(L + 2) % 3] for the launch after it. A
// level whose frontier is empty or more than half of the u live vertices
// (the host's sweep rule) ends the chain: it records its index in
// end_level, and the launches after it return at once.
struct ColorChain {
  uint32_t ring[3];
  uint32_t ended;
  uint32_t end_level;
  uint32_t pad[3];
};

__global__ __launch_bounds__(WG) void k_color_chain(uint32_t* F0, uint32_t* F1, const uint64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ adj, OpArgs a, uint32_t* inq0,
                                                    uint32_t* inq1, bool first_clear, ColorChain* ctl, uint32_t u,
                                                    uint32_t L, uint32_t* err, unsigned long long* esum) {
  __shared__ uint32_t s_v[CHUNK_N];
  __shared__ uint64_t s_b[CHUNK_N];
  __shared__ uint32_t s_S[CHUNK_N + 1];
  __shared__ uint64_t s_w[WG / WAVE];
  __shared__ uint32_t s_out[CHUNK_QCAP];
  __shared__ uint32_t s_on, s_ob, s_f, s_stop;
  if (threadIdx.x == 0) {
    const uint32_t ended = ctl->ended;
    const uint32_t f = ctl->ring[L % 3];
    s_f = f;
    s_stop = ended || f == 0 || (double)f * 2.0 > (double)u;
    if (!ended && s_stop && blockIdx.x == 0) {
      ctl->ended = 1u;
      ctl->end_level = L;
    }
  }
  __syncthreads();
  if (s_stop) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->ring[(L + 2) % 3] = 0u;
  BlockQueue q{s_out, &s_on, &s_ob, CHUNK_QCAP};
  bq_init(q);
  OpArgs b = a;
  b.inq = (L & 1) ? inq1 : inq0;
  b.inq_clear = (L == 0 && !first_clear) ? nullptr : ((L & 1) ? inq0 : inq1);
  uint64_t nsum = 0;
  uint32_t nmax = 0;
  uint32_t* N = (L & 1) ? F0 : F1;
  uint32_t* nc = &ctl->ring[(L + 1) % 3];
  chunk_body<OP_COLOR>((L & 1) ? F1 : F0, s_f, off, adj, b, q, N, nc, err, esum, s_v, s_b, s_S, s_w, nsum, nmax);
  bq_flush(q, N, nc);
}

constexpr uint32_t kColorChain = 8;  // launches per chain of colour push levels

// start of a chain: ring[0] = *fcount, the other slots and the end record zeroed
__global__ void k_color_chain_init(ColorChain* ctl, const uint32_t* fcount) {
  if (threadIdx.x == 0) {
    ctl->ring[0] = *fcount;
    ctl->ring[1] = 0u;
    ctl->ring[2] = 0u;
    ctl->ended = 0u;
    ctl->end_level = 0u;
  }
}

// end of a chain of `len` launches: the level it stopped at (len if none
// stopped) to *level, that level's frontier count to *fcount
__global__ void k_color_chain_end(const ColorChain* ctl, uint32_t len, uint32_t* fcount, uint32_t* level) {
  if (threadIdx.x == 0) {
    const uint32_t e = ctl->ended ? ctl->end_level : len;
    *level = e;
    *fcount = ctl->ring[e % 3];
  }
}

// =====================================================================
// persistent multi-level reachability
// =====================================================================
// One persistent launch runs push levels back to back with a grid barrier
// between them, so a level costs a barrier instead of three launches and a
// host round trip. Wavefronts take 64-vertex chunks of the frontier from a
// counter (dynamic balance), map the chunk's edges to their sources with a
// wavefront scan of the degrees staged in LDS, claim targets with the same
// bitmap atomics as k_expand and stage claimed vertices in a
// wavefront-private LDS queue (one global atomic per flush). Frontiers
// whose degree sum exceeds e_max or that hold a vertex of out-degree above
// d_max, and the pull switch (Beamer: degree sum * alpha > unexpanded
// edges), end the launch: the host continues with k_expand / k_pull.

constexpr int COOP_WG = 1024;
constexpr int COOP_WAVES = COOP_WG / WAVE;
constexpr uint32_t COOP_QCAP = 256;           // staged entries per wavefront
constexpr uint32_t COOP_SPIN_LIMIT = 1u << 21;  // barrier polls before giving up (about a second)
constexpr uint32_t COOP_LOCAL = 8;              // rounds a wavefront expands its own claims within a level
enum CoopReason : uint32_t { COOP_EMPTY = 1, COOP_PULL = 2, COOP_BIG = 3, COOP_ERR = 4 };

// relaxed device-scope loads of values other workgroups change within a
// level (counters, claims); what a level wrote is ordered before the next
// level's reads by the release/acquire pair of coop_barrier
__device__ inline uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_relaxed64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The next level's frontier size, degree sum, maximum degree and the error
// flag, read by thread 0 and shared with the workgroup.
struct CoopLevel {
  uint32_t f, fmax, err;
  unsigned long long fe, lexp, lvert;
};

// slot of the level just finished (lexp, lvert) is the one before `slot`
__device__ inline void coop_read_level(const CoopCtl* ctl, uint32_t slot, CoopLevel* lv) {
  lv->f = ld_relaxed(&ctl->cnt[slot]);
  lv->fe = ld_relaxed64(&ctl->nsum[slot]);
  lv->fmax = ld_relaxed(&ctl->maxd[slot]);
  lv->err = ld_relaxed(&ctl->err);
  lv->lexp = ld_relaxed64(&ctl->lexp[slot]);
  lv->lvert = ld_relaxed64(&ctl->lvert[slot]);
}

// Barrier over the workgroups of a persistent launch (resident together on
// an otherwise idle device). Everything one workgroup writes and another
// reads inside the launch goes through device-scope atomics (counters,
// bitmap claims) or device-scope relaxed stores and loads (the frontier
// queues). The HIP memory model orders them across the barrier: each wave
// releases at agent scope before the workgroup barrier, thread 0 arrives
// with a relaxed atomic add, polls with relaxed loads and acquires at agent
// scope once all have arrived. The arrival counter only grows: barrier k of
// a launch waits for k * workgroups arrivals. Thread 0 then reads the next
// level's values into *lv. Returns false if the others did not arrive within
// the spin limit or one of them gave up: a launch that is not fully
// resident ends with an error instead of waiting forever.
__device__ bool coop_barrier(CoopCtl* ctl, uint32_t k, uint32_t next_slot, CoopLevel* lv) {
  __shared__ uint32_t s_ok;
  // every wave releases its writes at agent scope before the workgroup
  // arrives (the workgroup barrier alone does not wait for other waves'
  // stores), and thread 0 acquires after the others arrived
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    const uint32_t target = k * gridDim.x;
    const uint32_t arrived = __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived + 1 < target) {
      uint32_t spins = 0;
      while (ld_relaxed(&ctl->bar_count) < target) {
        if (++spins > COOP_SPIN_LIMIT || ld_relaxed(&ctl->err)) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    coop_read_level(ctl, next_slot, lv);
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// a full wavefront queue goes out on its own (one atomic per flush)
__device__ inline void coop_flush(uint32_t* sq, uint32_t& qn, uint32_t* q, uint32_t* qcount) {
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(qcount, qn);
  base = __shfl(base, 0);
  __builtin_amdgcn_wave_barrier();
  for (uint32_t k = lane; k < qn; k += WAVE) __hip_atomic_store(&q[base + k], sq[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_wave_barrier();
  qn = 0;
}

// LDS of a workgroup expanding one level (persistent or chained launches)
struct CoopLds {
  uint64_t ex[COOP_WAVES][WAVE];
  uint64_t beg[COOP_WAVES][WAVE];
  uint32_t v[COOP_WAVES][WAVE];
  uint32_t q[COOP_WAVES][COOP_QCAP];
  uint32_t qn[COOP_WAVES];
  unsigned long long sum[COOP_WAVES];
  unsigned long long lexp[COOP_WAVES], lvert[COOP_WAVES];
  uint32_t max[COOP_WAVES];
  uint32_t base;
};

// One level of a reach by every workgroup of the launch: the frontier F
// (f vertices, out-degree sum fe) into N, whose count, degree sum, largest
// degree and local-round totals go to ring slot s1; chunks beyond the
// first pass come from ctl->work[s0]. `waves` is the launch's wavefront
// count, the same in every workgroup.
template <int OP>
__device__ void coop_level(CoopLds& S, uint32_t n, const uint32_t* F, uint32_t* N, uint32_t f,
                           unsigned long long fe, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ adj, const OpArgs& a, CoopCtl* ctl, uint32_t s0,
                           uint32_t s1, uint32_t d_max) {
  const uint32_t wid = threadIdx.x / WAVE, lane = lane_id();
  const uint64_t below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  const uint32_t waves = gridDim.x * COOP_WAVES;
  const uint32_t gw = blockIdx.x * COOP_WAVES + wid;
  uint64_t my_sum = 0;
  uint32_t my_max = 0, qn = 0;
  // edge-balanced expansion of up to 64 sources (one per lane: v, row
  // start b, degree d); claims go to the wavefront's LDS queue, a full
  // queue out to the next frontier
  auto expand = [&](uint32_t v, uint64_t b, uint64_t d) {
    const uint64_t incl = wave_incl_scan64(d);
    const uint64_t total = __shfl(incl, WAVE - 1);
    S.ex[wid][lane] = incl - d;
    S.beg[wid][lane] = b;
    S.v[wid][lane] = v;
    __builtin_amdgcn_wave_barrier();
    uint32_t lo = 0;
    for (uint64_t e0 = 0; e0 < total; e0 += WAVE) {
      const uint64_t e = e0 + lane;
      bool push = false;
      uint32_t w = 0;
      if (e < total) {
        uint32_t hi = WAVE;  // last source whose edges start at or before e
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (S.ex[wid][mid] <= e) lo = mid; else hi = mid;
        }
        w = adj[S.beg[wid][lo] + (e - S.ex[wid][lo])];
        push = reach_claim<OP>(S.v[wid][lo], w, a);
        if (push) {
          const uint64_t dd = a.noff[w + 1] - a.noff[w];
          my_sum += dd;
          my_max = max(my_max, (uint32_t)min(dd, (uint64_t)0xFFFFFFFFu));
        }
      }
      const uint64_t mask = __ballot(push);
      if (mask) {
        if (push) S.q[wid][qn + __popcll(mask & below)] = w;
        qn += (uint32_t)__popcll(mask);
        if (qn > COOP_QCAP - WAVE) coop_flush(S.q[wid], qn, N, &ctl->cnt[s1]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    return total;
  };
  // chunk gw first (no atomic); beyond the first pass chunks come from a counter
  const uint64_t nchunks = ((uint64_t)f + WAVE - 1) / WAVE;
  uint64_t chunk = gw;
  while (chunk < nchunks) {
    const uint64_t i = chunk * WAVE + lane;
    uint32_t v = 0;
    uint64_t b = 0, d = 0;
    if (i < f) {
      v = __hip_atomic_load(&F[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (v < n) {  // always true; an id out of range would be skipped, not used as an index
        b = off[v];
        d = off[v + 1] - b;
      }
    }
    expand(v, b, d);
    if (nchunks <= waves) break;
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(&ctl->work[s0], 1u);
    chunk = (uint64_t)waves + __shfl(k, 0);
  }
  // Reachability needs each claimed vertex expanded once, not level by
  // level. On a thin frontier of low degree (at most 16 vertices per
  // wavefront, at most 4 edges per vertex: long-diameter graphs) a
  // wavefront goes on with its own claims for a few rounds while they fit
  // one round, so such a frontier takes fewer levels; other frontiers
  // wait for the next level, where all wavefronts share them. A claim of
  // degree above d_max stays for the next level (its degree was counted
  // for it).
  uint64_t loc_e = 0, loc_v = 0;
  const bool thin = f <= 16ull * waves && fe <= 4ull * f;
  for (uint32_t r = 0; thin && r < COOP_LOCAL && qn > 0 && qn <= WAVE; ++r) {
    const uint32_t take = min(qn, (uint32_t)WAVE);
    qn -= take;
    const bool have = lane < take;
    uint32_t v = have ? S.q[wid][qn + lane] : 0u;
    uint64_t b = 0, d = 0;
    if (have) {
      b = off[v];
      d = off[v + 1] - b;
    }
    __builtin_amdgcn_wave_barrier();
    const bool big = have && d > d_max;
    const uint64_t bm = __ballot(big);
    if (bm) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&ctl->cnt[s1], (uint32_t)__popcll(bm));
      base = __shfl(base, 0);
      if (big) __hip_atomic_store(&N[base + __popcll(bm & below)], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (have && !big) my_sum -= d;  // not in the next frontier after all (a.noff is off)
    loc_v += (uint32_t)__popcll(__ballot(have && !big));
    loc_e += expand(v, b, big ? 0 : d);
  }
  // the workgroup's staged entries, degree sum and maximum: one atomic each
  my_sum = wave_sum64(my_sum);
  for (int o = 32; o > 0; o >>= 1) my_max = max(my_max, __shfl_down(my_max, o));
  if (lane == 0) {
    S.sum[wid] = my_sum;
    S.max[wid] = my_max;
    S.qn[wid] = qn;
    S.lexp[wid] = loc_e;
    S.lvert[wid] = loc_v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long bs = 0, le = 0, lv = 0;
    uint32_t bm = 0, bq = 0;
    for (int k = 0; k < COOP_WAVES; ++k) {
      bs += S.sum[k];
      bm = max(bm, S.max[k]);
      le += S.lexp[k];
      lv += S.lvert[k];
      const uint32_t q = S.qn[k];
      S.qn[k] = bq;  // exclusive prefix
      bq += q;
    }
    if (bs) atomicAdd(&ctl->nsum[s1], bs);
    if (bm) atomicMax(&ctl->maxd[s1], bm);
    if (le) atomicAdd(&ctl->lexp[s1], le);
    if (lv) atomicAdd(&ctl->lvert[s1], lv);
    S.base = bq ? atomicAdd(&ctl->cnt[s1], bq) : 0u;
  }
  __syncthreads();
  {
    const uint32_t dst = S.base + S.qn[wid];
    for (uint32_t k = lane; k < qn; k += WAVE)
      __hip_atomic_store(&N[dst + k], S.q[wid][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// why a run of push levels on the device ends before the next level (the
// same in every workgroup: computed from values written before the level)
__device__ inline uint32_t coop_reason(uint32_t f, unsigned long long fe, uint32_t fmax, double m_rest, float alpha,
                                       unsigned long long e_max, uint32_t d_max) {
  if (f == 0) return COOP_EMPTY;
  if (f > 64 && (double)fe * alpha > m_rest) return COOP_PULL;
  if (fe > e_max || fmax > d_max) return COOP_BIG;
  return 0;
}

// the ring slot after next (last read before this level) cleared for reuse
__device__ inline void coop_clear_slot(CoopCtl* ctl, uint32_t s2) {
  __hip_atomic_store(&ctl->cnt[s2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&ctl->work[s2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&ctl->nsum[s2], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&ctl->maxd[s2], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&ctl->lexp[s2], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&ctl->lvert[s2], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int OP>
__global__ __launch_bounds__(COOP_WG) void k_reach_coop(uint32_t n, uint32_t* F0, uint32_t* F1, const uint64_t* __restrict__ off,
                                                        const uint32_t* __restrict__ adj, OpArgs a, CoopCtl* ctl,
                                                        double m_rest, float alpha, unsigned long long e_max,
                                                        uint32_t d_max) {
  __shared__ CoopLds S;
  __shared__ CoopLevel s_lv;
  unsigned long long expanded = 0, verts = 0;
  uint32_t L = 0;
  if (threadIdx.x == 0) coop_read_level(ctl, 0, &s_lv);  // written by the init kernel before the launch
  __syncthreads();
  for (;;) {
    const uint32_t s0 = L % 3, s1 = (L + 1) % 3, s2 = (L + 2) % 3;
    const uint32_t f = s_lv.f, fmax = s_lv.fmax;
    const unsigned long long fe = s_lv.fe;
    // the local rounds of the level before (slot s0 holds what they added)
    expanded += s_lv.lexp;
    verts += s_lv.lvert;
    m_rest -= (double)s_lv.lexp;
    if (s_lv.err) return;  // a workgroup timed out at a barrier
    const uint32_t reason = coop_reason(f, fe, fmax, m_rest, alpha, e_max, d_max);
    if (reason) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->levels = L;
        ctl->reason = reason;
        ctl->expanded = expanded;
        ctl->verts = verts;
      }
      return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) coop_clear_slot(ctl, s2);
    coop_level<OP>(S, n, (L & 1) ? F1 : F0, (L & 1) ? F0 : F1, f, fe, off, adj, a, ctl, s0, s1, d_max);
    expanded += fe;
    verts += f;
    m_rest -= (double)fe;
    ++L;
    if (!coop_barrier(ctl, L, L % 3, &s_lv)) {
      if (threadIdx.x == 0) {
        atomicOr(&ctl->err, 1u);
        if (blockIdx.x == 0) {
          ctl->levels = L;
          ctl->reason = COOP_ERR;
          ctl->expanded = expanded;
          ctl->verts = verts;
        }
      }
      return;
    }
  }
}

// Chained push levels: level L of a run as a launch of its own, enqueued
// back to back with the others of the chain and no host round trip in
// between; the kernel boundaries order one level's writes before the next
// level's reads, so no grid barrier or fence is needed and the grid need
// not be resident at once. Each launch reads its level's values from ring
// slot L % 3 and the remaining-edge estimate from ctl->mrest[L % 3]; the
// first launch whose level ends the run records levels and reason, and the
// launches after it return at once. Block 0 keeps the run's totals.
template <int OP>
__global__ __launch_bounds__(COOP_WG) void k_chain_level(uint32_t n, uint32_t* F0, uint32_t* F1,
                                                         const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ adj, OpArgs a, CoopCtl* ctl,
                                                         float alpha, unsigned long long e_max, uint32_t d_max,
                                                         uint32_t L) {
  __shared__ CoopLds S;
  __shared__ CoopLevel s_lv;
  __shared__ double s_m;
  if (threadIdx.x == 0) {
    const uint32_t ended = ctl->reason;  // set by an earlier launch of the chain
    coop_read_level(ctl, L % 3, &s_lv);
    s_lv.err = (ended || s_lv.err) ? 1u : 0u;
    s_m = ctl->mrest[L % 3] - (double)s_lv.lexp;
  }
  __syncthreads();
  if (s_lv.err) return;
  const uint32_t s0 = L % 3, s1 = (L + 1) % 3, s2 = (L + 2) % 3;
  const uint32_t f = s_lv.f;
  const unsigned long long fe = s_lv.fe;
  const uint32_t reason = coop_reason(f, fe, s_lv.fmax, s_m, alpha, e_max, d_max);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the local rounds of the level before, then (unless the run ends here) this level
    ctl->expanded += s_lv.lexp + (reason ? 0ull : fe);
    ctl->verts += s_lv.lvert + (reason ? 0ull : f);
    if (reason) {
      ctl->levels = L;
      ctl->reason = reason;
    } else {
      ctl->mrest[s1] = s_m - (double)fe;
      coop_clear_slot(ctl, s2);
    }
  }
  if (reason) return;
  coop_level<OP>(S, n, (L & 1) ? F1 : F0, (L & 1) ? F0 : F1, f, fe, off, adj, a, ctl, s0, s1, d_max);
}

// after a chain of `len` launches that did not end the run: its levels,
// and the local rounds of its last level in the totals
__global__ void k_chain_end(CoopCtl* ctl, uint32_t len) {
  if (threadIdx.x == 0 && ctl->reason == 0) {
    ctl->levels = len;
    ctl->expanded += ctl->lexp[len % 3];
    ctl->verts += ctl->lvert[len % 3];
  }
}

// Narrow push levels: frontiers whose vertices have at most NARROW_DEG
// out-edges each (planted-SCC graphs, lattices), chained like
// k_chain_level but with no degree scan and no edge-balanced tiling. A
// wavefront takes 64 frontier vertices, lane L loading the row of vertex
// L. Lanes form eight groups of eight; in each step lane k of group g
// takes edge e0 + k of the rows of vertices g, g + 8, ..., g + 56, so
// eight lanes read eight consecutive entries of one row, and each lane has
// eight claims in flight together (reach_batch). A wavefront steps to the
// largest degree among its 64 rows. Claims are staged in the wavefront's
// LDS queue, one global atomic per flush; the next frontier's degree sum
// and maximum take one atomic each per wavefront. The chain ends, as
// k_chain_level's does, at an empty frontier, at the pull switch, or at a
// frontier holding a vertex of degree above NARROW_DEG (the host expands
// that level edge-balanced), and at a frontier of more than e_max out-edges
// (NARROW_EDGES: the edge-balanced expansion is faster on larger frontiers;
// planted 5e7 was fastest at 2^19 among 2^16 .. 2^20).
constexpr uint32_t NARROW_DEG = 64;
constexpr uint64_t NARROW_EDGES = 1ull << 19;
constexpr int NARROW_WG = 256;
constexpr uint32_t NARROW_QCAP = 1024;  // staged claims per wavefront (a step adds at most 512)
constexpr uint32_t kNarrowChain = 8;    // launches per chain (opts.chain_levels overrides)

template <int OP>
__global__ __launch_bounds__(NARROW_WG) void k_narrow_level(uint32_t* F0, uint32_t* F1,
                                                            const uint64_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj, OpArgs a,
                                                            CoopCtl* ctl, float alpha, unsigned long long e_max,
                                                            uint32_t L) {
  __shared__ uint32_t s_q[NARROW_WG / WAVE][NARROW_QCAP];
  __shared__ CoopLevel s_lv;
  __shared__ double s_m;
  if (threadIdx.x == 0) {
    const uint32_t ended = ctl->reason;  // set by an earlier launch of the chain
    coop_read_level(ctl, L % 3, &s_lv);
    s_lv.err = (ended || s_lv.err) ? 1u : 0u;
    s_m = ctl->mrest[L % 3];
  }
  __syncthreads();
  if (s_lv.err) return;
  const uint32_t s1 = (L + 1) % 3, s2 = (L + 2) % 3;
  const uint32_t f = s_lv.f;
  const unsigned long long fe = s_lv.fe;
  const uint32_t reason = coop_reason(f, fe, s_lv.fmax, s_m, alpha, e_max, NARROW_DEG);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (reason) {
      ctl->levels = L;
      ctl->reason = reason;
    } else {
      ctl->expanded += fe;
      ctl->verts += f;
      ctl->mrest[s1] = s_m - (double)fe;
      coop_clear_slot(ctl, s2);
    }
  }
  if (reason) return;
  const uint32_t* F = (L & 1) ? F1 : F0;
  uint32_t* N = (L & 1) ? F0 : F1;
  const uint32_t wid = threadIdx.x / WAVE, lane = lane_id(), k = lane & 7, grp = lane >> 3;
  WaveQueue q{s_q[wid], 0};
  uint64_t nsum = 0;
  uint32_t nmax = 0;
  const uint64_t stride = (uint64_t)gridDim.x * NARROW_WG;
  for (uint64_t wb = (uint64_t)blockIdx.x * NARROW_WG + (uint64_t)wid * WAVE; wb < f; wb += stride) {
    const uint64_t i = wb + lane;
    uint32_t v = 0, d = 0;
    uint64_t b = 0;
    if (i < f) {
      v = F[i];
      b = off[v];
      d = (uint32_t)min(off[v + 1] - b, (uint64_t)0xFFFFFFFFu);
    }
    uint32_t dm = d;
    for (int o = 32; o > 0; o >>= 1) dm = max(dm, (uint32_t)__shfl_xor(dm, o));
    for (uint32_t e0 = 0; e0 < dm; e0 += 8) {
      uint32_t uu[8], xx[8];
      uint64_t pp[8];
      bool ok[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int src = grp + 8 * r;
        const uint32_t vr = __shfl(v, src), dr = __shfl(d, src);
        const uint64_t br = __shfl((unsigned long long)b, src);
        const uint32_t e = e0 + k;
        ok[r] = e < dr;
        uu[r] = vr;
        pp[r] = br + e;
        xx[r] = ok[r] ? adj[pp[r]] : 0u;
      }
      if (q.n > NARROW_QCAP - 8 * WAVE) wq_flush(q, N, &ctl->cnt[s1]);
      reach_batch<OP, 8, WaveQueue>(uu, xx, ok, pp, a, q, nsum, nmax);
    }
  }
  wq_flush(q, N, &ctl->cnt[s1]);
  // the next frontier's degree sum and maximum: one atomic each per
  // workgroup (the wavefronts of a launch finish together, and adds to one
  // address from all of them would queue)
  __shared__ unsigned long long s_sum[NARROW_WG / WAVE];
  __shared__ uint32_t s_max[NARROW_WG / WAVE];
  nsum = wave_sum64(nsum);
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, (uint32_t)__shfl_xor(nmax, o));
  if (lane == 0) {
    s_sum[wid] = nsum;
    s_max[wid] = nmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long bs = 0;
    uint32_t bm = 0;
    for (int w = 0; w < NARROW_WG / WAVE; ++w) {
      bs += s_sum[w];
      bm = max(bm, s_max[w]);
    }
    if (bs) atomicAdd(&ctl->nsum[s1], bs);
    if (bm) atomicMax(&ctl->maxd[s1], bm);
  }
}

// control block for a launch: level 0 is the host's frontier (count at
// *fcount, degree sum and maximum known)
// =====================================================================
// asynchronous reachability
// =====================================================================
// One launch runs a whole reach without levels. Work items are row slices
// (AS_SLICE edges) of claimed vertices in one linear queue: every vertex is
// claimed once, so a reach never appends more than n + m / AS_SLICE entries.
// A wavefront takes 64 queue slots at a time and waits for them to be
// written; whatever entries are ready it expands as one batch (their edges
// concatenated, four per lane, the loads of the four issued together),
// claims targets with the bitmap atomics of the level kernels, and stages
// the claimed vertices' slices in LDS before reserving queue slots for them.
// Slots are written once with agent-scope stores and polled with
// agent-scope loads; the consumer writes them back to AS_EMPTY, so the
// queue is empty again when the reach ends. One counter, `pending`, holds
// the entries reserved and not yet expanded. A producer adds its entries to
// it, and waits for that atomic to return, before it writes their slots; a
// batch subtracts itself only after its children are counted. So a consumer
// can subtract an entry only after the producer's addition for it, and the
// counter reaches 0 exactly when the reach is over: a wavefront waiting on
// slots that will never be written sees that and leaves. Waits are bounded:
// after AS_SPIN polls without progress the kernel sets `err` and every
// wavefront leaves; a reach that ends with `pending` above 0 sets it too.
constexpr uint32_t AS_SLICE = 256;
constexpr unsigned long long AS_EMPTY = ~0ull;
constexpr int AS_WG = 256;
constexpr int AS_WAVES = AS_WG / WAVE;
constexpr uint32_t AS_STAGE = 256;
constexpr uint32_t AS_SPIN = 1u << 24;

struct AsyncCtl {
  uint32_t head, tail, err, pad;
  unsigned long long pending;  // entries reserved and not yet expanded
  unsigned long long edges;    // edges expanded (statistics)
  unsigned long long verts;    // entries expanded
  unsigned long long claims;   // vertices claimed
};

__device__ inline unsigned long long ld_agent64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_async_init(AsyncCtl* ctl) {
  if (threadIdx.x == 0) *ctl = AsyncCtl{};
}

// n queue slots reserved and counted in `pending`; the slot index is
// returned only after both atomics have returned, so the count is in place
// before any of the slots is written
__device__ inline uint32_t as_reserve(uint32_t n, AsyncCtl* ctl) {
  uint32_t base = 0;
  if (lane_id() == 0) {
    base = atomicAdd(&ctl->tail, n);
    const unsigned long long before = atomicAdd(&ctl->pending, (unsigned long long)n);
    if (before == AS_EMPTY) base = 0xFFFFFFFFu;  // never true; keeps the wait for the return
  }
  return __shfl(base, 0);
}

// Slots for the staged entries, written with agent-scope stores; returns how
// many were reserved. A queue that would overflow sets err instead.
__device__ inline uint32_t as_flush(unsigned long long* st, uint32_t n, unsigned long long* Q, uint32_t qcap,
                                    AsyncCtl* ctl) {
  if (n == 0) return 0;
  const uint32_t lane = lane_id();
  const uint32_t base = as_reserve(n, ctl);
  if ((uint64_t)base + n > qcap) {
    if (lane == 0) atomicOr(&ctl->err, 2u);
  } else {
    for (uint32_t i = lane; i < n; i += WAVE) st_agent64(&Q[base + i], st[i]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return n;
}

// the slices of the lanes' claimed vertices (take, deg) into the stage;
// every lane calls together. *made counts the slots reserved by flushes.
__device__ inline void as_stage(unsigned long long* st, uint32_t& sn, bool take, uint32_t v, uint64_t deg,
                                unsigned long long* Q, uint32_t qcap, AsyncCtl* ctl, uint32_t& made) {
  const bool any = take && deg;
  if (!__ballot(any)) return;
  const uint32_t lane = lane_id();
  const uint32_t k = any ? (uint32_t)((deg + AS_SLICE - 1) / AS_SLICE) : 0u;
  uint32_t incl = k;
  for (uint32_t o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  const uint32_t total = __shfl(incl, WAVE - 1);
  if (sn + total > AS_STAGE) {
    made += as_flush(st, sn, Q, qcap, ctl);
    sn = 0;
  }
  if (total > AS_STAGE) {  // a hub's many slices: straight to the queue
    const uint32_t base = as_reserve(total, ctl);
    if ((uint64_t)base + total > qcap) {
      if (lane == 0) atomicOr(&ctl->err, 2u);
    } else {
      for (uint32_t s = 0; s < k; ++s) st_agent64(&Q[base + incl - k + s], ((unsigned long long)v << 32) | s);
    }
    made += total;
    return;
  }
  for (uint32_t s = 0; s < k; ++s) st[sn + incl - k + s] = ((unsigned long long)v << 32) | s;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  sn += total;
}

// the frontier F[0..*fcount) (vis bits already set) as the queue's first entries
__global__ __launch_bounds__(WG) void k_async_seed(const uint32_t* __restrict__ F, const uint32_t* fcount,
                                                   const uint64_t* __restrict__ off, unsigned long long* Q,
                                                   uint32_t qcap, AsyncCtl* ctl) {
  __shared__ unsigned long long s_st[WG / WAVE][AS_STAGE];
  const uint32_t f = *fcount;
  const uint32_t wid = threadIdx.x / WAVE;
  uint32_t sn = 0, made = 0;
  const uint32_t pad = (f + WG - 1) / WG * WG;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < pad; i += gridDim.x * WG) {
    const bool ok = i < f;
    const uint32_t v = ok ? F[i] : 0u;
    const uint64_t deg = ok ? off[v + 1] - off[v] : 0;
    as_stage(s_st[wid], sn, ok, v, deg, Q, qcap, ctl, made);
  }
  made += as_flush(s_st[wid], sn, Q, qcap, ctl);
}

template <int OP>
__global__ __launch_bounds__(AS_WG) void k_reach_async(const uint64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ adj, OpArgs a,
                                                       unsigned long long* Q, uint32_t qcap, AsyncCtl* ctl) {
  __shared__ uint64_t s_beg[AS_WAVES][WAVE];
  __shared__ uint32_t s_ex[AS_WAVES][WAVE], s_src[AS_WAVES][WAVE];
  __shared__ unsigned long long s_st[AS_WAVES][AS_STAGE];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t* cand = OP == OP_BW ? a.fwv : (OP == OP_PBW ? a.pcand : a.live);
  uint32_t* vis = OP == OP_FW ? a.fwv : a.bwv;
  const bool part = OP == OP_PBW || (OP == OP_FW && a.part != nullptr);
  uint64_t* sb = s_beg[wid];
  uint32_t* sx = s_ex[wid];
  uint32_t* su = s_src[wid];
  unsigned long long* st = s_st[wid];
  unsigned long long edges = 0, verts = 0, claims = 0;
  bool over = false;
  while (!over) {
    uint32_t u0 = 0;
    if (lane == 0) u0 = atomicAdd(&ctl->head, (uint32_t)WAVE);
    u0 = __shfl(u0, 0);
    if (u0 >= qcap) break;  // never written: the queue cannot hold that many entries
    const uint32_t slot = u0 + lane;
    bool mine = slot < qcap;  // slots of this range not yet expanded
    uint32_t spins = 0;
    while (__ballot(mine)) {
      const unsigned long long e = mine ? ld_agent64(&Q[slot]) : AS_EMPTY;
      const bool ready = mine && e != AS_EMPTY;
      if (!__ballot(ready)) {
        unsigned long long pend = 0;
        uint32_t err = 0;
        if (lane == 0) {
          pend = ld_agent64(&ctl->pending);
          err = __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        pend = __shfl(pend, 0);
        err = __shfl(err, 0);
        if (pend == 0 || err) {  // nothing outstanding anywhere: these slots stay empty
          over = true;
          break;
        }
        if (++spins > AS_SPIN) {
          if (lane == 0) atomicOr(&ctl->err, 1u);
          over = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      spins = 0;
      // the ready entries as one batch
      uint32_t v = 0, cnt = 0;
      uint64_t b = 0;
      if (ready) {
        v = (uint32_t)(e >> 32);
        const uint32_t sl = (uint32_t)e;
        const uint64_t re = off[v + 1];
        b = off[v] + (uint64_t)sl * AS_SLICE;
        cnt = (uint32_t)(min(b + (uint64_t)AS_SLICE, re) - b);
      }
      uint32_t incl = cnt;
      for (uint32_t o = 1; o < WAVE; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      const uint32_t total = __shfl(incl, WAVE - 1);
      sb[lane] = b;
      sx[lane] = incl - cnt;
      su[lane] = v;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      uint32_t sn = 0, made = 0;
      for (uint32_t base = 0; base < total; base += 4 * WAVE) {
        uint32_t w[4], u[4];
        bool ok[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t x = base + q * WAVE + lane;
          ok[q] = x < total;
          w[q] = u[q] = 0;
          if (ok[q]) {
            uint32_t lo = 0, hi = WAVE;  // last entry whose scan is <= x
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (sx[mid] <= x) lo = mid; else hi = mid;
            }
            u[q] = su[lo];
            w[q] = adj[sb[lo] + (x - sx[lo])];
          }
        }
        uint32_t cw[4], vw[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          cw[q] = ok[q] ? cand[w[q] >> 5] : 0u;
          vw[q] = ok[q] ? vis[w[q] >> 5] : 0xFFFFFFFFu;
        }
        bool c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t bit = 1u << (w[q] & 31);
          c[q] = (cw[q] & bit) && !(vw[q] & bit);
        }
        if (part) {
          uint32_t pu[4], pw[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            pu[q] = c[q] ? a.part[u[q]] : 0u;
            pw[q] = c[q] ? a.part[w[q]] : 0u;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) c[q] = c[q] && pu[q] == pw[q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t bit = 1u << (w[q] & 31);
          if (c[q]) c[q] = (atomicOr(&vis[w[q] >> 5], bit) & bit) == 0;
        }
        uint64_t d0[4], d1[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          d0[q] = c[q] ? off[w[q]] : 0;
          d1[q] = c[q] ? off[w[q] + 1] : 0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          claims += c[q];
          as_stage(st, sn, c[q], w[q], d1[q] - d0[q], Q, qcap, ctl, made);
        }
      }
      made += as_flush(st, sn, Q, qcap, ctl);
      if (ready) st_agent64(&Q[slot], AS_EMPTY);
      const uint32_t nready = (uint32_t)__popcll(__ballot(ready));
      // this batch out, after its children were counted in (as_reserve)
      if (lane == 0) atomicAdd(&ctl->pending, 0ull - (unsigned long long)nready);
      if (lane == 0) {
        edges += total;
        verts += nready;
      }
      mine = mine && !ready;
    }
  }
  claims = wave_sum64(claims);
  if (lane == 0) {
    if (edges) atomicAdd(&ctl->edges, edges);
    if (verts) atomicAdd(&ctl->verts, verts);
    if (claims) atomicAdd(&ctl->claims, claims);
  }
}

// after an asynchronous reach: its statistics into cnt64[0..1], its error
// into cnt[C_AERR]; both frontier counters zeroed (as a level-by-level
// reach leaves them)
__global__ void k_async_done(const AsyncCtl* ctl, uint32_t* cnt, unsigned long long* cnt64) {
  if (threadIdx.x == 0) {
    cnt64[0] = ctl->edges;
    cnt64[1] = ctl->claims;
    cnt[C_AERR] = ctl->err | (ctl->pending ? 4u : 0u);
    cnt[C_Q0] = 0;
    cnt[C_Q1] = 0;
  }
}

__global__ void k_coop_init(CoopCtl* ctl, const uint32_t* fcount, unsigned long long edges, uint32_t maxd,
                            double m_rest = 0.0) {
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; ++k) {
      ctl->cnt[k] = 0;
      ctl->work[k] = 0;
      ctl->nsum[k] = 0;
      ctl->maxd[k] = 0;
      ctl->lexp[k] = 0;
      ctl->lvert[k] = 0;
    }
    ctl->cnt[0] = *fcount;
    ctl->nsum[0] = edges;
    ctl->maxd[0] = maxd;
    ctl->expanded = 0;
    ctl->verts = 0;
    ctl->bar_count = 0;
    ctl->levels = ctl->reason = ctl->err = 0;
    ctl->mrest[0] = m_rest;
  }
}

// after a launch: the last frontier's size into the host's counter slot of
// the buffer that holds it, the other slot and the level counters zeroed
// for a following host level; results into the mapped host counters
__global__ void k_coop_publish(const CoopCtl* ctl, uint32_t* cnt_even, uint32_t* cnt_odd, unsigned long long* cnt64,
                               uint32_t* map, uint32_t seq) {
  if (threadIdx.x == 0) {
    const uint32_t L = ctl->levels;
    const uint32_t f = ctl->cnt[L % 3];
    *((L & 1) ? cnt_odd : cnt_even) = f;
    *((L & 1) ? cnt_even : cnt_odd) = 0;  // the host's next level appends there
    cnt64[0] = cnt64[1] = cnt64[2] = 0;
    const unsigned long long fe = ctl->nsum[L % 3], ex = ctl->expanded;
    map[32] = L;
    map[33] = ctl->reason;
    map[34] = ctl->err;
    map[35] = f;
    map[36] = (uint32_t)fe;
    map[37] = (uint32_t)(fe >> 32);
    map[38] = ctl->maxd[L % 3];
    map[39] = (uint32_t)ex;
    map[40] = (uint32_t)(ex >> 32);
    map[41] = (uint32_t)ctl->verts;
    map[42] = (uint32_t)(ctl->verts >> 32);
  }
  publish_seq(map, seq);
}

// =====================================================================
// trim-1 first pass + degree-product pivot
// =====================================================================

// One pass over all vertices: in- or out-degree 0 -> decided with its own
// id; the rest form the live bitmap (written as two words per wavefront
// ballot). Each workgroup reduces its (max in*out, min id) pair.
// Without `labels` the labels stay for k_label_bits's dense pass. Besides
// the live bitmap it writes zeros to the words of inq, fwv and bwv
// it owns, and block 0 zeroes the ncnt counters cnt and the four cnt64:
// the state a solve starts from, without fill launches of its own.
__global__ __launch_bounds__(WG) void k_trim_first(uint32_t n, const uint64_t* __restrict__ fw_off,
                                                   const uint64_t* __restrict__ bw_off, uint32_t* __restrict__ live,
                                                   uint32_t* __restrict__ labels, uint64_t* __restrict__ part,
                                                   uint32_t* __restrict__ inq, uint32_t* __restrict__ fwv,
                                                   uint32_t* __restrict__ bwv, uint32_t* __restrict__ cnt,
                                                   uint32_t ncnt, unsigned long long* __restrict__ cnt64) {
  if (blockIdx.x == 0) {
    for (uint32_t i = threadIdx.x; i < ncnt; i += WG) cnt[i] = 0;
    if (threadIdx.x < 4) cnt64[threadIdx.x] = 0;
  }
  __shared__ uint64_t s_p[WG];
  __shared__ uint32_t s_id[WG];
  __shared__ uint32_t s_red[WG / WAVE];
  uint32_t dead = 0;
  uint64_t best = 0;
  uint32_t best_id = 0xFFFFFFFFu;
  // TRIM_U vertices per thread and step (one per 256-vertex tile), their
  // offset loads issued together
  constexpr int TRIM_U = 4;
  constexpr uint64_t SPAN = (uint64_t)WG * TRIM_U;
  const uint64_t n_pad = ((uint64_t)n + SPAN - 1) / SPAN * SPAN;
  const uint64_t n_pad256 = ((uint64_t)n + WG - 1) / WG * WG;
  const uint32_t lane = lane_id();
  for (uint64_t base = (uint64_t)blockIdx.x * SPAN; base < n_pad; base += (uint64_t)gridDim.x * SPAN) {
    uint64_t fo[TRIM_U], fo1[TRIM_U], bo[TRIM_U], bo1[TRIM_U];
#pragma unroll
    for (int k = 0; k < TRIM_U; ++k) {
      const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
      const bool ok = v < n;
      fo[k] = ok ? fw_off[v] : 0;
      fo1[k] = ok ? fw_off[v + 1] : 0;
      bo[k] = ok ? bw_off[v] : 0;
      bo1[k] = ok ? bw_off[v + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < TRIM_U; ++k) {
      const uint64_t v64 = base + (uint64_t)k * WG + threadIdx.x;
      const uint64_t out = fo1[k] - fo[k];
      const uint64_t in = bo1[k] - bo[k];
      const bool is_live = v64 < n && in != 0 && out != 0;
      if (v64 < n) {
        const uint32_t v = (uint32_t)v64;
        if (labels) labels[v] = is_live ? UNDEC : v;
        if (is_live) {
          const uint64_t p = in * out;
          if (p > best || (p == best && v < best_id)) {
            best = p;
            best_id = v;
          }
        }
      }
      const uint64_t mask = __ballot(is_live);
      const uint64_t first = v64 - lane;  // the wavefront's first vertex of tile k
      // words up to the 256-vertex padding get written (zero past n), as the
      // bitmaps are allocated; the steps' 1024-vertex padding owns none
      if (first < n_pad256 && (lane == 0 || lane == 32)) {
        const uint64_t w = (first >> 5) + (lane >> 5);
        live[w] = (uint32_t)(mask >> lane);
        inq[w] = 0;
        fwv[w] = 0;
        bwv[w] = 0;
      }
      dead += (v64 < n && !is_live);
    }
  }
  uint32_t dead_sum = block_sum(dead, s_red);
  s_p[threadIdx.x] = best;
  s_id[threadIdx.x] = best_id;
  __syncthreads();
  for (int s = WG / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      uint64_t p2 = s_p[threadIdx.x + s];
      uint32_t i2 = s_id[threadIdx.x + s];
      if (p2 > s_p[threadIdx.x] || (p2 == s_p[threadIdx.x] && i2 < s_id[threadIdx.x])) {
        s_p[threadIdx.x] = p2;
        s_id[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s_p[0];
    part[3 * blockIdx.x + 1] = s_id[0];
    part[3 * blockIdx.x + 2] = dead_sum;
  }
}

constexpr uint32_t RED_U = 16;  // single-workgroup reductions: partials per thread loaded together

// What k_pivot_reduce seeds for FW-BW from the pivot p it finds, into
// bitmaps and counters k_trim_first zeroed: bit p of fwv and bwv, the
// forward frontier queue = [p] with *qcount = 1, and the exclusive scan of
// its out-degree S = [0, deg(p)] (the first level's, so the forward reach
// starts with its expansion). No pivot: *qcount = 0, S = [0].
struct FwbwSeed {
  uint32_t* fwv;
  uint32_t* bwv;
  uint32_t* queue;
  uint32_t* qcount;
  const uint64_t* off;
  uint64_t* S;
};

// (max product, smallest id) over workgroup partials -> out[0] (0xFFFFFFFF
// if none); summed partial counts -> *count_out
__global__ __launch_bounds__(WG) void k_pivot_reduce(const uint64_t* part, uint32_t nparts, uint32_t* out,
                                                     uint32_t* count_out, FwbwSeed seed = {}) {
  __shared__ uint64_t s_p[WG];
  __shared__ uint32_t s_id[WG];
  __shared__ uint32_t s_red[WG / WAVE];
  uint64_t best = 0;
  uint32_t best_id = 0xFFFFFFFFu;
  uint32_t total = 0;
  // RED_U partials per thread and step, loaded together
  for (uint32_t i0 = threadIdx.x; i0 < nparts; i0 += WG * RED_U) {
    uint64_t x[RED_U][3];
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      const uint32_t i = i0 + r * WG;
      x[r][0] = i < nparts ? part[3 * i] : 0;
      x[r][1] = i < nparts ? part[3 * i + 1] : 0xFFFFFFFFull;
      x[r][2] = i < nparts ? part[3 * i + 2] : 0;
    }
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      const uint64_t p = x[r][0];
      const uint32_t id = (uint32_t)x[r][1];
      total += (uint32_t)x[r][2];
      if (id != 0xFFFFFFFFu && (p > best || (p == best && id < best_id))) {
        best = p;
        best_id = id;
      }
    }
  }
  total = block_sum(total, s_red);
  if (threadIdx.x == 0 && count_out) *count_out = total;
  s_p[threadIdx.x] = best;
  s_id[threadIdx.x] = best_id;
  __syncthreads();
  for (int s = WG / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      uint64_t p2 = s_p[threadIdx.x + s];
      uint32_t i2 = s_id[threadIdx.x + s];
      if (i2 != 0xFFFFFFFFu && (s_id[threadIdx.x] == 0xFFFFFFFFu || p2 > s_p[threadIdx.x] ||
                                (p2 == s_p[threadIdx.x] && i2 < s_id[threadIdx.x]))) {
        s_p[threadIdx.x] = p2;
        s_id[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t p = s_id[0];
    *out = p;
    if (seed.queue) {
      const bool any = p != 0xFFFFFFFFu;
      if (any) {
        seed.fwv[p >> 5] = 1u << (p & 31);
        seed.bwv[p >> 5] = 1u << (p & 31);
        seed.queue[0] = p;
        seed.S[1] = seed.off[p + 1] - seed.off[p];
      }
      *seed.qcount = any ? 1u : 0u;
      seed.S[0] = 0;
    }
  }
}

// =====================================================================
// pull (bottom-up) reachability levels
// =====================================================================

// Pull levels over all vertices (grid-stride, no barrier inside the loop).
// Forward (DIR 0): a live v not yet reached is reached when any in-neighbour
// is. Backward (DIR 1): a forward-reached v not yet reached backward is
// reached when any out-neighbour is. One random bit per scanned edge; claims
// made earlier in the same sweep are visible. Newly reached vertices are
// written as a bitmap (each wavefront owns its two words, no atomics); per
// workgroup partials: claimed count, edges scanned, next-level degree sum
// (noff, optional) of the claimed vertices, or with `sbm` the number of
// candidates left unreached, whose bits it receives (written like fnew) so
// that a short remainder can be compacted into a list for k_pull_list.
// Block 0 zeroes the four list counters `lcnt`, which the sweep does not use.
// With `gate`, the sweep does nothing unless *gate is set.
constexpr int PULL_ILP = 2;  // vertices in flight per thread (2 beat 1 and 4 on R-MAT 24 and the 4096-graph batch)
constexpr uint32_t PULL_FLUSH = 4;                     // k_pull_list: iterations between checks of its LDS queues
constexpr uint32_t PULL_QCAP = PULL_FLUSH * WG * PULL_ILP;  // entries of each LDS queue (8 KiB)

template <int DIR, int ILP>
__global__ __launch_bounds__(WG) void k_pull(uint32_t n, const uint32_t* __restrict__ cand, uint32_t* vis,
                                             const uint64_t* __restrict__ roff, const uint32_t* __restrict__ radj,
                                             const uint64_t* __restrict__ noff, uint32_t* __restrict__ fnew,
                                             uint64_t* __restrict__ part, const uint32_t* __restrict__ pid,
                                             uint32_t* cursor, uint32_t* __restrict__ sbm, uint32_t* lcnt,
                                             const uint32_t* gate) {
  if (gate && *gate == 0) return;  // a speculative sweep that lost (SpecGate)
  // A candidate costs a chain of dependent loads (bitmap word, row offsets,
  // in-neighbour, its visited bit, its partition). Each thread carries
  // ILP vertices (one per 256-vertex tile of its workgroup's span)
  // through every step together, so their loads overlap.
  __shared__ uint64_t s_red[WG / WAVE];
  uint64_t claimed = 0, edges = 0, nd = 0;
  constexpr uint64_t SPAN = (uint64_t)WG * ILP;
  const uint64_t n_pad = ((uint64_t)n + SPAN - 1) / SPAN * SPAN;
  const uint32_t lane = lane_id();
  for (uint64_t base = (uint64_t)blockIdx.x * SPAN; base < n_pad; base += (uint64_t)gridDim.x * SPAN) {
    uint32_t vw[ILP], pv[ILP];
    uint64_t it[ILP], end[ILP];
    bool act[ILP], hit[ILP], cnd[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
      const uint64_t word = v >> 5;
      vw[k] = (v < n) ? vis[word] : 0xFFFFFFFFu;
      const uint32_t cw = (v < n) ? cand[word] : 0u;
      const uint32_t bit = 1u << (v & 31);
      act[k] = (cw & bit) && !(vw[k] & bit);
      cnd[k] = act[k];
      hit[k] = false;
      it[k] = end[k] = 0;
      pv[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      if (act[k]) {
        const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
        it[k] = roff[v];
        end[k] = roff[v + 1];
        if (pid) pv[k] = pid[v];
      }
    }
    // scan the rows together, one in-neighbour of every unfinished row per step
    for (;;) {
      bool more = false;
      uint32_t u[ILP];
#pragma unroll
      for (int k = 0; k < ILP; ++k) {
        act[k] = act[k] && it[k] < end[k];
        more |= act[k];
        u[k] = act[k] ? radj[it[k]] : 0u;
      }
      if (!more) break;
      uint32_t uw[ILP];
#pragma unroll
      for (int k = 0; k < ILP; ++k) uw[k] = act[k] ? vis[u[k] >> 5] : 0u;
#pragma unroll
      for (int k = 0; k < ILP; ++k) {
        if (!act[k]) continue;
        ++edges;
        ++it[k];
        if ((uw[k] >> (u[k] & 31)) & 1u) {
          if (!pid || pid[u[k]] == pv[k]) {
            hit[k] = true;
            act[k] = false;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
      if (hit[k]) {
        ++claimed;
        if (noff) nd += noff[v + 1] - noff[v];
      }
      const bool left = cnd[k] && !hit[k];
      if (sbm && left) ++nd;
      // only this wavefront sets bits in its two words during a pull level,
      // so the claim is a plain store of (old | new) per word
      const uint64_t mask = __ballot(hit[k]);
      const uint64_t lmask = __ballot(left);
      const uint64_t first = v - lane;  // the wavefront's first vertex of tile k
      const uint64_t word0 = first >> 5;
      if (first < n) {  // the padding past n owns no bitmap words
        if (lane == 0) {
          fnew[word0] = (uint32_t)mask;
          if ((uint32_t)mask) vis[word0] = vw[k] | (uint32_t)mask;
          if (sbm) sbm[word0] = (uint32_t)lmask;
        }
        if (lane == 32) {
          fnew[word0 + 1] = (uint32_t)(mask >> 32);
          if ((uint32_t)(mask >> 32)) vis[word0 + 1] = vw[k] | (uint32_t)(mask >> 32);
          if (sbm) sbm[word0 + 1] = (uint32_t)(lmask >> 32);
        }
      }
    }
  }
  claimed = block_sum64(claimed, s_red);
  edges = block_sum64(edges, s_red);
  nd = block_sum64(nd, s_red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = claimed;
    part[3 * blockIdx.x + 1] = edges;
    part[3 * blockIdx.x + 2] = nd;
    if (blockIdx.x == 0) {
      *cursor = 0;  // for a compaction of this level's bitmap
      if (lcnt) lcnt[0] = lcnt[1] = lcnt[2] = lcnt[3] = 0;
    }
  }
}

// A pull level over a list of candidates: the vertices the previous pull
// level left unreached (each once, none reached yet). The scan is k_pull's;
// a claim sets the vertex's bit with an atomic OR (list neighbours need not
// share a bitmap word) and, when `claimed_q` is given, appends it there
// (the frontier of a push level that may follow). Candidates still
// unreached go to `out`. Several list levels run back to back without the
// host: each reads its candidate count from the device, block 0 zeroes the
// counter `zero` (neither read nor written here: the next level's output),
// and every block adds its scanned edges to *edges_sum when it is given
// (statistics only: one add per workgroup on one address is not free).
template <int DIR>
__global__ __launch_bounds__(WG) void k_pull_list(const uint32_t* __restrict__ list, const uint32_t* list_cnt,
                                                  uint32_t* vis, const uint64_t* __restrict__ roff,
                                                  const uint32_t* __restrict__ radj, const uint32_t* __restrict__ pid,
                                                  uint32_t* __restrict__ claimed_q, uint32_t* claimed_cnt,
                                                  uint32_t* __restrict__ out, uint32_t* out_cnt, uint32_t* zero,
                                                  uint32_t* zero2, unsigned long long* edges_sum,
                                                  const uint32_t* prev_cnt) {
  __shared__ uint64_t s_red[WG / WAVE];
  __shared__ uint32_t s_cq[PULL_QCAP], s_sq[PULL_QCAP];
  __shared__ uint32_t s_cn, s_cb, s_sn, s_sb;
  BlockQueue cq{s_cq, &s_cn, &s_cb, PULL_QCAP}, sq{s_sq, &s_sn, &s_sb, PULL_QCAP};
  bq_init(cq);
  bq_init(sq);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *zero = 0;
    if (zero2) *zero2 = 0;
  }
  const uint32_t cnt = *list_cnt;
  // every candidate of the previous level stayed unreached: it claimed
  // nothing, so nothing is left to reach. This level returns (leaving its
  // output count zero), and the batch's later levels find no candidates.
  if (prev_cnt && *prev_cnt == cnt) return;
  uint32_t iter = 0;
  uint64_t edges = 0;
  constexpr uint32_t SPAN = (uint32_t)WG * PULL_ILP;
  const uint32_t cnt_pad = (cnt + SPAN - 1) / SPAN * SPAN;
  for (uint32_t base = blockIdx.x * SPAN; base < cnt_pad; base += gridDim.x * SPAN) {
    uint32_t vv[PULL_ILP], pv[PULL_ILP];
    uint64_t it[PULL_ILP], end[PULL_ILP];
    bool act[PULL_ILP], hit[PULL_ILP], cnd[PULL_ILP];
#pragma unroll
    for (int k = 0; k < PULL_ILP; ++k) {
      const uint32_t i = base + (uint32_t)k * WG + threadIdx.x;
      cnd[k] = act[k] = i < cnt;
      vv[k] = act[k] ? list[i] : 0u;
      hit[k] = false;
      it[k] = end[k] = 0;
      pv[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < PULL_ILP; ++k) {
      if (act[k]) {
        it[k] = roff[vv[k]];
        end[k] = roff[vv[k] + 1];
        if (pid) pv[k] = pid[vv[k]];
      }
    }
    for (;;) {
      bool more = false;
      uint32_t u[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) {
        act[k] = act[k] && it[k] < end[k];
        more |= act[k];
        u[k] = act[k] ? radj[it[k]] : 0u;
      }
      if (!more) break;
      uint32_t uw[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) uw[k] = act[k] ? vis[u[k] >> 5] : 0u;
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) {
        if (!act[k]) continue;
        ++edges;
        ++it[k];
        if ((uw[k] >> (u[k] & 31)) & 1u) {
          if (!pid || pid[u[k]] == pv[k]) {
            hit[k] = true;
            act[k] = false;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PULL_ILP; ++k) {
      if (hit[k]) atomicOr(&vis[vv[k] >> 5], 1u << (vv[k] & 31));
      if (claimed_q) bq_push(cq, hit[k], vv[k]);
      bq_push(sq, cnd[k] && !hit[k], vv[k]);
    }
    if ((++iter % PULL_FLUSH) == 0) {
      if (claimed_q) bq_maybe_flush(cq, claimed_q, claimed_cnt, PULL_FLUSH * SPAN);
      bq_maybe_flush(sq, out, out_cnt, PULL_FLUSH * SPAN);
    }
  }
  if (claimed_q) bq_flush(cq, claimed_q, claimed_cnt);
  bq_flush(sq, out, out_cnt);
  if (edges_sum) {
    edges = block_sum64(edges, s_red);
    if (threadIdx.x == 0 && edges) atomicAdd(edges_sum, (unsigned long long)edges);
  }
}

// sums of the three per-workgroup partials: [0] -> *count, [1] -> cnt64[0], [2] -> cnt64[1]
__global__ __launch_bounds__(WG) void k_reduce3(const uint64_t* __restrict__ part, uint32_t nparts, uint32_t* count,
                                                unsigned long long* cnt64) {
  __shared__ uint64_t s_red[WG / WAVE];
  uint64_t a = 0, b = 0, d = 0;
  // RED_U partials per thread and step, loaded together
  for (uint32_t i0 = threadIdx.x; i0 < nparts; i0 += WG * RED_U) {
    uint64_t x[RED_U][3];
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      const uint32_t i = i0 + r * WG;
#pragma unroll
      for (int j = 0; j < 3; ++j) x[r][j] = i < nparts ? part[3 * i + j] : 0;
    }
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      a += x[r][0];
      b += x[r][1];
      d += x[r][2];
    }
  }
  a = block_sum64(a, s_red);
  b = block_sum64(b, s_red);
  d = block_sum64(d, s_red);
  if (threadIdx.x == 0) {
    *count = (uint32_t)a;
    cnt64[0] = b;
    cnt64[1] = d;
  }
}

// After a pull sweep (its reduction and survivor listing in one launch):
// every workgroup sums the sweep's per-workgroup partials (claims, scanned
// edges, candidates left; at most RED_U * WG of them), block 0 writes the
// sums to *claims, cnt64[0] and cnt64[1], and when at most list_max
// candidates are left the workgroups list their bits in sbm as
// k_bitmap_to_queue does (*cursor zero before the launch). With `gate`,
// nothing unless *gate is set. Launch b2q_grid(words) workgroups.
__global__ __launch_bounds__(WG) void k_sweep_list(const uint64_t* __restrict__ part, uint32_t nparts, uint32_t* claims,
                                                   unsigned long long* cnt64, const uint32_t* __restrict__ sbm,
                                                   uint64_t words, uint32_t* __restrict__ list, uint32_t* cursor,
                                                   unsigned long long list_max, const uint32_t* gate) {
  __shared__ uint64_t s_red[WG / WAVE];
  __shared__ unsigned long long s_left;
  __shared__ uint32_t s_wave[WG / WAVE];
  __shared__ uint32_t s_base;
  if (gate && *gate == 0) return;  // a speculative sweep that lost (SpecGate)
  uint64_t a = 0, b = 0, d = 0;
  for (uint32_t i0 = threadIdx.x; i0 < nparts; i0 += WG * RED_U) {
    uint64_t x[RED_U][3];
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      const uint32_t i = i0 + r * WG;
#pragma unroll
      for (int j = 0; j < 3; ++j) x[r][j] = i < nparts ? part[3 * i + j] : 0;
    }
#pragma unroll
    for (uint32_t r = 0; r < RED_U; ++r) {
      a += x[r][0];
      b += x[r][1];
      d += x[r][2];
    }
  }
  a = block_sum64(a, s_red);
  b = block_sum64(b, s_red);
  d = block_sum64(d, s_red);
  if (threadIdx.x == 0) {
    s_left = d;
    if (blockIdx.x == 0) {
      *claims = (uint32_t)a;
      cnt64[0] = b;
      cnt64[1] = d;
    }
  }
  __syncthreads();
  if (s_left > list_max) return;  // too many for list levels: the host sweeps again
  for (uint64_t base = (uint64_t)blockIdx.x * B2Q_CHUNK; base < words; base += (uint64_t)gridDim.x * B2Q_CHUNK) {
    uint32_t w[B2Q_STEPS][4];
#pragma unroll
    for (uint32_t t = 0; t < B2Q_STEPS; ++t) {
      const uint64_t i0 = base + (uint64_t)t * WG * 4 + (uint64_t)threadIdx.x * 4;
      if (i0 + 4 <= words) {
        const uint4 q = *reinterpret_cast<const uint4*>(sbm + i0);  // i0 is a multiple of 4 words
        w[t][0] = q.x;
        w[t][1] = q.y;
        w[t][2] = q.z;
        w[t][3] = q.w;
      } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) w[t][k] = i0 + k < words ? sbm[i0 + k] : 0u;
      }
    }
    chunk_to_list(w, base, list, cursor, s_wave, &s_base);
  }
}

// degrees of a frontier of at most 4096 vertices and their exclusive scan
// (S[f] = total) in one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_deg_scan_small(const uint32_t* __restrict__ F, const uint32_t* fcount,
                                                         const uint64_t* __restrict__ off, uint64_t* __restrict__ S) {
  __shared__ uint64_t s[1024];
  const uint32_t f = *fcount;
  const uint32_t t = threadIdx.x;
  uint64_t d[4];
  uint64_t sum = 0;
  for (int k = 0; k < 4; ++k) {
    uint32_t i = 4 * t + k;
    d[k] = (i < f) ? off[F[i] + 1] - off[F[i]] : 0;
    sum += d[k];
  }
  s[t] = sum;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    uint64_t x = (t >= o) ? s[t - o] : 0;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  uint64_t run = s[t] - sum;  // exclusive prefix of this thread's block
  for (int k = 0; k < 4; ++k) {
    uint32_t i = 4 * t + k;
    if (i <= f) S[i] = run;
    run += d[k];
  }
}

// Exclusive scan of the frontier's out-degrees, S[0..f] (S[f] = total), in
// one pass with decoupled look-back (Merrill and Garland, "Single-pass
// parallel prefix scan with decoupled look-back", 2016). Workgroups take
// tiles of DS_TILE entries in ticket order, so every earlier tile belongs
// to a workgroup that is already running. A tile publishes its aggregate,
// then its inclusive prefix once the earlier tiles' values are known; one
// wavefront reads 64 earlier status words at a time. Status words carry
// the launch's epoch (bits 50..63), so they need no clearing between
// launches; values are below 2^48. The last ticket of a launch resets the
// ticket counter for the next one. The launch has exactly `tiles`
// workgroups, ceil((f + 1) / DS_TILE) for the f the host knows; a device
// count that does not fit sets *err. With tstart, each entry i of non-zero
// degree also writes tstart[t] = i for every expansion tile t < tcap (tiles
// of 2^tile_shift edges) whose first edge is one of its own, so k_expand
// reads a tile's frontier window instead of searching the scan for it.
constexpr uint32_t DS_ITEMS = 8;
constexpr uint32_t DS_TILE = WG * DS_ITEMS;
constexpr unsigned long long DS_AGG = 1ull << 48, DS_INC = 2ull << 48, DS_FLAGS = 3ull << 48;
constexpr unsigned long long DS_VALUE = (1ull << 48) - 1;
constexpr uint32_t DS_EPOCHS = 1u << 14;
constexpr uint32_t DS_SPIN = 1u << 22;

__global__ __launch_bounds__(WG) void k_deg_scan(const uint32_t* __restrict__ F, const uint32_t* fcount,
                                                 const uint64_t* __restrict__ off, uint64_t* __restrict__ S,
                                                 unsigned long long* st, uint32_t* ticket, uint32_t tiles,
                                                 uint32_t epoch, uint32_t* err, uint32_t* __restrict__ tstart,
                                                 uint32_t tile_shift, uint64_t tcap) {
  __shared__ uint32_t s_t;
  __shared__ uint64_t s_w[WG / WAVE];
  __shared__ uint64_t s_excl;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  if (threadIdx.x == 0) {
    const uint32_t t = atomicAdd(ticket, 1u);
    if (t + 1 == tiles) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_t = t;
  }
  __syncthreads();
  const uint32_t t = s_t;
  if (t >= tiles) {  // a ticket counter out of step (not expected): report, write nothing
    if (threadIdx.x == 0) *err = 1u;
    return;
  }
  const uint32_t f = *fcount;
  if (t == 0 && threadIdx.x == 0 && (uint64_t)f + 1 > (uint64_t)tiles * DS_TILE) *err = 1u;
  const uint64_t i0 = (uint64_t)t * DS_TILE + (uint64_t)threadIdx.x * DS_ITEMS;
  uint32_t v[DS_ITEMS];
  uint64_t d[DS_ITEMS];
  uint64_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < DS_ITEMS; ++k) v[k] = i0 + k < f ? F[i0 + k] : 0u;
#pragma unroll
  for (uint32_t k = 0; k < DS_ITEMS; ++k) {
    d[k] = i0 + k < f ? off[v[k] + 1] - off[v[k]] : 0;
    sum += d[k];
  }
  // workgroup scan of the threads' sums
  const uint64_t incl = wave_incl_scan64(sum);
  if (lane == WAVE - 1) s_w[wid] = incl;
  __syncthreads();
  uint64_t wave_off = 0, agg = 0;
  for (uint32_t k = 0; k < WG / WAVE; ++k) {
    wave_off += k < wid ? s_w[k] : 0;
    agg += s_w[k];
  }
  const unsigned long long tag = (unsigned long long)epoch << 50;
  if (wid == 0) {
    uint64_t excl = 0;
    if (t == 0) {
      if (lane == 0) __hip_atomic_store(&st[0], tag | DS_INC | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(&st[t], tag | DS_AGG | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int64_t p = (int64_t)t - 1;  // nearest predecessor not yet summed
      bool failed = false;
      for (;;) {
        const int64_t q = p - (int64_t)lane;
        unsigned long long w = tag | DS_INC;  // before tile 0: an inclusive prefix of 0
        if (q >= 0) {
          uint32_t spins = 0;
          for (;;) {
            w = __hip_atomic_load(&st[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((w & ~(DS_FLAGS | DS_VALUE)) == tag && (w & DS_FLAGS)) break;
            if (++spins > DS_SPIN) {
              failed = true;
              w = tag | DS_INC;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        const uint64_t incm = __ballot((w & DS_FLAGS) == DS_INC);
        const uint32_t j = incm ? (uint32_t)__ffsll((unsigned long long)incm) - 1 : WAVE - 1;
        uint64_t x = lane <= j ? (w & DS_VALUE) : 0;
        x = __shfl(wave_sum64(x), 0);
        excl += x;
        if (incm) break;
        p -= WAVE;
      }
      if (__ballot(failed) && lane == 0) *err = 1u;
      if (lane == 0) __hip_atomic_store(&st[t], tag | DS_INC | (excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  uint64_t run = s_excl + wave_off + incl - sum;
#pragma unroll
  for (uint32_t k = 0; k < DS_ITEMS; ++k) {
    if (i0 + k <= f) S[i0 + k] = run;
    if (tstart && d[k]) {
      // the tiles whose first edge lies in [run, run + d)
      const uint64_t mask = (1ull << tile_shift) - 1;
      const uint64_t t1 = min((run + d[k] - 1) >> tile_shift, tcap - 1);
      for (uint64_t t = (run + mask) >> tile_shift; t <= t1; ++t) tstart[t] = (uint32_t)(i0 + k);
    }
    run += d[k];
  }
}

// zero a queue counter and the 64-bit level counters
__global__ void k_level_reset(uint32_t* cnt_slot, unsigned long long* cnt64, uint32_t* s0 = nullptr,
                              uint32_t* s1 = nullptr) {
  if (threadIdx.x == 0) {
    *cnt_slot = 0;
    cnt64[0] = 0;
    cnt64[1] = 0;
    cnt64[2] = 0;
    if (s0) *s0 = 0;
    if (s1) *s1 = 0;
  }
}

// The start of a reach from one vertex whose bit is already set: queue =
// [*vdev] with *qcount = 1 (0 and nothing else when *vdev is 0xFFFFFFFF),
// the level counters k_level_reset zeroes, and the first level's degree
// scan S = [0, deg]
__global__ void k_reach_start(const uint32_t* vdev, uint32_t* queue, uint32_t* qcount, uint32_t* cnt_slot,
                              unsigned long long* cnt64, const uint64_t* __restrict__ off, uint64_t* S) {
  if (threadIdx.x == 0) {
    const uint32_t v = *vdev;
    const bool any = v != 0xFFFFFFFFu;
    if (any) {
      queue[0] = v;
      S[1] = off[v + 1] - off[v];
    }
    *qcount = any ? 1u : 0u;
    S[0] = 0;
    *cnt_slot = 0;
    cnt64[0] = cnt64[1] = cnt64[2] = 0;
  }
}

// After FW-BW: scc = bwv (backward-reached within forward-reached), live &=
// ~scc. Each workgroup takes chunks of FIN_CHUNK words (launch
// finish_grid(words) workgroups, at most FINISH_PARTS) and leaves the
// partials of |scc| (part[2b]) and of its smallest vertex (part[2b + 1])
// that k_label_bits reduces. With `list`, the vertices left live are
// appended to it as k_bitmap_to_queue appends (*cursor zero before the
// launch). Block 0 zeroes the given counters `zero` (of the kernels after it).
constexpr uint32_t FINISH_PARTS = WG;
constexpr uint32_t FIN_STEPS = 2;  // 16-byte words per thread and chunk
constexpr uint64_t FIN_CHUNK = (uint64_t)WG * 4 * FIN_STEPS;
struct Zero3 {
  uint32_t* p[3];
};
__global__ __launch_bounds__(WG) void k_fwbw_finish(uint32_t* live, const uint32_t* __restrict__ bwv, uint64_t words,
                                                    uint64_t* __restrict__ part, uint32_t* __restrict__ list,
                                                    uint32_t* cursor, Zero3 zero) {
  __shared__ uint32_t s_wave[WG / WAVE];
  __shared__ uint32_t s_base;
  __shared__ uint32_t s_red[WG / WAVE];
  if (blockIdx.x == 0 && threadIdx.x < 3 && zero.p[threadIdx.x]) *zero.p[threadIdx.x] = 0;
  uint32_t count = 0, mn = 0xFFFFFFFFu;
  for (uint64_t base = (uint64_t)blockIdx.x * FIN_CHUNK; base < words; base += (uint64_t)gridDim.x * FIN_CHUNK) {
    uint32_t w[FIN_STEPS][4];  // live words after the SCC left
#pragma unroll
    for (uint32_t t = 0; t < FIN_STEPS; ++t) {
      const uint64_t i0 = base + (uint64_t)t * WG * 4 + (uint64_t)threadIdx.x * 4;
      uint32_t sc[4];
      if (i0 + 4 <= words) {  // i0 is a multiple of 4 words: 16-byte accesses
        const uint4 a = *reinterpret_cast<const uint4*>(bwv + i0);
        const uint4 l = *reinterpret_cast<const uint4*>(live + i0);
        sc[0] = a.x, sc[1] = a.y, sc[2] = a.z, sc[3] = a.w;
        w[t][0] = l.x & ~a.x, w[t][1] = l.y & ~a.y, w[t][2] = l.z & ~a.z, w[t][3] = l.w & ~a.w;
        *reinterpret_cast<uint4*>(live + i0) = make_uint4(w[t][0], w[t][1], w[t][2], w[t][3]);
      } else {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
          const bool in = i0 + k < words;
          sc[k] = in ? bwv[i0 + k] : 0u;
          w[t][k] = in ? live[i0 + k] & ~sc[k] : 0u;
          if (in) live[i0 + k] = w[t][k];
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        count += __popc(sc[k]);
        if (sc[k]) mn = min(mn, (uint32_t)((i0 + k) * 32 + __ffs(sc[k]) - 1));
      }
    }
    if (list) chunk_to_list(w, base, list, cursor, s_wave, &s_base);
  }
  count = block_sum(count, s_red);
  mn = block_min(mn, s_red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = count;
    part[2 * blockIdx.x + 1] = mn;
  }
}

// Labels of the pivot's SCC (the bits of bm): every workgroup reduces
// k_fwbw_finish's nparts partials (at most WG) to the SCC's smallest vertex
// m and its size, which block 0 writes to *minv_out and *count_out; then
// labels[v] = m for the SCC. With `live` (the live bitmap the SCC has left,
// after a first trim that wrote no labels) every label is written, in full
// lines: m in the SCC, UNDEC for live vertices, and their own id for the
// vertices the first trim decided (neither live nor in the SCC).
__device__ void label_bits(uint32_t bid, uint32_t nb, uint32_t n, const uint32_t* __restrict__ bm,
                                                   const uint64_t* __restrict__ part, uint32_t nparts,
                                                   uint32_t* minv_out, uint32_t* count_out,
                                                   uint32_t* __restrict__ labels,
                                                   const uint32_t* __restrict__ live) {
  __shared__ uint32_t s_red[WG / WAVE];
  __shared__ uint32_t s_m;
  {
    const bool have = threadIdx.x < nparts;
    uint32_t c = have ? (uint32_t)part[2 * threadIdx.x] : 0u;
    uint32_t x = have ? (uint32_t)part[2 * threadIdx.x + 1] : 0xFFFFFFFFu;
    c = block_sum(c, s_red);
    x = block_min(x, s_red);
    if (threadIdx.x == 0) {
      s_m = x;
      if (bid == 0) {
        *minv_out = x;
        *count_out = c;
      }
    }
    __syncthreads();
  }
  const uint32_t m = s_m;
  if (live && ((uintptr_t)labels & 15) == 0) {
    // dense pass into 16-byte aligned labels: four consecutive vertices per
    // thread and store (one bitmap word holds them), four stores in flight
    constexpr int LQ = 4;
    const uint64_t quads = (uint64_t)n / 4;
    uint4* lab4 = reinterpret_cast<uint4*>(labels);
    for (uint64_t base = bid * (uint64_t)WG * LQ; base < quads; base += (uint64_t)nb * WG * LQ) {
      uint32_t bw[LQ], lw[LQ];
#pragma unroll
      for (int k = 0; k < LQ; ++k) {
        const uint64_t qd = base + (uint64_t)k * WG + threadIdx.x;
        bw[k] = qd < quads ? bm[qd >> 3] : 0u;
        lw[k] = qd < quads ? live[qd >> 3] : 0u;
      }
#pragma unroll
      for (int k = 0; k < LQ; ++k) {
        const uint64_t qd = base + (uint64_t)k * WG + threadIdx.x;
        if (qd >= quads) break;
        const uint32_t sh = (uint32_t)(qd & 7) * 4;
        uint32_t o[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          const uint32_t v = (uint32_t)(qd * 4 + j);
          o[j] = ((bw[k] >> (sh + j)) & 1u) ? m : (((lw[k] >> (sh + j)) & 1u) ? UNDEC : v);
        }
        lab4[qd] = make_uint4(o[0], o[1], o[2], o[3]);
      }
    }
    // the last n % 4 vertices
    if (bid == 0 && threadIdx.x < (n & 3u)) {
      const uint32_t v = (uint32_t)(quads * 4 + threadIdx.x);
      labels[v] = bm_get(bm, v) ? m : (bm_get(live, v) ? UNDEC : v);
    }
    return;
  }
  constexpr int LU = 8;  // vertices per thread and step (one per 256-vertex tile), word loads issued together
  for (uint64_t base = bid * (uint64_t)WG * LU; base < n; base += (uint64_t)nb * WG * LU) {
    uint32_t bw[LU], lw[LU];
#pragma unroll
    for (int k = 0; k < LU; ++k) {
      const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
      bw[k] = v < n ? bm[v >> 5] : 0u;
      lw[k] = (live && v < n) ? live[v >> 5] : 0u;
    }
#pragma unroll
    for (int k = 0; k < LU; ++k) {
      const uint64_t v = base + (uint64_t)k * WG + threadIdx.x;
      if (v >= n) break;
      const bool in = (bw[k] >> (v & 31)) & 1u;
      if (live)
        labels[v] = in ? m : (((lw[k] >> (v & 31)) & 1u) ? UNDEC : (uint32_t)v);
      else if (in)
        labels[v] = m;
    }
  }
}

__global__ __launch_bounds__(WG) void k_label_bits(uint32_t n, const uint32_t* __restrict__ bm,
                                                   const uint64_t* __restrict__ part, uint32_t nparts,
                                                   uint32_t* minv_out, uint32_t* count_out,
                                                   uint32_t* __restrict__ labels,
                                                   const uint32_t* __restrict__ live = nullptr) {
  label_bits(blockIdx.x, gridDim.x, n, bm, part, nparts, minv_out, count_out, labels, live);
}

// =====================================================================
// trimming
// =====================================================================

enum TrimMode : int { TRIM_ONE = 0, TRIM_NORMAL = 1 };

// TRIM_ONE: decide v when it has no live in-neighbour or no live
// out-neighbour (trim_1_exhaustive's degree-0 test). TRIM_NORMAL: decide v
// when it has neither (trim_1_normal's has_degree test). Decided vertices
// get their own id and go to `out`. Members of Q first drop their inq bit.
template <int MODE>
__global__ __launch_bounds__(WG) void k_trim_sweep(const uint32_t* __restrict__ Q, const uint32_t* qcount,
                                                   uint32_t* live, uint32_t* inq, const uint64_t* __restrict__ fw_off,
                                                   const uint32_t* __restrict__ fw_adj, const uint64_t* __restrict__ bw_off,
                                                   const uint32_t* __restrict__ bw_adj, uint32_t* __restrict__ labels,
                                                   uint32_t* __restrict__ out, uint32_t* out_cnt) {
  __shared__ uint32_t s_out[8 * WG];
  __shared__ uint32_t s_on, s_ob;
  BlockQueue q{s_out, &s_on, &s_ob, 8 * WG};
  uint32_t iter = 0;
  bq_init(q);
  const uint32_t qn = *qcount;
  const uint32_t q_pad = (qn + WG - 1) / WG * WG;
  for (uint32_t base = blockIdx.x * WG; base < q_pad; base += gridDim.x * WG) {
    uint32_t i = base + threadIdx.x;
    bool dec = false;
    uint32_t v = 0;
    if (i < qn) {
      v = Q[i];
      if (inq) atomicAnd(&inq[v >> 5], ~(1u << (v & 31)));
      if (bm_get(live, v)) {
        bool has_in = false, has_out = false;
        for (uint64_t it = bw_off[v], end = bw_off[v + 1]; it < end; ++it)
          if (bm_get(live, bw_adj[it])) { has_in = true; break; }
        if (MODE == TRIM_NORMAL ? !has_in : has_in) {
          for (uint64_t it = fw_off[v], end = fw_off[v + 1]; it < end; ++it)
            if (bm_get(live, fw_adj[it])) { has_out = true; break; }
        }
        bool trim = (MODE == TRIM_NORMAL) ? (!has_in && !has_out) : (!has_in || !has_out);
        if (trim && bm_unclaim(live, v)) {
          labels[v] = v;
          dec = true;
        }
      }
    }
    bq_push(q, dec, v);
    if ((++iter & 3) == 0) bq_maybe_flush(q, out, out_cnt, 4 * WG);  // <= 4 appends per thread since the last check
  }
  bq_flush(q, out, out_cnt);
}

// live neighbour (excluding v) count in one direction, up to 2; *only = the single one
__device__ inline int live_nbrs_upto2(uint32_t v, const uint64_t* off, const uint32_t* adj, const uint32_t* live,
                                      uint32_t* only) {
  int c = 0;
  uint32_t first = 0xFFFFFFFFu;
  for (uint64_t it = off[v], end = off[v + 1]; it < end; ++it) {
    uint32_t w = adj[it];
    if (w == v || w == first) continue;
    if (bm_get(live, w)) {
      if (c == 0) { first = w; c = 1; }
      else return 2;
    }
  }
  *only = first;
  return c;
}

// Trim-2: {v,u} is an SCC when u is v's only live neighbour and v is u's
// only live neighbour in one direction (in-edges for DIR 0, out-edges for
// DIR 1) -- nothing else live can enter (leave) the pair, and the edges
// between them make it strongly connected. The pair is claimed through the
// smaller id; one direction per launch keeps concurrent claims consistent.
template <int DIR>
__global__ __launch_bounds__(WG) void k_trim2(const uint32_t* __restrict__ Q, const uint32_t* qcount, uint32_t* live,
                                              const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                              uint32_t* __restrict__ labels, uint32_t* __restrict__ out,
                                              uint32_t* out_cnt) {
  __shared__ uint32_t s_out[8 * WG];
  __shared__ uint32_t s_on, s_ob;
  BlockQueue q{s_out, &s_on, &s_ob, 8 * WG};
  uint32_t iter = 0;
  bq_init(q);
  const uint32_t qn = *qcount;
  const uint32_t q_pad = (qn + WG - 1) / WG * WG;
  for (uint32_t base = blockIdx.x * WG; base < q_pad; base += gridDim.x * WG) {
    uint32_t i = base + threadIdx.x;
    bool dec = false;
    uint32_t lo = 0, hi = 0;
    if (i < qn) {
      uint32_t v = Q[i];
      uint32_t u = 0, back = 0;
      if (bm_get(live, v) && live_nbrs_upto2(v, off, adj, live, &u) == 1 &&
          live_nbrs_upto2(u, off, adj, live, &back) == 1 && back == v) {
        lo = min(u, v);
        hi = max(u, v);
        if (v == lo && bm_unclaim(live, lo)) {
          if (bm_unclaim(live, hi)) {
            labels[lo] = lo;
            labels[hi] = lo;
            dec = true;
          } else {
            // cannot happen when one direction runs per launch; restore
            atomicOr(&live[lo >> 5], 1u << (lo & 31));
          }
        }
      }
    }
    bq_push(q, dec, lo);
    bq_push(q, dec, hi);
    if ((++iter & 1) == 0) bq_maybe_flush(q, out, out_cnt, 4 * WG);  // <= 4 appends per thread since the last check
  }
  bq_flush(q, out, out_cnt);
}

// =====================================================================
// colour propagation
// =====================================================================

// key of v in round r: ~rotr(v, r). Taking the max of keys selects the
// same root as the reference's min of rotr(v, r); 0 marks decided vertices.
__device__ inline uint32_t color_key(uint32_t v, uint32_t r) { return ~rotr32(v, r); }

// Colour rounds keep, per vertex, one 64-bit word: the partition id in the
// high half (0 without refinement) and the colour in the low half (0 for a
// decided vertex). A neighbour's partition and colour then come from one
// random read, and atomicMax on the word raises the colour of a vertex whose
// partition never changes within a round.
__device__ inline unsigned long long color_word(uint32_t p, uint32_t c) { return ((unsigned long long)p << 32) | c; }

__global__ void k_color_init_all(uint32_t n, const uint32_t* __restrict__ live, const uint32_t* __restrict__ pid,
                                 unsigned long long* __restrict__ cp, uint32_t r) {
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG)
    cp[v] = bm_get(live, (uint32_t)v) ? color_word(pid ? pid[v] : 0u, color_key((uint32_t)v, r)) : 0ull;
}

// re-key the previous round's live list: still live -> new key (and current
// partition), decided -> 0
__global__ void k_color_init_list(const uint32_t* __restrict__ U, const uint32_t* ucount, const uint32_t* __restrict__ live,
                                  const uint32_t* __restrict__ pid, unsigned long long* __restrict__ cp, uint32_t r) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) {
    uint32_t v = U[i];
    cp[v] = bm_get(live, v) ? color_word(pid ? pid[v] : 0u, color_key(v, r)) : 0ull;
  }
}

// pull sweep: v takes the max colour over its in-neighbours of the same
// partition (decided ones hold 0); only v's thread raises cp[v]; changed
// vertices go to `out`. With `prune` (one bit per bw_adj entry), marked
// in-edges are skipped and an in-neighbour of another partition gets its
// edge marked (partitions only split, decided vertices stay decided).
constexpr int CP_BATCH = 8;  // in-neighbours a colour pull reads per step

__global__ __launch_bounds__(WG) void k_color_pull(const uint32_t* __restrict__ U, const uint32_t* ucount,
                                                   const uint64_t* __restrict__ bw_off, const uint32_t* __restrict__ bw_adj,
                                                   unsigned long long* cp, uint32_t* __restrict__ out, uint32_t* out_cnt,
                                                   uint32_t* prune) {
  __shared__ uint32_t s_out[8 * WG];
  __shared__ uint32_t s_on, s_ob;
  BlockQueue q{s_out, &s_on, &s_ob, 8 * WG};
  uint32_t iter = 0;
  bq_init(q);
  const uint32_t u = *ucount;
  const uint32_t u_pad = (u + WG - 1) / WG * WG;
  for (uint32_t base = blockIdx.x * WG; base < u_pad; base += gridDim.x * WG) {
    uint32_t i = base + threadIdx.x;
    bool changed = false;
    uint32_t v = 0;
    if (i < u) {
      v = U[i];
      const unsigned long long w0 = cp[v];
      const uint32_t c0 = (uint32_t)w0, pv = (uint32_t)(w0 >> 32);
      uint32_t c = c0;
      // CP_BATCH in-neighbours per step: their ids, then their words, each
      // group of loads in flight together
      for (uint64_t it = bw_off[v], end = bw_off[v + 1]; it < end; it += CP_BATCH) {
        uint32_t w[CP_BATCH];
        unsigned long long x[CP_BATCH];
        bool in[CP_BATCH];
        uint32_t pw[CP_BATCH];
#pragma unroll
        for (int k = 0; k < CP_BATCH; ++k) {
          in[k] = it + k < end;
          pw[k] = prune && in[k] ? prune[(it + k) >> 5] : 0u;
          w[k] = in[k] ? bw_adj[it + k] : v;
        }
#pragma unroll
        for (int k = 0; k < CP_BATCH; ++k) {
          in[k] = in[k] && !((pw[k] >> ((it + k) & 31)) & 1u);
          x[k] = in[k] ? cp[w[k]] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < CP_BATCH; ++k) {
          if ((uint32_t)(x[k] >> 32) == pv) c = max(c, (uint32_t)x[k]);
          else if (prune && in[k]) atomicOr(&prune[(it + k) >> 5], 1u << ((it + k) & 31));
        }
      }
      if (c > c0 && c0 != 0u) {
        atomicMax(&cp[v], color_word(pv, c));
        changed = true;
      }
    }
    bq_push(q, changed, v);
    if ((++iter & 3) == 0) bq_maybe_flush(q, out, out_cnt, 4 * WG);  // <= 4 appends per thread since the last check
  }
  bq_flush(q, out, out_cnt);
}

// partition refinement after a colour round: the vertices of one SCC share
// their ancestors and hence their colour, so (partition, colour) refines the
// partition; the colour's root id names the new partition
__global__ void k_refine_pid(const uint32_t* __restrict__ U, const uint32_t* ucount, const uint32_t* __restrict__ live,
                             const unsigned long long* __restrict__ cp, uint32_t r, uint32_t* __restrict__ pid) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) {
    uint32_t v = U[i];
    if (bm_get(live, v)) pid[v] = rotl32(~(uint32_t)cp[v], r);
  }
}

// roots: live v whose colour is its own key; decided with id v
__global__ __launch_bounds__(WG) void k_color_roots(const uint32_t* __restrict__ U, const uint32_t* ucount, uint32_t* live,
                                                    const unsigned long long* __restrict__ cp, uint32_t* __restrict__ labels,
                                                    uint32_t r, uint32_t* __restrict__ out, uint32_t* out_cnt) {
  __shared__ uint32_t s_out[8 * WG];
  __shared__ uint32_t s_on, s_ob;
  BlockQueue q{s_out, &s_on, &s_ob, 8 * WG};
  uint32_t iter = 0;
  bq_init(q);
  const uint32_t u = *ucount;
  const uint32_t u_pad = (u + WG - 1) / WG * WG;
  for (uint32_t base = blockIdx.x * WG; base < u_pad; base += gridDim.x * WG) {
    uint32_t i = base + threadIdx.x;
    bool root = false;
    uint32_t v = 0;
    if (i < u) {
      v = U[i];
      if ((uint32_t)cp[v] == color_key(v, r) && bm_unclaim(live, v)) {
        labels[v] = v;
        root = true;
      }
    }
    bq_push(q, root, v);
    if ((++iter & 3) == 0) bq_maybe_flush(q, out, out_cnt, 4 * WG);  // <= 4 appends per thread since the last check
  }
  bq_flush(q, out, out_cnt);
}

// =====================================================================
// canonical labels
// =====================================================================

// atomicMin(minid[rep], v) with lanes grouped by rep: the wavefront peels
// one representative per pass (leader = first pending lane), reduces the
// group's minimum across lanes and issues one atomic per group. Large
// components therefore cost one atomic per wavefront, not one per vertex.
// All lanes of the wavefront call it (uniform control flow).
// At most two peeling passes: groups of at least 8 lanes are common inside
// large components; the remaining lanes issue their own atomics.
__device__ inline void grouped_atomic_min(bool valid, uint32_t rep, uint32_t v, uint32_t* minid) {
  uint64_t pending = __ballot(valid);
  const uint32_t lane = lane_id();
  for (int pass = 0; pass < 2 && pending; ++pass) {
    uint32_t leader = __ffsll((unsigned long long)pending) - 1;
    uint32_t r = __shfl(rep, leader);
    bool mine = valid && rep == r;
    uint64_t grp = __ballot(mine);
    if (__popcll(grp) < 8) break;
    uint32_t m = wave_min(mine ? v : 0xFFFFFFFFu);
    if (lane == 0) atomicMin(&minid[r], m);
    pending &= ~grp;
    if (mine) valid = false;
  }
  if (valid && ((pending >> lane) & 1ull)) atomicMin(&minid[rep], v);
}

// minid[rep] = min { v : labels[v] == rep } over v in [0,n) or over list[0..count)
__global__ __launch_bounds__(WG) void k_canon_min(uint32_t n, const uint32_t* __restrict__ list, uint32_t count,
                                                  const uint32_t* __restrict__ labels, uint32_t* minid) {
  const uint64_t items = list ? count : n;
  const uint64_t pad = (items + WG - 1) / WG * WG;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < pad; base += (uint64_t)gridDim.x * WG) {
    uint64_t i = base + threadIdx.x;
    bool valid = i < items;
    uint32_t v = valid ? (list ? list[i] : (uint32_t)i) : 0xFFFFFFFFu;
    uint32_t rep = valid ? labels[v] : 0xFFFFFFFFu;
    if (rep >= n) valid = false;  // not a vertex id: left as is
    grouped_atomic_min(valid, rep, v, minid);
  }
}

// labels[v] <- minid[labels[v]] over [0,n) or over list[0..count)
__global__ void k_canon_apply(uint32_t n, const uint32_t* __restrict__ list, uint32_t count, uint32_t* __restrict__ labels,
                              const uint32_t* __restrict__ minid) {
  const uint64_t items = list ? count : n;
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < items; i += (uint64_t)gridDim.x * WG) {
    uint32_t v = list ? list[i] : (uint32_t)i;
    uint32_t rep = labels[v];
    if (rep >= n) continue;
    labels[v] = minid[rep];
  }
}

// minid[list[i]] = 0xFFFFFFFF for i < count (the representatives)
__global__ void k_fill_list(const uint32_t* __restrict__ list, uint32_t count, uint32_t* __restrict__ minid) {
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < count; i += gridDim.x * WG) minid[list[i]] = 0xFFFFFFFFu;
}

// =====================================================================
// residual: induced subgraph + Tarjan on one wavefront
// =====================================================================

__global__ void k_sub_ids(const uint32_t* __restrict__ U, const uint32_t* ucount, uint32_t* __restrict__ subid) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) subid[U[i]] = i;
}

// out-degree inside the subgraph (mask given as a bitmap or a byte array)
__global__ void k_sub_deg(const uint32_t* __restrict__ U, const uint32_t* ucount, const uint64_t* __restrict__ off,
                          const uint32_t* __restrict__ adj, const uint32_t* __restrict__ live_bm,
                          const uint8_t* __restrict__ mask8, uint64_t* __restrict__ deg) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i <= u; i += gridDim.x * WG) {
    if (i == u) { deg[i] = 0; continue; }
    uint32_t v = U[i];
    uint64_t d = 0;
    for (uint64_t it = off[v], end = off[v + 1]; it < end; ++it) {
      uint32_t w = adj[it];
      d += live_bm ? bm_get(live_bm, w) : (mask8[w] != 0);
    }
    deg[i] = d;
  }
}

__global__ void k_sub_fill(const uint32_t* __restrict__ U, const uint32_t* ucount, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ adj, const uint32_t* __restrict__ live_bm,
                           const uint8_t* __restrict__ mask8, const uint32_t* __restrict__ subid,
                           const uint64_t* __restrict__ sub_off, uint32_t* __restrict__ sub_adj) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) {
    uint32_t v = U[i];
    uint64_t p = sub_off[i];
    for (uint64_t it = off[v], end = off[v + 1]; it < end; ++it) {
      uint32_t w = adj[it];
      bool in = live_bm ? bm_get(live_bm, w) : (mask8[w] != 0);
      if (in) sub_adj[p++] = subid[w];
    }
  }
}

// Iterative Tarjan (scc/tarjan.hpp:20-122) by lane 0 of one wavefront over
// a device CSR; work = 6n words: index, low, stack, frame vertex, frame
// iterator (u64). labels = minimum member id (callers take min_element).
__global__ void k_tarjan(uint32_t n, const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                         uint32_t* __restrict__ work, uint32_t* __restrict__ labels) {
  if (threadIdx.x != 0) return;
  uint32_t* index = work;
  uint32_t* low = work + n;
  uint32_t* st = work + 2 * (uint64_t)n;
  uint32_t* fu = work + 3 * (uint64_t)n;
  uint64_t* fit = reinterpret_cast<uint64_t*>(work + ((4 * (uint64_t)n + 1) & ~1ull));  // 8-byte aligned
  const uint32_t ON = 0x80000000u;  // low[] high bit: vertex is on the stack
  for (uint32_t i = 0; i < n; ++i) index[i] = 0xFFFFFFFFu;
  uint32_t count = 0, sp = 0, fp = 0;
  for (uint32_t root = 0; root < n; ++root) {
    if (index[root] != 0xFFFFFFFFu) continue;
    index[root] = count;
    low[root] = count | ON;
    ++count;
    st[sp++] = root;
    fu[fp] = root;
    fit[fp] = off[root];
    ++fp;
    while (fp) {
      uint32_t u = fu[fp - 1];
      uint64_t it = fit[fp - 1];
      if (it < off[u + 1]) {
        uint32_t v = adj[it];
        fit[fp - 1] = it + 1;
        if (index[v] == 0xFFFFFFFFu) {
          index[v] = count;
          low[v] = count | ON;
          ++count;
          st[sp++] = v;
          fu[fp] = v;
          fit[fp] = off[v];
          ++fp;
        } else if ((low[v] & ON) && index[v] < (low[u] & ~ON)) {
          low[u] = index[v] | ON;
        }
        continue;
      }
      uint32_t lu = low[u] & ~ON;
      if (lu == index[u]) {
        uint32_t end = sp, beg;
        do {
          --sp;
          low[st[sp]] &= ~ON;
        } while (st[sp] != u);
        beg = sp;
        uint32_t mn = 0xFFFFFFFFu;
        for (uint32_t i = beg; i < end; ++i) mn = min(mn, st[i]);
        for (uint32_t i = beg; i < end; ++i) labels[st[i]] = mn;
      }
      --fp;
      if (fp) {
        uint32_t p = fu[fp - 1];
        uint32_t lp = low[p];
        if (lu < (lp & ~ON)) low[p] = lu | (lp & ON);
      }
    }
  }
}

// Tail solver: colour rounds on a compacted subgraph inside one workgroup
// of 1024 threads (block barriers, no host round trips). Vertex i of the
// subgraph is super[i] (any order; subid maps back); fo/fa and bo/ba are its out- and
// in-edge CSRs. Round r keys live vertices with ~rotr(super[i], r), sweeps
// max-colours along in-edges to a fixpoint, decides the roots and closes
// backward over out-edges within each colour, then writes the minimum
// member as the label of every vertex decided in the round. state: 0 live,
// 2 decided this round, 1 decided. Rounds are capped by max_rounds (0: none);
// *status = 1 if the cap was hit. A subgraph of at most TAIL_LDS_V vertices
// and TAIL_LDS_E edges per direction runs from LDS (its CSRs as u32, the
// colours, states and minima), the rest from the global buffers.
constexpr uint32_t TAIL_LDS_V = 4096;  // every tail (at most 4095 vertices)
constexpr uint32_t TAIL_LDS_E = 7168;  // edges per direction staged in LDS

// the tail solver's rounds over CSRs fo/fa (out) and bo/ba (in), with the
// per-vertex colours, states and minima at col/state/minv (LDS or global)
template <typename Off, typename St>
__device__ void tail_rounds(uint32_t sn, const uint32_t* __restrict__ super, const uint32_t* __restrict__ subid,
                            const Off* fo, const Off* bo, const uint32_t* fa, const uint32_t* ba, uint32_t* col,
                            St* state, uint32_t* minv, uint32_t* __restrict__ labels, uint32_t r0,
                            uint32_t max_rounds, uint32_t* status, uint32_t* live, uint32_t* s_flag,
                            uint32_t* s_live) {
  const uint32_t tid = threadIdx.x, T = blockDim.x;
  for (uint32_t i = tid; i < sn; i += T) state[i] = 0;
  uint32_t r = r0;
  for (uint32_t round = 0;; ++round) {
    __syncthreads();
    if (tid == 0) *s_live = 0;
    __syncthreads();
    for (uint32_t i = tid; i < sn; i += T)
      if (state[i] == 0) { *s_live = 1; break; }
    __syncthreads();
    if (!*s_live) break;
    if (max_rounds && round >= max_rounds) {
      if (tid == 0) *status = 1;
      break;
    }
    ++r;
    for (uint32_t i = tid; i < sn; i += T) col[i] = state[i] == 0 ? ~rotr32(super[i], r) : 0u;
    // forward: max colour over in-neighbours, in place, to a fixpoint
    for (;;) {
      __syncthreads();
      if (tid == 0) *s_flag = 0;
      __syncthreads();
      for (uint32_t i = tid; i < sn; i += T) {
        if (state[i] != 0) continue;
        uint32_t c0 = col[i], c = c0;
        for (uint64_t it = bo[i], end = bo[i + 1]; it < end; ++it) c = max(c, col[ba[it]]);
        if (c > c0) {
          col[i] = c;
          *s_flag = 1;
        }
      }
      __syncthreads();
      if (!*s_flag) break;
    }
    // roots, then backward closure within a colour over out-edges
    for (uint32_t i = tid; i < sn; i += T)
      if (state[i] == 0 && col[i] == ~rotr32(super[i], r)) {
        state[i] = 2;
        minv[i] = super[i];
      }
    for (;;) {
      __syncthreads();
      if (tid == 0) *s_flag = 0;
      __syncthreads();
      for (uint32_t i = tid; i < sn; i += T) {
        if (state[i] != 0) continue;
        uint32_t ci = col[i];
        for (uint64_t it = fo[i], end = fo[i + 1]; it < end; ++it) {
          uint32_t j = fa[it];
          if (state[j] == 2 && col[j] == ci) {
            state[i] = 2;
            *s_flag = 1;
            break;
          }
        }
      }
      __syncthreads();
      if (!*s_flag) break;
    }
    // minimum member per colour: the root of colour c is the vertex whose
    // key is c, at compact index subid[root]
    for (uint32_t i = tid; i < sn; i += T) {
      if (state[i] != 2) continue;
      atomicMin(&minv[subid[rotl32(~col[i], r)]], super[i]);
    }
    __syncthreads();
    for (uint32_t i = tid; i < sn; i += T) {
      if (state[i] != 2) continue;
      labels[super[i]] = minv[subid[rotl32(~col[i], r)]];
    }
    __syncthreads();
    for (uint32_t i = tid; i < sn; i += T)
      if (state[i] == 2) state[i] = 1;
  }
  // every decided vertex leaves the live set
  __syncthreads();
  for (uint32_t i = tid; i < sn; i += T)
    if (state[i] == 1) atomicAnd(&live[super[i] >> 5], ~(1u << (super[i] & 31)));
}

__global__ __launch_bounds__(1024) void k_tail_solve(const uint32_t* sn_ptr, const uint32_t* __restrict__ super,
                                                     const uint32_t* __restrict__ subid,
                                                     const uint64_t* __restrict__ fo, const uint64_t* __restrict__ bo,
                                                     const uint32_t* __restrict__ sub, uint64_t cap,
                                                     uint32_t* col, uint32_t* state, uint32_t* minv,
                                                     uint32_t* __restrict__ labels, uint32_t r0, uint32_t max_rounds,
                                                     uint32_t* status, uint32_t* live, const uint32_t* skip,
                                                     const uint32_t* room) {
  __shared__ uint32_t s_flag, s_live;
  const uint32_t tid = threadIdx.x, T = blockDim.x;
  // the sizes and the sub-CSR come from the launches before (k_tail_fill)
  const uint32_t sn = *sn_ptr;
  if (*skip || *room) return;
  const uint64_t sm = fo[sn];
  if (2 * sm + 2 > cap) return;
  const uint32_t* __restrict__ fa = sub;
  const uint32_t* __restrict__ ba = sub + sm + 1;
  if (sn <= TAIL_LDS_V && sm <= TAIL_LDS_E) {
    __shared__ uint32_t s_fo[TAIL_LDS_V + 1], s_bo[TAIL_LDS_V + 1];
    __shared__ uint32_t s_fa[TAIL_LDS_E], s_ba[TAIL_LDS_E];
    __shared__ uint32_t s_col[TAIL_LDS_V], s_minv[TAIL_LDS_V];
    __shared__ uint8_t s_state[TAIL_LDS_V];
    for (uint32_t i = tid; i <= sn; i += T) {
      s_fo[i] = (uint32_t)fo[i];
      s_bo[i] = (uint32_t)bo[i];
    }
    for (uint32_t e = tid; e < (uint32_t)sm; e += T) {
      s_fa[e] = fa[e];
      s_ba[e] = ba[e];
    }
    // the rounds start with a block barrier
    tail_rounds(sn, super, subid, s_fo, s_bo, s_fa, s_ba, s_col, s_state, s_minv, labels, r0, max_rounds, status,
                live, &s_flag, &s_live);
  } else {
    tail_rounds(sn, super, subid, fo, bo, fa, ba, col, state, minv, labels, r0, max_rounds, status, live, &s_flag,
                &s_live);
  }
}

// tail subgraph, one pass: compact ids, live out- and in-degrees (entry u = 0)
// Nothing is done when there are more than `limit` live vertices (the
// tail scan reports that case). *iso counts live vertices without a live
// neighbour in either direction: trim_1_normal's decisions. One wavefront
// per vertex: tail vertices can have long rows.
__device__ void tail_prep(uint32_t bid, uint32_t nb, uint32_t n, const uint32_t* __restrict__ U,
                          const uint32_t* ucount, uint32_t limit, const uint64_t* __restrict__ fw_off,
                          const uint32_t* __restrict__ fw_adj, const uint64_t* __restrict__ bw_off,
                          const uint32_t* __restrict__ bw_adj, const uint32_t* __restrict__ live,
                          uint32_t* __restrict__ subid, uint64_t* __restrict__ degf, uint64_t* __restrict__ degb,
                          uint32_t* iso, uint32_t* bad) {
  __shared__ uint32_t s_iso;
  const uint32_t u = *ucount;
  if (u > limit) return;
  const uint32_t lane = lane_id();
  const uint32_t wpb = WG / WAVE;
  if (threadIdx.x == 0) s_iso = 0;
  __syncthreads();
  if (bid == 0 && threadIdx.x == 0) {
    degf[u] = 0;
    degb[u] = 0;
  }
  for (uint32_t i = bid * wpb + threadIdx.x / WAVE; i < u; i += nb * wpb) {
    const uint32_t v = U[i];
    if (v >= n) {  // not a vertex: reported, never used as an index
      if (lane == 0) *bad = 1;
      continue;
    }
    uint32_t df = 0, db = 0;
    for (uint64_t it = fw_off[v] + lane, end = fw_off[v + 1]; it < end; it += WAVE) df += bm_get(live, fw_adj[it]);
    for (uint64_t it = bw_off[v] + lane, end = bw_off[v + 1]; it < end; it += WAVE) db += bm_get(live, bw_adj[it]);
    df = wave_sum(df);
    db = wave_sum(db);
    if (lane == 0) {
      subid[v] = i;
      degf[i] = df;
      degb[i] = db;
      if (df == 0 && db == 0) atomicAdd(&s_iso, 1u);
    }
  }
  __syncthreads();
  if (iso && threadIdx.x == 0 && s_iso) atomicAdd(iso, s_iso);  // one global add per workgroup
}

__global__ __launch_bounds__(WG) void k_tail_prep(uint32_t n, const uint32_t* __restrict__ U, const uint32_t* ucount,
                                                  uint32_t limit, const uint64_t* __restrict__ fw_off,
                                                  const uint32_t* __restrict__ fw_adj,
                                                  const uint64_t* __restrict__ bw_off,
                                                  const uint32_t* __restrict__ bw_adj,
                                                  const uint32_t* __restrict__ live, uint32_t* __restrict__ subid,
                                                  uint64_t* __restrict__ degf, uint64_t* __restrict__ degb,
                                                  uint32_t* iso, uint32_t* bad) {
  tail_prep(blockIdx.x, gridDim.x, n, U, ucount, limit, fw_off, fw_adj, bw_off, bw_adj, live, subid, degf, degb, iso,
            bad);
}

// The label pass of the last FW-BW (label_bits, blocks [0, lblocks); its
// bitmap bm, partials, outputs and lab_live as for k_label_bits) and the
// tail's degree pass (tail_prep, the other blocks; nothing counted as
// isolated) in one launch. Both need only the tail split before them: the
// split labels the vertices it decides with their own id, which is what
// the label pass writes for vertices neither live nor in the SCC.
__global__ __launch_bounds__(WG) void k_label_prep(uint32_t lblocks, uint32_t n, const uint32_t* __restrict__ bm,
                                                   const uint64_t* __restrict__ part, uint32_t nparts,
                                                   uint32_t* minv_out, uint32_t* count_out,
                                                   uint32_t* __restrict__ labels,
                                                   const uint32_t* __restrict__ lab_live,
                                                   const uint32_t* __restrict__ U, const uint32_t* ucount,
                                                   uint32_t limit, const uint64_t* __restrict__ fw_off,
                                                   const uint32_t* __restrict__ fw_adj,
                                                   const uint64_t* __restrict__ bw_off,
                                                   const uint32_t* __restrict__ bw_adj,
                                                   const uint32_t* __restrict__ live, uint32_t* __restrict__ subid,
                                                   uint64_t* __restrict__ degf, uint64_t* __restrict__ degb,
                                                   uint32_t* bad) {
  if (blockIdx.x < lblocks)
    label_bits(blockIdx.x, lblocks, n, bm, part, nparts, minv_out, count_out, labels, lab_live);
  else
    tail_prep(blockIdx.x - lblocks, gridDim.x - lblocks, n, U, ucount, limit, fw_off, fw_adj, bw_off, bw_adj, live,
              subid, degf, degb, nullptr, bad);
}

// trim_1_normal over a list of live vertices, for the tail: a vertex with
// no live neighbour in either direction is decided (own id, live bit
// cleared); the others go to `rest`, staged per workgroup in LDS, so the
// decided ones number *lcount - *rest_cnt (not counted here: a count per
// workgroup on one address costs more than the scans). Removing
// such a vertex changes no other vertex's test: it is nobody's live
// neighbour. A group of 16 lanes per vertex (four per wavefront) reads 16
// entries of its out-row and of its in-row per step, stopping at the first
// live neighbour.
__global__ __launch_bounds__(WG) void k_tail_split(uint32_t n, const uint32_t* __restrict__ L, const uint32_t* lcount,
                                                   const uint64_t* __restrict__ fw_off,
                                                   const uint32_t* __restrict__ fw_adj,
                                                   const uint64_t* __restrict__ bw_off,
                                                   const uint32_t* __restrict__ bw_adj, uint32_t* live,
                                                   uint32_t* __restrict__ labels, uint32_t* __restrict__ rest,
                                                   uint32_t* rest_cnt, uint32_t* bad) {
  constexpr uint32_t GL = 16;              // lanes per vertex
  constexpr uint32_t VPW = WAVE / GL;      // vertices per wavefront
  constexpr uint32_t VPB = VPW * (WG / WAVE);  // vertices per workgroup and step
  __shared__ uint32_t s_q[2 * VPB];
  __shared__ uint32_t s_n, s_b;
  BlockQueue q{s_q, &s_n, &s_b, 2 * VPB};
  bq_init(q);
  const uint32_t lane = lane_id();
  const uint32_t gl = lane % GL;
  const uint64_t gmask = ((1ull << GL) - 1) << (lane - gl);
  const uint32_t cnt = *lcount;
  const uint32_t cnt_pad = (cnt + VPB - 1) / VPB * VPB;
  for (uint32_t base = blockIdx.x * VPB; base < cnt_pad; base += gridDim.x * VPB) {
    const uint32_t i = base + threadIdx.x / GL;
    const uint32_t li = i < cnt ? L[i] : 0u;
    if (i < cnt && li >= n && gl == 0) *bad = 1;  // not a vertex: reported, never used as an index
    const bool valid = i < cnt && li < n;
    const uint32_t v = valid ? li : 0u;
    uint64_t fb = 0, fe = 0, bb = 0, be = 0;
    if (valid) {
      fb = fw_off[v];
      fe = fw_off[v + 1];
      bb = bw_off[v];
      be = bw_off[v + 1];
    }
    bool found = false;
    for (uint64_t c = gl;; c += GL) {
      const bool act = valid && !found && (fb + c - gl < fe || bb + c - gl < be);
      if (!__any(act)) break;
      bool hit = false;
      if (act) {
        const bool fx = fb + c < fe, bx = bb + c < be;
        const uint32_t x = fx ? fw_adj[fb + c] : 0u;
        const uint32_t y = bx ? bw_adj[bb + c] : 0u;
        hit = (fx && bm_get(live, x)) || (bx && bm_get(live, y));
      }
      found = found || (__ballot(hit) & gmask) != 0;
    }
    const bool keep = gl == 0 && valid && found;
    if (gl == 0 && valid && !found) {
      labels[v] = v;
      atomicAnd(&live[v >> 5], ~(1u << (v & 31)));
    }
    bq_push(q, keep, v);
    bq_maybe_flush(q, rest, rest_cnt, VPB);  // at most one append per vertex of the step
  }
  bq_flush(q, rest, rest_cnt);
}

// in-place exclusive scans of two arrays of u+1 <= 4096 entries in one
// workgroup (entry u becomes the total); *flag = 0 for the tail solver. A
// u (*ucount) above `limit` (at most 4095) is reported through *err and
// nothing is written.
__global__ __launch_bounds__(1024) void k_tail_scan(const uint32_t* ucount, uint32_t limit, uint64_t* __restrict__ a,
                                                    uint64_t* __restrict__ b, uint32_t* flag, uint32_t* err) {
  __shared__ uint64_t sa[1024], sb[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t u = *ucount;
  if (t == 0) *flag = 0;
  if (u > limit || u >= 4096) {
    if (t == 0) *err = 1;
    return;
  }
  uint64_t da[4], db[4], suma = 0, sumb = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = 4 * t + k;
    da[k] = i <= u ? a[i] : 0;
    db[k] = i <= u ? b[i] : 0;
    suma += da[k];
    sumb += db[k];
  }
  sa[t] = suma;
  sb[t] = sumb;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    const uint64_t xa = t >= o ? sa[t - o] : 0, xb = t >= o ? sb[t - o] : 0;
    __syncthreads();
    sa[t] += xa;
    sb[t] += xb;
    __syncthreads();
  }
  uint64_t ra = sa[t] - suma, rb = sb[t] - sumb;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = 4 * t + k;
    if (i <= u) {
      a[i] = ra;
      b[i] = rb;
    }
    ra += da[k];
    rb += db[k];
  }
}

// both directions' rows of the tail subgraph in one launch
// The sub-CSR's two adjacency arrays share `sub` (cap entries): forward
// at 0, backward after the forward total fo[u] and one spare entry. Nothing
// is done when *skip is set; when they do not fit, *room is set instead.
__global__ __launch_bounds__(WG) void k_tail_fill(uint32_t n, const uint32_t* __restrict__ U, const uint32_t* ucount,
                                                  const uint64_t* __restrict__ fw_off,
                                                  const uint32_t* __restrict__ fw_adj,
                                                  const uint64_t* __restrict__ bw_off,
                                                  const uint32_t* __restrict__ bw_adj,
                                                  const uint32_t* __restrict__ live,
                                                  const uint32_t* __restrict__ subid, const uint64_t* __restrict__ fo,
                                                  const uint64_t* __restrict__ bo, uint32_t* __restrict__ sub,
                                                  uint64_t cap, const uint32_t* skip, uint32_t* room, uint32_t* bad) {
  const uint32_t u = *ucount;
  if (*skip) return;
  const uint64_t sm = fo[u];
  if (2 * sm + 2 > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *room = 1;
    return;
  }
  uint32_t* __restrict__ fa = sub;
  uint32_t* __restrict__ ba = sub + sm + 1;
  const uint32_t lane = lane_id();
  const uint32_t wpb = WG / WAVE;
  const uint64_t below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  // one wavefront per vertex; live neighbours in row order
  for (uint32_t i = blockIdx.x * wpb + threadIdx.x / WAVE; i < u; i += gridDim.x * wpb) {
    const uint32_t v = U[i];
    if (v >= n) {  // not a vertex: reported, never used as an index
      if (lane == 0) *bad = 1;
      continue;
    }
    uint64_t p = fo[i];
    for (uint64_t b = fw_off[v], end = fw_off[v + 1]; b < end; b += WAVE) {
      const uint64_t it = b + lane;
      const uint32_t w = it < end ? fw_adj[it] : 0u;
      const bool keep = it < end && bm_get(live, w);
      const uint64_t m = __ballot(keep);
      if (keep) fa[p + __popcll(m & below)] = subid[w];
      p += __popcll(m);
    }
    p = bo[i];
    for (uint64_t b = bw_off[v], end = bw_off[v + 1]; b < end; b += WAVE) {
      const uint64_t it = b + lane;
      const uint32_t w = it < end ? bw_adj[it] : 0u;
      const bool keep = it < end && bm_get(live, w);
      const uint64_t m = __ballot(keep);
      if (keep) ba[p + __popcll(m & below)] = subid[w];
      p += __popcll(m);
    }
  }
}

// labels[U[i]] = U[sub_label[i]]
__global__ void k_res_apply(const uint32_t* __restrict__ U, const uint32_t* ucount, const uint32_t* __restrict__ sublab,
                            uint32_t* __restrict__ labels) {
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) labels[U[i]] = U[sublab[i]];
}

// pivot among live vertices: max in*out, smallest id on ties. LIVE = false:
// original degrees (select_pivot_from_head, scc/pivot.hpp:28-47); LIVE =
// true: live neighbours only (select_pivot_from_degree, scc/pivot.hpp:53-74,
// whose degree arrays count exactly the live neighbours after an exhaustive
// trim). Restricted to one rank.
template <bool LIVE>
__global__ __launch_bounds__(WG) void k_pivot_live(const uint32_t* __restrict__ U, const uint32_t* ucount,
                                                   const uint64_t* __restrict__ fw_off, const uint32_t* __restrict__ fw_adj,
                                                   const uint64_t* __restrict__ bw_off, const uint32_t* __restrict__ bw_adj,
                                                   const uint32_t* __restrict__ live, uint64_t* __restrict__ part) {
  __shared__ uint64_t s_p[WG];
  __shared__ uint32_t s_id[WG];
  uint64_t best = 0;
  uint32_t best_id = 0xFFFFFFFFu;
  uint32_t u = *ucount;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < u; i += gridDim.x * WG) {
    uint32_t v = U[i];
    uint64_t p;
    if (LIVE) {
      uint64_t dout = 0, din = 0;
      for (uint64_t it = fw_off[v], e = fw_off[v + 1]; it < e; ++it) dout += bm_get(live, fw_adj[it]);
      for (uint64_t it = bw_off[v], e = bw_off[v + 1]; it < e; ++it) din += bm_get(live, bw_adj[it]);
      p = dout * din;
    } else {
      p = (fw_off[v + 1] - fw_off[v]) * (bw_off[v + 1] - bw_off[v]);
    }
    if (best_id == 0xFFFFFFFFu || p > best || (p == best && v < best_id)) { best = p; best_id = v; }
  }
  s_p[threadIdx.x] = best;
  s_id[threadIdx.x] = best_id;
  __syncthreads();
  for (int s = WG / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      uint64_t p2 = s_p[threadIdx.x + s];
      uint32_t i2 = s_id[threadIdx.x + s];
      if (i2 != 0xFFFFFFFFu && (s_id[threadIdx.x] == 0xFFFFFFFFu || p2 > s_p[threadIdx.x] ||
                                (p2 == s_p[threadIdx.x] && i2 < s_id[threadIdx.x]))) {
        s_p[threadIdx.x] = p2;
        s_id[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s_p[0];
    part[3 * blockIdx.x + 1] = s_id[0];
    part[3 * blockIdx.x + 2] = 0;
  }
}

__global__ void k_mask_to_bitmap(uint32_t n, const uint8_t* mask, uint32_t* bm) {
  const uint64_t n_pad = ((uint64_t)n + WG - 1) / WG * WG;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < n_pad; base += (uint64_t)gridDim.x * WG) {
    uint64_t v = base + threadIdx.x;
    bool in = v < n && mask[v] != 0;
    uint64_t bits = __ballot(in);
    uint32_t lane = lane_id();
    uint64_t word0 = (base + threadIdx.x - lane) >> 5;
    if (lane == 0) bm[word0] = (uint32_t)bits;
    if (lane == 32) bm[word0 + 1] = (uint32_t)(bits >> 32);
  }
}

// =====================================================================
// partitioned multi-pivot FW-BW
// =====================================================================
// Every live vertex carries a partition id; SCCs never cross partitions.
// A round picks one pivot per partition, runs forward and backward
// reachability from all pivots at once (claims stay inside the source's
// partition), decides FW & BW (the pivots' SCCs) and splits every partition
// into FW-only, BW-only and neither (Fleischer, Hendrickson, Pinar 2000).

// half-octave of a positive product: 2 * floor(log2 p) + the bit after the leading one (0 for p = 0)
__device__ inline uint32_t prod_bucket(uint64_t p) {
  if (p == 0) return 0;
  const uint32_t b = 63 - __clzll(p);
  return 1 + 2 * b + (b ? (uint32_t)((p >> (b - 1)) & 1u) : 0u);
}

__device__ inline uint32_t float_order(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline unsigned long long wave_max64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_down(v, off);
    v = o > v ? o : v;
  }
  return v;
}

// Pivot priority of every vertex for a whole solve (the degrees and the
// MLP score do not change within it): min(in*out, 2^32-1) for the plain
// rule (smallest id on ties), else the plain rule's half-octave of in*out in
// bits 24..31, then the ordered score (learned) or nothing (random: a
// seeded hash of the round fills bits 0..23 when the key is formed), and
// bit 0 set so a vertex of product 0 still has a priority.
__global__ void k_part_prio(uint32_t n, const uint64_t* __restrict__ fw_off, const uint64_t* __restrict__ bw_off,
                            const float* __restrict__ score, bool bucketed, uint32_t* __restrict__ prio) {
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG) {
    const uint64_t prod = (fw_off[v + 1] - fw_off[v]) * (bw_off[v + 1] - bw_off[v]);
    uint32_t pri;
    if (bucketed) pri = (prod_bucket(prod) << 24) | (score ? float_order(score[v]) >> 8 : 0u) | 1u;
    else pri = prod > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)prod;
    prio[v] = pri;
  }
}

// the pivot key of vertex v of priority pri (prio[v]): (priority << 32) |
// ~v, so the largest key is the highest priority and, among equals, the
// smallest id
__device__ inline unsigned long long part_key_of(uint32_t pri, uint32_t v, uint64_t rnd_seed) {
  if (rnd_seed) {
    uint64_t z = ((uint64_t)rnd_seed << 32 | v) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ull;
    pri |= (uint32_t)(z >> 40);
  }
  return ((unsigned long long)pri << 32) | (uint32_t)~v;
}

__device__ inline unsigned long long part_key(const uint32_t* __restrict__ prio, uint32_t v, uint64_t rnd_seed) {
  return part_key_of(prio[v], v, rnd_seed);
}

// Per-workgroup aggregation of (slot, value) updates in an LDS hash table
// (linear probing over AGG_SLOTS entries); a full table falls back to the
// global atomic. Many vertices update the same slot (a partition's pivot
// key, an SCC's smallest member), so this turns one global atomic per
// vertex into one per workgroup and slot.
constexpr uint32_t AGG_SLOTS = 512;
constexpr uint32_t AGG_EMPTY = 0xFFFFFFFFu;

template <typename T>
__device__ inline void agg_init(uint32_t* key, T* val, T none) {
  for (uint32_t i = threadIdx.x; i < AGG_SLOTS; i += blockDim.x) {
    key[i] = AGG_EMPTY;
    val[i] = none;
  }
  __syncthreads();
}

// true if the update landed in the table
template <typename T, bool MAX>
__device__ inline bool agg_put(uint32_t* key, T* val, uint32_t k, T v) {
  uint32_t h = (k * 2654435761u) >> 23;  // 9 bits: AGG_SLOTS
  for (int probe = 0; probe < 8; ++probe) {
    uint32_t old = key[h];
    if (old == AGG_EMPTY) old = atomicCAS(&key[h], AGG_EMPTY, k);
    if (old == AGG_EMPTY || old == k) {
      if constexpr (MAX) atomicMax(&val[h], v);
      else atomicMin(&val[h], v);
      return true;
    }
    h = (h + 1) & (AGG_SLOTS - 1);
  }
  return false;
}

template <typename T, bool MAX>
__device__ inline void agg_flush(const uint32_t* key, const T* val, T* global) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < AGG_SLOTS; i += blockDim.x) {
    if (key[i] == AGG_EMPTY) continue;
    if constexpr (MAX) atomicMax(&global[key[i]], val[i]);
    else atomicMin(&global[key[i]], val[i]);
  }
}

// One update per lane (slot p, value x; lanes with ok false have none): the
// wavefront's largest group (the first lane's slot) makes one update with
// its extreme when it has at least 8 members; the rest go to the
// workgroup's LDS table one by one.
template <typename T, bool MAX>
__device__ inline void agg_wave_put(uint32_t* key, T* val, T* global, bool ok, uint32_t p, T x) {
  const uint64_t pending = __ballot(ok);
  if (!pending) return;
  bool left = ok;
  const uint32_t leader = __ffsll((unsigned long long)pending) - 1;
  const uint32_t r = __shfl(p, leader);
  const bool mine = ok && p == r;
  if (__popcll(__ballot(mine)) >= 8) {
    T m = mine ? x : (MAX ? (T)0 : (T)~(T)0);
    for (int o = 32; o > 0; o >>= 1) {
      const T t = __shfl_down(m, o);
      m = MAX ? (t > m ? t : m) : (t < m ? t : m);
    }
    if (lane_id() == 0 && !agg_put<T, MAX>(key, val, r, m)) {
      if constexpr (MAX) atomicMax(&global[r], m);
      else atomicMin(&global[r], m);
    }
    if (mine) left = false;
  }
  if (left && !agg_put<T, MAX>(key, val, p, x)) {
    if constexpr (MAX) atomicMax(&global[p], x);
    else atomicMin(&global[p], x);
  }
}

// pbest[p] = max over live v in p of part_key(v): the first round's keys
// (later rounds form them in k_part_split)
__global__ __launch_bounds__(WG) void k_part_key(uint32_t n, const uint32_t* __restrict__ live,
                                                 const uint32_t* __restrict__ pid, const uint32_t* __restrict__ prio,
                                                 uint64_t rnd_seed, unsigned long long* pbest) {
  __shared__ uint32_t s_key[AGG_SLOTS];
  __shared__ unsigned long long s_val[AGG_SLOTS];
  agg_init<unsigned long long>(s_key, s_val, 0ull);
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < pad; base += (uint64_t)gridDim.x * WG) {
    const uint64_t v = base + threadIdx.x;
    const bool valid = v < n && bm_get(live, (uint32_t)v);
    uint32_t p = 0;
    unsigned long long key = 0;
    if (valid) {
      p = pid[v];
      key = part_key(prio, (uint32_t)v, rnd_seed);
    }
    agg_wave_put<unsigned long long, true>(s_key, s_val, pbest, valid, p, key);
  }
  agg_flush<unsigned long long, true>(s_key, s_val, pbest);
}

// Partition ids are sparse inside a round: after a round, the part of
// partition q in class c (0 unreached, 1 forward only, 2 backward only) gets
// id 3 * newid[q] + c. The next round's pivot keys (pbest, one slot per
// sparse id) show which ids are in use; an exclusive scan of "slot used"
// gives newid, the compact index of each used partition. So no per-vertex
// pass writes per-partition flags.

// flags[q] = (pbest[q] != 0) for q < S, flags[S] = 0
__global__ void k_part_used(uint32_t S, const unsigned long long* __restrict__ pbest, uint32_t* __restrict__ flags) {
  for (uint32_t q = blockIdx.x * WG + threadIdx.x; q <= S; q += gridDim.x * WG) flags[q] = (q < S && pbest[q] != 0ull);
}

// pivot of every used partition into F[newid[q]], its forward and backward
// bits set. A key naming no vertex (id >= n) is reported through *err
// instead of being used as an index.
__global__ void k_part_emit(uint32_t S, uint32_t n, const unsigned long long* __restrict__ pbest,
                            const uint32_t* __restrict__ newid, uint32_t* __restrict__ F, uint32_t* fwv, uint32_t* bwv,
                            uint32_t* err) {
  for (uint32_t q = blockIdx.x * WG + threadIdx.x; q < S; q += gridDim.x * WG) {
    const unsigned long long k = pbest[q];
    if (k == 0ull) continue;
    const uint32_t v = ~(uint32_t)k;
    if (v >= n) {
      *err = 1u;
      continue;
    }
    F[newid[q]] = v;
    atomicOr(&fwv[v >> 5], 1u << (v & 31));
    atomicOr(&bwv[v >> 5], 1u << (v & 31));
  }
}

// after a round: pivot SCCs (forward and backward reached) contribute their
// smallest member to pmin[newid[pid]] (aggregated per workgroup in LDS);
// with pcount, every live vertex counts itself in its partition's
// (decided, forward-only, backward-only, unreached) counters. Only decided
// vertices read pid: otherwise a pass over bitmaps.
__global__ __launch_bounds__(WG) void k_part_classify(uint32_t n, const uint32_t* __restrict__ live,
                                                      const uint32_t* __restrict__ fwv, const uint32_t* __restrict__ bwv,
                                                      const uint32_t* __restrict__ pid, const uint32_t* __restrict__ newid,
                                                      uint32_t* pmin, uint32_t* pcount) {
  __shared__ uint32_t s_key[AGG_SLOTS];
  __shared__ uint32_t s_val[AGG_SLOTS];
  agg_init<uint32_t>(s_key, s_val, 0xFFFFFFFFu);
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  const uint32_t lane = lane_id();
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < pad; base += (uint64_t)gridDim.x * WG) {
    const uint64_t v = base + threadIdx.x;
    const bool is_live = v < n && bm_get(live, (uint32_t)v);
    bool dec = false;
    uint32_t k = 0;
    if (is_live) {
      const bool f = bm_get(fwv, (uint32_t)v), b = bm_get(bwv, (uint32_t)v);
      dec = f && b;
      if (dec || pcount) k = newid[pid[v]];
      if (pcount) atomicAdd(&pcount[4ull * k + (dec ? 0 : (f ? 1 : (b ? 2 : 3)))], 1u);
    }
    // the wavefront's largest group (leader's partition): one update with
    // its first (smallest) member; the other decided lanes update alone
    const uint64_t dm = __ballot(dec);
    if (dm) {
      const uint32_t leader = __ffsll((unsigned long long)dm) - 1;
      const uint32_t r = __shfl(k, leader);
      if (dec && k == r) {
        if (lane == leader && !agg_put<uint32_t, false>(s_key, s_val, r, (uint32_t)v)) atomicMin(&pmin[r], (uint32_t)v);
        dec = false;
      }
      if (dec && !agg_put<uint32_t, false>(s_key, s_val, k, (uint32_t)v)) atomicMin(&pmin[k], (uint32_t)v);
    }
  }
  agg_flush<uint32_t, false>(s_key, s_val, pmin);
}

// decided: label = pmin[newid[pid]], removed from live (each wavefront owns
// its two words); others: pid = 3 * newid[pid] + class. Per-workgroup
// decided counts.
__global__ __launch_bounds__(WG) void k_part_apply(uint32_t n, uint32_t* live, const uint32_t* __restrict__ fwv,
                                                   const uint32_t* __restrict__ bwv, uint32_t* __restrict__ pid,
                                                   const uint32_t* __restrict__ pmin, const uint32_t* __restrict__ newid,
                                                   uint32_t* __restrict__ labels, uint64_t* __restrict__ partials) {
  __shared__ uint64_t s_red[WG / WAVE];
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  const uint32_t lane = lane_id();
  uint64_t count = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < pad; base += (uint64_t)gridDim.x * WG) {
    const uint64_t v = base + threadIdx.x;
    const uint32_t lw = v < n ? live[v >> 5] : 0u;
    const bool is_live = v < n && ((lw >> (v & 31)) & 1u);
    bool dec = false;
    if (is_live) {
      const uint32_t k = newid[pid[v]];
      bool f = bm_get(fwv, (uint32_t)v), b = bm_get(bwv, (uint32_t)v);
      if (f && b) {
        dec = true;
        labels[v] = pmin[k];
        pid[v] = DEAD_PID;  // no longer live
        ++count;
      } else {
        pid[v] = 3u * k + (f ? 1u : (b ? 2u : 0u));
      }
    }
    const uint64_t mask = __ballot(dec);
    const uint64_t word0 = (base + threadIdx.x - lane) >> 5;
    if (lane == 0 && (uint32_t)mask) live[word0] = lw & ~(uint32_t)mask;
    if (lane == 32 && (uint32_t)(mask >> 32)) live[word0 + 1] = lw & ~(uint32_t)(mask >> 32);
  }
  count = block_sum64(count, s_red);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = count;
    partials[3 * blockIdx.x + 1] = 0;
    partials[3 * blockIdx.x + 2] = 0;
  }
}

// After a round (without outcome records): classification and
// relabelling in one pass over the vertices. A live vertex reached both
// forward and backward is decided: it leaves live (each wavefront owns its
// two words), its id goes to pmin[newid[pid]] (minimum, aggregated) and it
// is appended to dlist (count *dcount). Any other live vertex gets pid =
// 3 * newid[pid] + class (0 unreached, 1 forward only, 2 backward only)
// and, with `keys`, its pivot key goes to pbest[new pid] (maximum,
// aggregated): the next round's keys, over pbest zeroed for its ids.
// Per-workgroup decided counts into partials; the decided vertices are
// listed in dlist (count *dcount), staged in LDS with one atomic per flush.
constexpr int PS_U = 4;                               // k_part_split: vertices per thread and step
constexpr uint32_t PS_SPAN = (uint32_t)WG * PS_U;     // vertices per workgroup and step

__global__ __launch_bounds__(WG) void k_part_split(uint32_t n, uint32_t* live, const uint32_t* __restrict__ fwv,
                                                   const uint32_t* __restrict__ bwv, uint32_t* __restrict__ pid,
                                                   const uint32_t* __restrict__ newid, uint32_t* pmin,
                                                   const uint32_t* __restrict__ prio, uint64_t rnd_seed, bool keys,
                                                   unsigned long long* pbest, uint64_t* __restrict__ partials,
                                                   uint32_t* __restrict__ dlist, uint32_t* dcount) {
  __shared__ uint32_t s_mk[AGG_SLOTS];
  __shared__ uint32_t s_mv[AGG_SLOTS];
  __shared__ uint32_t s_bk[AGG_SLOTS];
  __shared__ unsigned long long s_bv[AGG_SLOTS];
  __shared__ uint64_t s_red[WG / WAVE];
  __shared__ uint32_t s_dq[2 * PS_SPAN];
  __shared__ uint32_t s_dn, s_db;
  BlockQueue dq{s_dq, &s_dn, &s_db, 2 * PS_SPAN};
  agg_init<uint32_t>(s_mk, s_mv, 0xFFFFFFFFu);
  agg_init<unsigned long long>(s_bk, s_bv, 0ull);
  bq_init(dq);
  // PS_U vertices per thread and step (one per 256-vertex tile): the
  // loads of each kind are issued for all of them before any is used
  constexpr uint64_t SPAN = PS_SPAN;
  const uint64_t pad = ((uint64_t)n + SPAN - 1) / SPAN * SPAN;
  const uint32_t lane = lane_id();
  uint64_t count = 0;
  for (uint64_t base = (uint64_t)blockIdx.x * SPAN; base < pad; base += (uint64_t)gridDim.x * SPAN) {
    uint32_t lw[PS_U], fw[PS_U], bw[PS_U], p[PS_U], k[PS_U], pr[PS_U];
    bool is_live[PS_U];
#pragma unroll
    for (int u = 0; u < PS_U; ++u) {
      const uint64_t v = base + (uint64_t)u * WG + threadIdx.x;
      lw[u] = v < n ? live[v >> 5] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PS_U; ++u) {
      const uint64_t v = base + (uint64_t)u * WG + threadIdx.x;
      is_live[u] = v < n && ((lw[u] >> (v & 31)) & 1u);
      fw[u] = is_live[u] ? fwv[v >> 5] : 0u;
      bw[u] = is_live[u] ? bwv[v >> 5] : 0u;
      p[u] = is_live[u] ? pid[v] : 0u;
      pr[u] = is_live[u] && keys ? prio[v] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PS_U; ++u) k[u] = is_live[u] ? newid[p[u]] : 0u;
#pragma unroll
    for (int u = 0; u < PS_U; ++u) {
      const uint64_t v = base + (uint64_t)u * WG + threadIdx.x;
      bool dec = false;
      uint32_t np = 0;
      unsigned long long key = 0;
      if (is_live[u]) {
        const bool f = (fw[u] >> (v & 31)) & 1u, b = (bw[u] >> (v & 31)) & 1u;
        dec = f && b;
        if (!dec) {
          np = 3u * k[u] + (f ? 1u : (b ? 2u : 0u));
          pid[v] = np;
          if (keys) key = part_key_of(pr[u], (uint32_t)v, rnd_seed);
        }
      }
      agg_wave_put<uint32_t, false>(s_mk, s_mv, pmin, dec, k[u], (uint32_t)v);
      if (keys) agg_wave_put<unsigned long long, true>(s_bk, s_bv, pbest, is_live[u] && !dec, np, key);
      bq_push(dq, dec, (uint32_t)v);
      const uint64_t mask = __ballot(dec);
      if (mask) {
        const uint64_t word0 = (v - lane) >> 5;
        if (lane == 0 && (uint32_t)mask) live[word0] = lw[u] & ~(uint32_t)mask;
        if (lane == 32 && (uint32_t)(mask >> 32)) live[word0 + 1] = lw[u] & ~(uint32_t)(mask >> 32);
        count += dec;
      }
    }
    bq_maybe_flush(dq, dlist, dcount, PS_SPAN);  // a step lists at most PS_SPAN vertices
  }
  bq_flush(dq, dlist, dcount);
  agg_flush<uint32_t, false>(s_mk, s_mv, pmin);
  if (keys) agg_flush<unsigned long long, true>(s_bk, s_bv, pbest);
  count = block_sum64(count, s_red);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = count;
    partials[3 * blockIdx.x + 1] = 0;
    partials[3 * blockIdx.x + 2] = 0;
  }
}

// the vertices k_part_split decided (forward and backward reached: every
// reached vertex was live when the round began), listed in dlist: label =
// the smallest decided id of their partition (pid still names the round's
// partition); PL_U per thread and step, each kind of load issued for all
__global__ void k_part_label(const uint32_t* __restrict__ dlist, const uint32_t* dcount, uint32_t* __restrict__ pid,
                             const uint32_t* __restrict__ newid, const uint32_t* __restrict__ pmin,
                             uint32_t* __restrict__ labels) {
  constexpr int PL_U = 4;
  constexpr uint64_t SPAN = (uint64_t)WG * PL_U;
  const uint32_t cnt = *dcount;
  for (uint64_t base = blockIdx.x * SPAN; base < cnt; base += (uint64_t)gridDim.x * SPAN) {
    uint32_t v[PL_U], k[PL_U];
    bool ok[PL_U];
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      const uint64_t i = base + (uint64_t)u * WG + threadIdx.x;
      ok[u] = i < cnt;
      v[u] = ok[u] ? dlist[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < PL_U; ++u) k[u] = ok[u] ? pid[v[u]] : 0u;
#pragma unroll
    for (int u = 0; u < PL_U; ++u) k[u] = ok[u] ? newid[k[u]] : 0u;
#pragma unroll
    for (int u = 0; u < PL_U; ++u) {
      if (!ok[u]) continue;
      labels[v[u]] = pmin[k[u]];
      pid[v[u]] = DEAD_PID;  // no longer live
    }
  }
}

// partition ids at the start of the multi-pivot rounds: DEAD_PID for the
// vertices that are not live; the live keep theirs (preset: the graph of a
// batch) or get 0
__global__ void k_pid_init(uint32_t n, const uint32_t* __restrict__ live, uint32_t* __restrict__ pid, bool keep) {
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG) {
    if (!bm_get(live, (uint32_t)v)) pid[v] = DEAD_PID;
    else if (!keep) pid[v] = 0u;
  }
}

// pivot outcome records of one round (KSPN_LOG_WORDS u32 each), one per used partition
__global__ void k_part_log(uint32_t S, const unsigned long long* __restrict__ pbest, const uint32_t* __restrict__ newid,
                           const uint32_t* __restrict__ pcount, uint32_t round, uint32_t* __restrict__ log, uint32_t first,
                           uint32_t cap) {
  for (uint32_t q = blockIdx.x * WG + threadIdx.x; q < S; q += gridDim.x * WG) {
    if (pbest[q] == 0ull) continue;
    const uint32_t k = newid[q];
    if (first + k >= cap) continue;
    uint32_t* r = log + (uint64_t)(first + k) * KSPN_LOG_WORDS;
    const uint32_t* c = pcount + 4ull * k;
    r[0] = ~(uint32_t)pbest[q];
    r[1] = round;
    r[2] = c[0] + c[1] + c[2] + c[3];
    r[3] = c[0];
    r[4] = c[1];
    r[5] = c[2];
    r[6] = c[3];
    r[7] = 0;
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t count, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < count; i += (uint64_t)gridDim.x * WG) p[i] = v;
}

// out[list[i]] = vals[i]
__global__ void k_scatter_f32(const uint32_t* __restrict__ list, const uint32_t* count, const float* __restrict__ vals,
                              float* __restrict__ out) {
  uint32_t cnt = *count;
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < cnt; i += gridDim.x * WG) out[list[i]] = vals[i];
}

// batches: pid[v] = i for the vertices of graph i (one workgroup per graph)
__global__ void k_batch_pid(uint32_t count, const uint32_t* __restrict__ vtx_off, uint32_t* __restrict__ pid) {
  for (uint32_t g = blockIdx.x; g < count; g += gridDim.x) {
    const uint32_t b = vtx_off[g], e = vtx_off[g + 1];
    for (uint32_t v = b + threadIdx.x; v < e; v += WG) pid[v] = g;
  }
}

// batch labels hold the SCC's smallest batch-global id, a vertex of the same
// graph: subtracting the graph's first id gives the local canonical id
__global__ void k_batch_localize(uint32_t count, const uint32_t* __restrict__ vtx_off, uint32_t* __restrict__ labels) {
  for (uint32_t g = blockIdx.x; g < count; g += gridDim.x) {
    const uint32_t b = vtx_off[g], e = vtx_off[g + 1];
    for (uint32_t v = b + threadIdx.x; v < e; v += WG) labels[v] -= b;
  }
}

// =====================================================================
// host side
// =====================================================================

namespace {

template <class T>
int dalloc(T** p, uint64_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return KSPN_E_OOM;
  }
  return KSPN_OK;
}

}  // namespace

thread_local char g_last_error[1024] = "";

void note_error(const char* what, const char* detail, const char* file, int line) {
  const char* base = std::strrchr(file, '/');
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%s:%d)", what, detail, base ? base + 1 : file, line);
}

void clear_error() { g_last_error[0] = 0; }

namespace {

enum KernId : int {
  K_TRIM_FIRST = 0, K_PIVOT, K_BITMAP, K_FRONTIER_DEG, K_SCAN, K_EXPAND_FW, K_EXPAND_BW, K_PULL_FW, K_PULL_BW,
  K_TRIM_SWEEP, K_TRIM2, K_EXPAND_NEIGH, K_COLOR_INIT, K_COLOR_PULL, K_EXPAND_COLOR, K_COLOR_ROOTS,
  K_EXPAND_COLOR_BW, K_LIVE_LIST, K_CANON, K_RESIDUAL, K_TARJAN, K_LABEL, K_TAIL, K_COOP, K_PULL_LIST,
  K_CHUNK_FW, K_CHUNK_BW, K_COUNT
};
const char* kKernNames[K_COUNT] = {"trim_first", "pivot_reduce", "bitmap_ops", "frontier_deg", "scan", "expand_fw",
                                   "expand_bw", "pull_fw", "pull_bw", "trim_sweep", "trim2", "expand_neigh",
                                   "color_init", "color_pull", "expand_color", "color_roots", "expand_color_bw",
                                   "live_list", "canonicalize", "residual_build", "tarjan", "label_scc", "tail_solve",
                                   "reach_coop", "pull_list", "expand_chunk_fw", "expand_chunk_bw"};

struct Prof {
  Ctx* c;
  hipStream_t s;
  int idx = -1;
  Prof(Ctx* c_, hipStream_t s_, int kern, double bytes) : c(c_), s(s_) {
    if (!c->profile) return;
    EventPair ep{};
    if (c->free_events.size() >= 2) {
      ep.a = c->free_events.back();
      c->free_events.pop_back();
      ep.b = c->free_events.back();
      c->free_events.pop_back();
    } else {
      if (hipEventCreate(&ep.a) != hipSuccess || hipEventCreate(&ep.b) != hipSuccess) return;
    }
    ep.kern = kern;
    ep.phase = c->cur_phase;
    ep.bytes = bytes;
    (void)hipEventRecord(ep.a, s);
    c->pending.push_back(ep);
    idx = (int)c->pending.size() - 1;
  }
  void add_bytes(double b) {
    if (idx >= 0) c->pending[idx].bytes += b;
  }
  ~Prof() {
    if (idx >= 0) (void)hipEventRecord(c->pending[idx].b, s);
  }
};

void prof_collect(Ctx* c) {
  for (auto& ep : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ep.a, ep.b) == hipSuccess) {
      c->kt[ep.kern].launches++;
      c->kt[ep.kern].ms += ms;
      c->kt[ep.kern].bytes += ep.bytes;
      if (ep.phase >= 0 && ep.phase < KSPN_PHASE_MAX) {
        c->phase_bytes[ep.phase] += ep.bytes;
        c->phase_kms[ep.phase] += ms;
      }
    }
    c->free_events.push_back(ep.a);
    c->free_events.push_back(ep.b);
  }
  c->pending.clear();
}

inline uint32_t expand_grid(Ctx* c, uint64_t total, uint32_t ept = EXP_EPT) {
  const uint64_t tile = (uint64_t)WG * ept;
  uint64_t tiles = (total + tile - 1) / tile;
  uint64_t cap = (uint64_t)c->num_cu * 16;
  if (tiles < 1) tiles = 1;
  return (uint32_t)std::min<uint64_t>(tiles, cap);
}

// workgroups of a k_fwbw_finish launch over `words` words (its partials)
inline uint32_t finish_grid(uint64_t words) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((words + FIN_CHUNK - 1) / FIN_CHUNK, FINISH_PARTS));
}

// workgroups of a k_bitmap_to_queue launch over `words` words: one chunk each
inline uint32_t b2q_grid(uint64_t words) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((words + B2Q_CHUNK - 1) / B2Q_CHUNK, 1u << 20));
}

inline uint32_t stream_grid(Ctx* c, uint64_t items) {
  uint64_t g = (items + WG - 1) / WG;
  uint64_t cap = (uint64_t)c->num_cu * 16;
  if (g < 1) g = 1;
  return (uint32_t)std::min<uint64_t>(g, cap);
}

}  // namespace

int ctx_reserve(Ctx* c, uint32_t n, uint64_t m) {
  if (n <= c->cap_n && m <= c->cap_m && c->live) return KSPN_OK;
  uint32_t nn = std::max(n, c->cap_n);
  uint64_t mm = std::max(m, c->cap_m);
  if (nn < 64) nn = 64;
  uint64_t words = ((uint64_t)nn + 255) / 256 * 8 + 16;  // covers the 256-vertex padding of ballot writers
  KSPN_TRY(dalloc(&c->live, words));
  KSPN_TRY(dalloc(&c->fwv, words));
  KSPN_TRY(dalloc(&c->bwv, words));
  KSPN_TRY(dalloc(&c->inq, words));
  KSPN_TRY(dalloc(&c->inq2, words));
  if (!c->cchain) KSPN_TRY(dalloc(&c->cchain, 16));
  KSPN_TRY(dalloc(&c->fnew, words));
  KSPN_TRY(dalloc(&c->sbm, words));
  KSPN_TRY(dalloc(&c->qa, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->qb, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->qc, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->ulist, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->col, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->deg, (uint64_t)nn + 64));
  KSPN_TRY(dalloc(&c->scan, (uint64_t)nn + 64));
  {
    // look-back scan: one status word per tile, then its ticket counter
    const uint64_t tiles = ((uint64_t)nn + DS_TILE) / DS_TILE + 1;
    KSPN_TRY(dalloc(&c->ds_st, tiles + 2));
    KSPN_HIP(hipMemset(c->ds_st, 0, (tiles + 2) * sizeof(unsigned long long)));
    c->ds_ticket = reinterpret_cast<uint32_t*>(c->ds_st + tiles + 1);
    c->ds_epoch = 0;
  }
  KSPN_TRY(dalloc(&c->part, (uint64_t)c->num_cu * 64 * 2 + 64));
  size_t b1 = 0, b2 = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, b1, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)(nn + 1)) != hipSuccess)
    return KSPN_E_HIP;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, b2, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)(words + 1)) != hipSuccess)
    return KSPN_E_HIP;
  // partition flag scans: 3 flags per partition, partitions <= vertices
  size_t b3 = 0;
  const uint64_t flags_max = std::min<uint64_t>(3ull * nn + 1, 0x7FFFFFFFull);
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, b3, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)flags_max) != hipSuccess)
    return KSPN_E_HIP;
  size_t need = std::max(std::max(b1, b2), b3) + 256;
  if (need > c->cub_bytes) {
    if (c->cub_tmp) (void)hipFree(c->cub_tmp);
    c->cub_tmp = nullptr;
    if (hipMalloc(&c->cub_tmp, need) != hipSuccess) return KSPN_E_OOM;
    c->cub_bytes = need;
  }
  // buffers sized by cap_n and allocated on first use: drop them so they are reallocated
  for (uint32_t** p : {&c->tj, &c->pid, &c->pmin, &c->pflags, &c->pnew, &c->pcnt, &c->vprio})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  for (void** p : {(void**)&c->pbest, (void**)&c->vscore, (void**)&c->vfeat, (void**)&c->sub_off, (void**)&c->cpw})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  c->cap_n = nn;
  c->cap_m = mm;
  return KSPN_OK;
}

// Wait until the publication with sequence number seq is visible in the
// mapped host buffer. Polling the sequence word returns about 5 us sooner
// than a stream synchronisation; after 20 ms of polling the stream is
// synchronised instead, which also reports a failed launch.
int wait_published(Ctx* c, hipStream_t s, uint32_t seq, uint32_t bank = 0) {
  const volatile uint32_t* m = c->h_map + MAP_BANK * bank;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spins = 0;; ++spins) {
    if (m[63] == seq) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return KSPN_OK;
    }
    if ((spins & 1023u) == 1023u && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
    __builtin_ia32_pause();
  }
  KSPN_HIP(hipStreamSynchronize(s));
  if (m[63] != seq) {
    note_error("k_publish", "counters not published after the stream completed", __FILE__, __LINE__);
    return KSPN_E_HIP;
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return KSPN_OK;
}

// publish the device counters (and *extra) to host-mapped memory, wait,
// and copy them into h_cnt / h_u64[0] (extra) / h_u64[1] (cnt64[0]).
// With part: reduce nparts per-workgroup partials first (into *pcount and
// cnt64[0..1]); with reset: zero *reset and cnt64[0..2] afterwards.
// The two halves of publish: the launch (into map bank `bank`; with `gate`
// the kernel publishes only if *gate is set, and the caller takes the
// sequence number back when it learns that it did not) and the wait + copy.
uint32_t publish_launch(Ctx* c, hipStream_t s, const uint64_t* extra, const uint64_t* part, uint32_t nparts,
                        uint32_t* pcount, uint32_t* reset, uint32_t* reset2, uint32_t bank = 0,
                        const uint32_t* gate = nullptr, SpecGate sg = {}) {
  const uint32_t seq = ++c->pub_seq;
  k_publish<<<1, WG, 0, s>>>(c->cnt, c->cnt64, extra, c->d_map + MAP_BANK * bank, seq, part, nparts, pcount, reset,
                             reset2, gate, sg);
  return seq;
}

int publish_collect(Ctx* c, hipStream_t s, uint32_t seq, uint32_t bank = 0) {
  KSPN_HIP(hipGetLastError());
  KSPN_TRY(wait_published(c, s, seq, bank));
  const volatile uint32_t* m = c->h_map + MAP_BANK * bank;
  for (int i = 0; i < C_COUNT_; ++i) c->h_cnt[i] = m[i];
  c->h_u64[0] = (uint64_t)m[48] | ((uint64_t)m[49] << 32);
  c->h_u64[1] = (uint64_t)m[50] | ((uint64_t)m[51] << 32);
  c->h_u64[2] = (uint64_t)m[52] | ((uint64_t)m[53] << 32);
  c->h_u64[3] = (uint64_t)m[54] | ((uint64_t)m[55] << 32);
  return KSPN_OK;
}

int publish(Ctx* c, hipStream_t s, const uint64_t* extra, const uint64_t* part = nullptr, uint32_t nparts = 0,
            uint32_t* pcount = nullptr, uint32_t* reset = nullptr, uint32_t* reset2 = nullptr) {
  return publish_collect(c, s, publish_launch(c, s, extra, part, nparts, pcount, reset, reset2));
}

int read_counters(Ctx* c, hipStream_t s, int count) {
  (void)count;
  return publish(c, s, nullptr);
}

int exclusive_scan_u64(Ctx* c, const uint64_t* in, uint64_t* out, uint64_t items, hipStream_t s) {
  Prof p(c, s, K_SCAN, 16.0 * items);
  size_t bytes = c->cub_bytes;
  KSPN_HIP(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, bytes, in, out, (int)items, s));
  return KSPN_OK;
}

int exclusive_scan_u32(Ctx* c, const uint32_t* in, uint32_t* out, uint64_t items, hipStream_t s) {
  Prof p(c, s, K_SCAN, 8.0 * items);
  size_t bytes = c->cub_bytes;
  KSPN_HIP(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, bytes, in, out, (int)items, s));
  return KSPN_OK;
}

// ascending list of set bits of `bitmap` into `list`; count to *count_out
// (host) and to cnt[C_AUX] (device)
int build_live_list(Ctx* c, uint32_t n, const uint32_t* bitmap, uint32_t* list, uint32_t* count_out, hipStream_t s) {
  uint64_t words = (n + 31) / 32;
  uint32_t* pc = reinterpret_cast<uint32_t*>(c->deg);   // words+1 popcounts
  uint32_t* pos = reinterpret_cast<uint32_t*>(c->scan); // scanned
  {
    Prof p(c, s, K_LIVE_LIST, 4.0 * words);
    k_bm_popc<<<stream_grid(c, words + 1), WG, 0, s>>>(bitmap, words, pc);
  }
  KSPN_TRY(exclusive_scan_u32(c, pc, pos, words + 1, s));
  {
    Prof p(c, s, K_LIVE_LIST, 8.0 * words);
    k_bm_scatter<<<stream_grid(c, words), WG, 0, s>>>(bitmap, words, pos, list);
  }
  k_copy_u32<<<1, 64, 0, s>>>(&c->cnt[C_AUX], pos + words);
  KSPN_TRY(publish(c, s, nullptr));
  *count_out = c->h_cnt[C_AUX];
  return KSPN_OK;
}

namespace {

struct Run {
  Ctx* c;
  const kspn_graph* g;
  uint32_t* labels;
  kspn_opts o;
  kspn_stats* st;
  hipStream_t s;
  uint32_t n;
  uint64_t words;
  uint32_t live_n = 0;  // host view of the live count
  uint32_t rotation = 0;
  bool col_valid = false;  // col is 0 for every vertex outside ulist (prev_u entries)
  uint32_t prev_u = 0;
  uint32_t rounds = 0;  // colour rounds so far
  bool pid_preset = false;  // partition ids already set (batched graphs)
  uint32_t log_count = 0;   // pivot outcome records written
  bool pid_active = false;  // pid partitions the live vertices (after multi-pivot FW-BW)
  uint32_t np0 = 1;         // number of preset partitions
  const uint32_t* vtx_off = nullptr;  // batch: graph i owns [vtx_off[i], vtx_off[i+1]); labels are made local
  unsigned long long coop_edges = 0;  // frontiers the persistent kernel takes (degree sum, largest degree)
  uint32_t coop_deg = 0;
  uint32_t coop_wg = 0;               // its workgroups
  uint32_t pull_ilp = PULL_ILP;       // launch shapes (kspn_opts)
  uint32_t ept = EXP_EPT;
  uint32_t chain = 0;                 // KSPN_FLAG_CHAIN: launches per chain of push levels (0: persistent kernel)
  uint32_t narrow_len = 0;            // KSPN_FLAG_NARROW: launches per chain of narrow levels (0: off)
  bool prune = false;                 // multi-pivot rounds with KSPN_FLAG_PRUNE: the edge bitmaps are current
  bool dead_pid = false;              // pid holds DEAD_PID for every vertex that is not live (multi-pivot rounds)
  bool pivot_live_deg = false;        // pivots by live degree product (after an exhaustive trim)
  // the first trim's decided count and pivot are still on the device
  // (cnt[C_DECIDED], cnt[C_FLAG]); the first publication of the forward
  // reach brings them to the host
  bool live_pending = false;
  // fwbw(tail = true) leaves its label pass to tail_spec (k_label_prep)
  bool label_pending = false;
  uint32_t label_nparts = 0;  // k_fwbw_finish's partials
  bool label_dense = false;   // write every label (the first trim wrote none)
  uint32_t trim_decided = 0, first_pivot = 0;
};

int select_pivot_live(Run& R, uint32_t* pivot);

// S = c->scan[0..f]: exclusive scan of the out-degrees of queue F (count on
// device at fcount, host copy f); S[f] is the total. One workgroup below
// 4096 entries, else the single-pass look-back scan (hipCUB's device scan
// where a total could reach 2^48).
// With `tiles` (expansion tiles of 2^tile_shift edges), the look-back scan
// also writes their first frontier entries to c->tstart and *tiles is set;
// the other scans leave *tiles false.
int frontier_scan(Run& R, const uint32_t* F, const uint32_t* fcount_dev, uint32_t f, const uint64_t* off,
                  bool* tiles_out = nullptr, uint32_t tile_shift = 11) {
  Ctx* c = R.c;
  Prof p(c, R.s, K_FRONTIER_DEG, 28.0 * f);
  if (tiles_out) *tiles_out = false;
  if (f < 4096) {  // S[0..f] fits the 4096 slots of one workgroup
    k_deg_scan_small<<<1, 1024, 0, R.s>>>(F, fcount_dev, off, c->scan);
    return KSPN_OK;
  }
  if (R.g->m >= DS_VALUE) {
    k_frontier_deg<<<stream_grid(c, (uint64_t)f + 1), WG, 0, R.s>>>(F, fcount_dev, off, c->deg);
    return exclusive_scan_u64(c, c->deg, c->scan, (uint64_t)f + 1, R.s);
  }
  if (++c->ds_epoch >= DS_EPOCHS) {
    // status words of this epoch may remain from DS_EPOCHS launches ago
    KSPN_HIP(hipMemsetAsync(c->ds_st, 0, ((uint64_t)(c->cap_n + DS_TILE) / DS_TILE + 1) * sizeof(unsigned long long),
                            R.s));
    c->ds_epoch = 1;
  }
  const uint32_t tiles = (uint32_t)(((uint64_t)f + DS_TILE) / DS_TILE);
  uint32_t* tstart = nullptr;
  if (tiles_out) {
    // every tile of the largest level: m edges, tiles of at least 2^9 edges
    const uint64_t need = (R.g->m >> 9) + 2;
    if (c->tstart_cap < need) {
      c->tstart_cap = 0;
      KSPN_TRY(dalloc(&c->tstart, need));
      c->tstart_cap = need;
    }
    if (tile_shift >= 9) {
      tstart = c->tstart;
      *tiles_out = true;
    }
  }
  k_deg_scan<<<tiles, WG, 0, R.s>>>(F, fcount_dev, off, c->scan, c->ds_st, c->ds_ticket, tiles, c->ds_epoch,
                                    &c->cnt[C_SERR], tstart, tile_shift, c->tstart_cap);
  return KSPN_OK;
}

// Edge-balanced expansion of queue F (count on device at fcount, host copy
// f) over CSR (off, adj). Appends to `next` with its counter at next_cnt.
// Returns the expanded edge count through *edges.
template <int OP>
int expand(Run& R, const uint32_t* F, const uint32_t* fcount_dev, uint32_t f, const uint64_t* off, const uint32_t* adj,
           const OpArgs& a, uint32_t* next, uint32_t* next_cnt, int kern, double state_bytes_per_edge,
           uint64_t* edges, bool known_total = false, uint64_t total_hint = 0) {
  Ctx* c = R.c;
  if (f == 0) {
    if (edges) *edges = 0;
    return KSPN_OK;
  }
  if constexpr (OP == OP_COLOR || OP == OP_COLOR_BW) {
    if (R.o.flags & KSPN_FLAG_CHUNKED) {
      // chunks of the frontier, no degree scan and no publication of its
      // edge count (statistics: summed into cnt64[3], read by color_round;
      // profile mode reads it around the launch for the level's bytes)
      unsigned long long e_before = 0, e_after = 0;
      if (c->profile) {
        KSPN_HIP(hipMemcpyAsync(&e_before, c->cnt64 + 3, sizeof(e_before), hipMemcpyDeviceToHost, R.s));
        KSPN_HIP(hipStreamSynchronize(R.s));
      }
      {
        Prof p(c, R.s, kern, 20.0 * f);
        const uint32_t cgrid = (uint32_t)std::min<uint64_t>(((uint64_t)f + CHUNK_N - 1) / CHUNK_N,
                                                           (uint64_t)c->num_cu * 16);
        k_expand_chunk<OP><<<cgrid, WG, 0, R.s>>>(F, fcount_dev, off, adj, a, next, next_cnt, &c->cnt[C_ERR],
                                                   nullptr, 0, R.st || c->profile ? c->cnt64 + 3 : nullptr);
      }
      if (c->profile) {
        KSPN_HIP(hipMemcpyAsync(&e_after, c->cnt64 + 3, sizeof(e_after), hipMemcpyDeviceToHost, R.s));
        KSPN_HIP(hipStreamSynchronize(R.s));
        if (!c->pending.empty())
          c->pending.back().bytes = 20.0 * f + (4.0 + state_bytes_per_edge) * (double)(e_after - e_before);
      }
      if (edges) *edges = 0;
      return KSPN_OK;
    }
  }
  bool tiles = false;
  KSPN_TRY(frontier_scan(R, F, fcount_dev, f, off, &tiles, 11));  // k_expand<OP>: tiles of WG * 8 edges
  uint64_t total = total_hint;
  if (!known_total) {
    KSPN_TRY(publish(c, R.s, c->scan + f));
    if (c->h_cnt[C_SERR]) {
      note_error("k_deg_scan", "frontier count or ticket out of step with the launch", __FILE__, __LINE__);
      return KSPN_E_HIP;
    }
    total = c->h_u64[0];
  }
  if (edges) *edges = total;
  if (total == 0) return KSPN_OK;
  {
    Prof p(c, R.s, kern, 4.0 * f + 8.0 * f + 8.0 * f + (4.0 + state_bytes_per_edge) * total);
    k_expand<OP><<<expand_grid(c, total), WG, 0, R.s>>>(F, fcount_dev, c->scan, off, adj, a, next, next_cnt,
                                                         tiles ? c->tstart : nullptr, c->tstart_cap);
  }
  if (R.st) R.st->edges_scanned += total;
  return KSPN_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

// ---------------------------------------------------------------- FW-BW
// Reachability from the pivot forward (restricted to live vertices), then
// backward (restricted to reached ones); the backward set is the pivot's
// SCC and is decided with id = pivot. Each level is push (edge-balanced
// expansion of the frontier) or pull (sweep over all vertices), switched
// by Beamer's rule: pull when frontier edges * alpha > remaining edges,
// back to push when frontier * beta < live.
// Reachability in direction dir (0 forward, 1 backward) from the frontier
// the caller put in qa (count f0, also in cnt[C_Q0]; their vis bits set).
// Single partition: forward claims live vertices, backward claims
// forward-reached ones. With partitions (pid != nullptr) both directions
// claim live vertices of the source's partition.
// frontiers the persistent kernel takes by default: degree sum and largest degree bounds
constexpr uint64_t kCoopEdges = 1ull << 20;
constexpr uint32_t kCoopDeg = 2048;
constexpr uint32_t kCoopWorkgroups = 128;
constexpr uint32_t kChainLevels = 16;  // launches per chain of push levels (KSPN_FLAG_CHAIN)

struct CoopOut {
  uint32_t levels, reason, f, maxd;
  uint64_t edges, expanded, verts;
};

// push levels of one reachability in one persistent launch, starting from
// the frontier F (count at cnt[fc]) whose degree sum and maximum are known
// (KSPN_FLAG_CHAIN: up to R.chain levels as launches of their own, enqueued
// together, then one publication)
int run_coop(Run& R, int op, const OpArgs& a, const uint64_t* off, const uint32_t* adj, uint32_t* F, uint32_t* N,
             int fc, int nc, uint64_t edges, uint32_t maxd, double m_rest, double alpha, CoopOut* out,
             bool narrow = false) {
  Ctx* c = R.c;
  k_coop_init<<<1, 64, 0, R.s>>>(c->coop, &c->cnt[fc], edges, maxd, m_rest);
  KSPN_HIP(hipGetLastError());
  const float al = (float)alpha;
  if (narrow) {
    // narrow levels (every frontier degree <= NARROW_DEG): a chain of
    // R.narrow_len launches of four workgroups per CU (a launch whose
    // frontier is small, or that finds the run ended, costs less with fewer
    // workgroups: lattice 256^3 21.4 -> 19.8 ms against eight per CU; one
    // per CU was no faster, sixteen took 54 ms)
    Prof p(c, R.s, K_COOP, 0.0);
    const uint32_t grid = (uint32_t)c->num_cu * 4;
    for (uint32_t L = 0; L < R.narrow_len; ++L) {
      if (op == OP_FW)
        k_narrow_level<OP_FW><<<grid, NARROW_WG, 0, R.s>>>(F, N, off, adj, a, c->coop, al, NARROW_EDGES, L);
      else if (op == OP_BW)
        k_narrow_level<OP_BW><<<grid, NARROW_WG, 0, R.s>>>(F, N, off, adj, a, c->coop, al, NARROW_EDGES, L);
      else
        k_narrow_level<OP_PBW><<<grid, NARROW_WG, 0, R.s>>>(F, N, off, adj, a, c->coop, al, NARROW_EDGES, L);
    }
    k_chain_end<<<1, 64, 0, R.s>>>(c->coop, R.narrow_len);
    KSPN_HIP(hipGetLastError());
  } else if (R.chain) {
    Prof p(c, R.s, K_COOP, 0.0);
    const dim3 grid(R.coop_wg), block(COOP_WG);
    for (uint32_t L = 0; L < R.chain; ++L) {
      if (op == OP_FW)
        k_chain_level<OP_FW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, al, R.coop_edges, R.coop_deg, L);
      else if (op == OP_BW)
        k_chain_level<OP_BW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, al, R.coop_edges, R.coop_deg, L);
      else
        k_chain_level<OP_PBW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, al, R.coop_edges, R.coop_deg, L);
    }
    k_chain_end<<<1, 64, 0, R.s>>>(c->coop, R.chain);
    KSPN_HIP(hipGetLastError());
  } else {
    Prof p(c, R.s, K_COOP, 0.0);
    const dim3 grid(R.coop_wg), block(COOP_WG);
    if (op == OP_FW)
      k_reach_coop<OP_FW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, m_rest, al, R.coop_edges, R.coop_deg);
    else if (op == OP_BW)
      k_reach_coop<OP_BW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, m_rest, al, R.coop_edges, R.coop_deg);
    else
      k_reach_coop<OP_PBW><<<grid, block, 0, R.s>>>(R.n, F, N, off, adj, a, c->coop, m_rest, al, R.coop_edges, R.coop_deg);
    KSPN_HIP(hipGetLastError());
  }
  const uint32_t seq = ++c->pub_seq;
  k_coop_publish<<<1, 64, 0, R.s>>>(c->coop, &c->cnt[fc], &c->cnt[nc], c->cnt64, c->d_map, seq);
  KSPN_HIP(hipGetLastError());
  KSPN_TRY(wait_published(c, R.s, seq));
  const volatile uint32_t* m = c->h_map;
  out->levels = m[32];
  out->reason = m[33];
  out->f = m[35];
  out->edges = (uint64_t)m[36] | ((uint64_t)m[37] << 32);
  out->maxd = m[38];
  out->expanded = (uint64_t)m[39] | ((uint64_t)m[40] << 32);
  out->verts = (uint64_t)m[41] | ((uint64_t)m[42] << 32);
  if (m[34] || out->reason == COOP_ERR) {
    note_error("k_reach_coop", "grid barrier timed out (workgroups not resident together)", __FILE__, __LINE__);
    return KSPN_E_HIP;
  }
  if (c->profile && !c->pending.empty()) c->pending.back().bytes = 20.0 * out->verts + 4.125 * (double)out->expanded;
  return KSPN_OK;
}

// With `seeded`, the frontier queue, the level counters and the first
// level's degree scan are already on the device (k_pivot_reduce's FwbwSeed,
// k_reach_start).
// A whole reach as one asynchronous launch (KSPN_FLAG_ASYNC): the
// frontier F (count cnt[fc], vis bits set) seeds the slot queue.
int reach_async(Run& R, int op, const OpArgs& a, const uint64_t* off, const uint32_t* adj, const uint32_t* F, int fc,
                uint32_t f_hint) {
  Ctx* c = R.c;
  const uint64_t qcap = (uint64_t)R.n + R.g->m / AS_SLICE + 64;
  if (qcap > 0xFFFFFFF0ull) {
    note_error("reach_async", "graph too large for the 32-bit slot queue", __FILE__, __LINE__);
    return KSPN_E_UNIMPLEMENTED;
  }
  if (!c->actl) KSPN_HIP(hipMalloc(&c->actl, sizeof(AsyncCtl)));
  if (qcap > c->aq_cap) {
    if (c->aq) (void)hipFree(c->aq);
    c->aq = nullptr;
    c->aq_cap = 0;
    KSPN_TRY(dalloc(&c->aq, qcap));
    KSPN_HIP(hipMemsetAsync(c->aq, 0xFF, qcap * sizeof(unsigned long long), R.s));  // every slot empty
    c->aq_cap = qcap;
  }
  AsyncCtl* ctl = static_cast<AsyncCtl*>(c->actl);
  const uint32_t grid = R.o.coop_workgroups ? R.o.coop_workgroups : 512u;
  {
    Prof p(c, R.s, K_COOP, 0.0);
    k_async_init<<<1, 64, 0, R.s>>>(ctl);
    k_async_seed<<<stream_grid(c, std::max<uint32_t>(f_hint, 1)), WG, 0, R.s>>>(F, &c->cnt[fc], off, c->aq,
                                                                              (uint32_t)qcap, ctl);
    if (op == OP_FW) k_reach_async<OP_FW><<<grid, AS_WG, 0, R.s>>>(off, adj, a, c->aq, (uint32_t)qcap, ctl);
    else if (op == OP_BW) k_reach_async<OP_BW><<<grid, AS_WG, 0, R.s>>>(off, adj, a, c->aq, (uint32_t)qcap, ctl);
    else k_reach_async<OP_PBW><<<grid, AS_WG, 0, R.s>>>(off, adj, a, c->aq, (uint32_t)qcap, ctl);
    k_async_done<<<1, 64, 0, R.s>>>(ctl, c->cnt, c->cnt64);
  }
  // published, then C_AERR and cnt64[0..2] are zeroed again
  KSPN_TRY(publish(c, R.s, nullptr, nullptr, 0, nullptr, &c->cnt[C_AERR]));
  if (c->h_cnt[C_AERR]) {
    // the queue may hold entries: empty it for the next reach
    KSPN_HIP(hipMemsetAsync(c->aq, 0xFF, c->aq_cap * sizeof(unsigned long long), R.s));
    note_error("reach_async",
               c->h_cnt[C_AERR] & 2u   ? "slot queue overflow"
               : c->h_cnt[C_AERR] & 1u ? "no progress within the spin limit"
                                       : "the reach ended with entries not expanded",
               __FILE__, __LINE__);
    return KSPN_E_HIP;
  }
  const uint64_t edges = c->h_u64[1];
  if (R.st) {
    R.st->edges_scanned += edges;
    R.st->fwbw_levels_coop += 1;
  }
  if (c->profile && !c->pending.empty()) {
    // the reach's profile entry: per claimed vertex its id, offsets and queue slot; per edge an id and a bit
    for (size_t i = c->pending.size(); i-- > 0;)
      if (c->pending[i].kern == K_COOP) {
        c->pending[i].bytes = 28.0 * (double)c->h_u64[2] + 4.125 * (double)edges;
        break;
      }
  }
  if (R.live_pending) {
    R.trim_decided = c->h_cnt[C_DECIDED];
    R.first_pivot = c->h_cnt[C_FLAG];
    R.live_n = R.n - R.trim_decided;
    R.live_pending = false;
  }
  return KSPN_OK;
}

// a push level of a reach with the tile width the options ask for
// (tstart: the scan's tile starts for tiles of WG * ept edges, or nullptr)
template <int OP>
void launch_expand_ept(uint32_t ept, uint32_t grid, hipStream_t s, const uint32_t* F, const uint32_t* fc,
                       const uint64_t* S, const uint64_t* off, const uint32_t* adj, const OpArgs& a, uint32_t* N,
                       uint32_t* nc, const uint32_t* tstart, uint64_t tcap) {
  if (ept == 2) k_expand<OP, 2><<<grid, WG, 0, s>>>(F, fc, S, off, adj, a, N, nc, tstart, tcap);
  else if (ept == 4) k_expand<OP, 4><<<grid, WG, 0, s>>>(F, fc, S, off, adj, a, N, nc, tstart, tcap);
  else if (ept == 16) k_expand<OP, 16><<<grid, WG, 0, s>>>(F, fc, S, off, adj, a, N, nc, tstart, tcap);
  else k_expand<OP, EXP_EPT><<<grid, WG, 0, s>>>(F, fc, S, off, adj, a, N, nc, tstart, tcap);
}

void launch_reach_expand(int op, uint32_t ept, uint32_t grid, hipStream_t s, const uint32_t* F, const uint32_t* fc,
                         const uint64_t* S, const uint64_t* off, const uint32_t* adj, const OpArgs& a, uint32_t* N,
                         uint32_t* nc, const uint32_t* tstart, uint64_t tcap) {
  if (op == OP_FW) launch_expand_ept<OP_FW>(ept, grid, s, F, fc, S, off, adj, a, N, nc, tstart, tcap);
  else if (op == OP_PBW) launch_expand_ept<OP_PBW>(ept, grid, s, F, fc, S, off, adj, a, N, nc, tstart, tcap);
  else launch_expand_ept<OP_BW>(ept, grid, s, F, fc, S, off, adj, a, N, nc, tstart, tcap);
}

// a pull sweep with the width the options ask for
template <int DIR>
void launch_pull_ilp(uint32_t ilp, uint32_t grid, hipStream_t s, uint32_t n, const uint32_t* cand, uint32_t* vis,
                     const uint64_t* roff, const uint32_t* radj, uint32_t* fnew, uint64_t* part, const uint32_t* pid,
                     uint32_t* cursor, uint32_t* sbm, uint32_t* lcnt, const uint32_t* gate) {
  if (ilp == 1) k_pull<DIR, 1><<<grid, WG, 0, s>>>(n, cand, vis, roff, radj, nullptr, fnew, part, pid, cursor, sbm, lcnt, gate);
  else if (ilp == 4) k_pull<DIR, 4><<<grid, WG, 0, s>>>(n, cand, vis, roff, radj, nullptr, fnew, part, pid, cursor, sbm, lcnt, gate);
  else k_pull<DIR, PULL_ILP><<<grid, WG, 0, s>>>(n, cand, vis, roff, radj, nullptr, fnew, part, pid, cursor, sbm, lcnt, gate);
}

int reach(Run& R, int dir, uint32_t f0, const uint32_t* pid, uint32_t* fw_levels_push, uint32_t* fw_levels_pull,
          bool seeded = false) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  const uint64_t* off = dir == 0 ? g->fw_off : g->bw_off;
  const uint32_t* adj = dir == 0 ? g->fw_adj : g->bw_adj;
  const uint64_t* roff = dir == 0 ? g->bw_off : g->fw_off;
  const uint32_t* radj = dir == 0 ? g->bw_adj : g->fw_adj;
  uint32_t* vis = dir == 0 ? c->fwv : c->bwv;
  // backward: within the forward sets, or (partitioned, KSPN_FLAG_FULL_BW) over every live vertex of the partition
  const bool full_bw = pid && (R.o.flags & KSPN_FLAG_FULL_BW);
  const uint32_t* cand = (dir == 0 || full_bw) ? c->live : c->fwv;
  OpArgs a{c->live, c->fwv, c->bwv, c->inq, c->col, R.labels, 0, off, c->cnt64 + 1, pid, cand,
           pid && R.dead_pid ? 1u : 0u};
  if (pid && R.prune && dir == 0) a.prune = c->prune_fw;
  uint32_t* F = c->qa;
  uint32_t* N = c->qb;
  int fc = C_Q0, nc = C_Q1;
  uint32_t f = f0;
  uint32_t surv_n = 0;  // candidates the last pull level left unreached
  uint64_t edges = 0;  // out-edges of the frontier in this direction (unknown for the pivot level)
  bool known = false;
  bool pull = false;
  const double alpha = R.o.alpha > 0 ? R.o.alpha : 14.0;
  const double beta = R.o.beta > 0 ? R.o.beta : 24.0;
  double m_rest = (double)g->m;  // edges not yet expanded or reached
  uint32_t maxd = 0xFFFFFFFFu;   // largest out-degree in the frontier (unknown)
  const bool coop = (R.coop_wg > 0 || R.chain) && !(R.o.flags & KSPN_FLAG_NO_COOP);
  const int op = dir == 0 ? OP_FW : (pid ? OP_PBW : OP_BW);
  if (R.o.flags & KSPN_FLAG_ASYNC) return reach_async(R, op, a, off, adj, F, fc, f0);
  // A pull sweep leaves its unreached candidates in the bitmap sbm. When
  // they are few, the following pull levels visit only them: they are
  // compacted into a list, and list levels run in batches of LIST_BATCH
  // launches without a host round trip, each passing the candidates it
  // leaves to the next (lists alternate between qc and ulist, both free
  // while reaching; their counters rotate over four slots, which the sweep
  // zeroes, so a level can compare its count with its predecessor's). The
  // last level of a batch also records what it claimed, the frontier if
  // push levels follow; a batch whose last level claims nothing has reached
  // everything.
  constexpr uint32_t LIST_BATCH = 4;
  uint32_t* lists[2] = {c->qc, c->ulist};
  const int lcnt[4] = {C_SURV0, C_SURV1, C_SURV2, C_SURV3};
  const uint32_t list_max = std::max<uint32_t>(R.n / 16, 1u);  // longest remainder worth a list
  bool list_mode = false;   // the next pull level visits a list
  bool list_level = false;  // the level just run visited lists
  uint32_t L = 0;           // list levels since the last sweep
  // A sweep over all vertices: claims -> cnt[claim_slot], edges ->
  // cnt64[0], candidates left -> cnt64[1], and their list (lists[0]) when
  // there are at most list_max of them, so a batch after it does nothing
  // otherwise. With `gate`, every launch is conditional on the SpecGate.
  auto launch_sweep = [&](int claim_slot, const uint32_t* gate, int* prof) {
    // up to 2^25 vertices one resident round of workgroups (2048 beat 1024,
    // 4096 and 16384 on R-MAT 24), above it two (the 4096-graph batch, n =
    // 6e8, lost 2.4 % with one)
    const uint32_t per_cu = R.o.sweep_wgs ? R.o.sweep_wgs : (R.n <= (1u << 25) ? 8u : 16u);
    const uint32_t grid = std::min<uint32_t>(stream_grid(c, R.n), (uint32_t)c->num_cu * per_cu);
    {
      // candidate and visited bits of every vertex; the offsets of each
      // candidate and the scanned edges are added once they are published
      Prof p(c, R.s, dir == 0 ? K_PULL_FW : K_PULL_BW, (double)R.n / 4.0);
      if (prof) *prof = p.idx;
      // no out-degree sum of the claimed vertices: it would read a third
      // offset array at scattered rows, and the first push level after the
      // pull levels sums its frontier's degrees itself
      if (dir == 0)
        launch_pull_ilp<0>(R.pull_ilp, grid, R.s, R.n, cand, vis, roff, radj, c->fnew, c->part, pid,
                           &c->cnt[C_CURSOR], c->sbm, &c->cnt[C_SURV0], gate);
      else
        launch_pull_ilp<1>(R.pull_ilp, grid, R.s, R.n, cand, vis, roff, radj, c->fnew, c->part, pid,
                           &c->cnt[C_CURSOR], c->sbm, &c->cnt[C_SURV0], gate);
    }
    Prof p(c, R.s, K_LIVE_LIST, 4.0 * R.words);
    k_sweep_list<<<b2q_grid(R.words), WG, 0, R.s>>>(c->part, grid, &c->cnt[claim_slot], c->cnt64, c->sbm, R.words,
                                                    lists[0], &c->cnt[lcnt[0]], list_max, gate);
  };
  // LIST_BATCH list levels from list level L on; the last records its
  // claims in `rec` (count cnt[C_LREC])
  auto launch_batch = [&](uint32_t* rec, uint32_t n_in, int* prof) {
    // the levels' scanned edges, for statistics and profiles only
    unsigned long long* esum = R.st || c->profile ? c->cnt64 + 2 : nullptr;
    const uint32_t lgrid = std::min<uint32_t>(stream_grid(c, n_in / PULL_ILP + 1),
                                              (uint32_t)c->num_cu * (R.o.list_wgs ? R.o.list_wgs : 4u));
    for (uint32_t j = 0; j < LIST_BATCH; ++j, ++L) {
      // the first level zeroes the claim count the last one records
      uint32_t* q = j + 1 == LIST_BATCH ? rec : nullptr;
      uint32_t* zero2 = j == 0 ? &c->cnt[C_LREC] : nullptr;
      // the previous level's candidate count (its slot is untouched until the next level)
      const uint32_t* prev = j == 0 ? nullptr : &c->cnt[lcnt[(L + 3) % 4]];
      Prof p(c, R.s, K_PULL_LIST, 0.0);
      if (prof) *prof = p.idx;
      if (dir == 0)
        k_pull_list<0><<<lgrid, WG, 0, R.s>>>(lists[L % 2], &c->cnt[lcnt[L % 4]], vis, roff, radj, pid, q,
                                              &c->cnt[C_LREC], lists[(L + 1) % 2], &c->cnt[lcnt[(L + 1) % 4]],
                                              &c->cnt[lcnt[(L + 2) % 4]], zero2, esum, prev);
      else
        k_pull_list<1><<<lgrid, WG, 0, R.s>>>(lists[L % 2], &c->cnt[lcnt[L % 4]], vis, roff, radj, pid, q,
                                              &c->cnt[C_LREC], lists[(L + 1) % 2], &c->cnt[lcnt[(L + 1) % 4]],
                                              &c->cnt[lcnt[(L + 2) % 4]], zero2, esum, prev);
    }
  };
  // Speculation: on a large graph, the first push level of a direction
  // (from the pivot) is followed at once by a sweep and its list batch,
  // conditional on the device evaluating the host's direction rule
  // (SpecGate, in the level's publication), and by their publication into
  // map bank 1. The host reads the push level's publication meanwhile; if
  // its rule then chooses pull, the sweep is already running, else the
  // gated launches did nothing and the host takes the sequence number of
  // their publication back.
  const bool speculate = !pid && !c->profile && R.n >= (1u << 20);
  uint32_t spec_seq = 0;  // the speculative publication's sequence number, while unresolved
  // each level's publication zeroes the counters of the level after it
  if (!seeded) k_level_reset<<<1, 64, 0, R.s>>>(&c->cnt[nc], c->cnt64);
  bool scanned = seeded;  // the first level's degree scan is done
  bool first_level = true;
  while (f > 0) {
    // direction: Beamer's rule on the frontier's edge count, which the
    // previous level summed on the device for the vertices it claimed
    if (!pull && known && (double)edges * alpha > m_rest && f > 64) pull = true;
    const bool spec_taken = spec_seq != 0 && pull;
    if (spec_seq != 0 && !pull) {
      c->pub_seq = spec_seq - 1;  // the gated publication did not happen
      spec_seq = 0;
    }
    const bool narrow = !pull && known && R.narrow_len && maxd <= NARROW_DEG && edges <= NARROW_EDGES &&
                        !(R.o.flags & KSPN_FLAG_NO_COOP);
    if (narrow || (!pull && known && coop && edges <= R.coop_edges && maxd <= R.coop_deg)) {
      // the following push levels on the device: narrow chained launches,
      // chained launches or one persistent launch
      CoopOut co{};
      KSPN_TRY(run_coop(R, op, a, off, adj, F, N, fc, nc, edges, maxd, m_rest, alpha, &co, narrow));
      if (co.levels & 1) {
        std::swap(F, N);
        std::swap(fc, nc);
      }
      f = co.f;
      edges = co.edges;
      maxd = co.maxd;
      m_rest -= (double)co.expanded;
      if (R.st) {
        R.st->edges_scanned += co.expanded;
        R.st->fwbw_levels_coop += co.levels;
      }
      if (co.levels > 0) continue;
      // nothing expanded (not expected: the bounds were checked): one level on the host
    }
    if (!pull) {
      // tile width: the options' or, by default, the widest that still
      // gives every CU a few tiles (narrow tiles fill the device on
      // mid-sized frontiers, wide ones keep more loads in flight per thread)
      uint32_t ept = R.ept;
      if (!ept) {
        const uint64_t cu = c->num_cu;
        ept = !known || edges >= cu * WG * 8 * 4 ? 8u : (edges >= cu * WG * 4 * 2 ? 4u : 2u);
      }
      // a frontier of small degrees: chunks of entries, no degree scan
      const bool chunked = (R.o.flags & KSPN_FLAG_CHUNKED) && known && !scanned && maxd <= CHUNK_MAXD;
      const int kern = chunked ? (dir == 0 ? K_CHUNK_FW : K_CHUNK_BW) : (dir == 0 ? K_EXPAND_FW : K_EXPAND_BW);
      bool tiles = false;  // the scan recorded the tiles' frontier windows
      if (!scanned && !chunked) {
        const uint32_t shift = ept == 2 ? 9u : (ept == 4 ? 10u : (ept == 16 ? 12u : 11u));
        KSPN_TRY(frontier_scan(R, F, &c->cnt[fc], f, off, &tiles, shift));
      }
      scanned = false;
      uint32_t grid = known ? expand_grid(c, edges, ept) : (uint32_t)c->num_cu * 8;
      Prof p(c, R.s, kern, 20.0 * f + (4.0 + 0.125) * (double)edges);
      const int rop = dir == 0 ? OP_FW : (pid ? OP_PBW : OP_BW);
      if (chunked) {
        const uint32_t cgrid = (uint32_t)std::min<uint64_t>(((uint64_t)f + CHUNK_N - 1) / CHUNK_N,
                                                           (uint64_t)c->num_cu * 16);
        if (rop == OP_FW)
          k_expand_chunk<OP_FW><<<cgrid, WG, 0, R.s>>>(F, &c->cnt[fc], off, adj, a, N, &c->cnt[nc], &c->cnt[C_ERR],
                                                        c->scan + f, edges, nullptr);
        else if (rop == OP_PBW)
          k_expand_chunk<OP_PBW><<<cgrid, WG, 0, R.s>>>(F, &c->cnt[fc], off, adj, a, N, &c->cnt[nc], &c->cnt[C_ERR],
                                                        c->scan + f, edges, nullptr);
        else
          k_expand_chunk<OP_BW><<<cgrid, WG, 0, R.s>>>(F, &c->cnt[fc], off, adj, a, N, &c->cnt[nc], &c->cnt[C_ERR],
                                                        c->scan + f, edges, nullptr);
      } else {
        launch_reach_expand(rop, ept, grid, R.s, F, &c->cnt[fc], c->scan, off, adj, a, N, &c->cnt[nc],
                            tiles ? c->tstart : nullptr, c->tstart_cap);
      }
      if (fw_levels_push) (*fw_levels_push)++;
    } else {
      bool batch = list_mode;  // a batch of list levels follows (after a sweep: when the remainder is short)
      bool swept = false;
      int sweep_prof = -1, list_prof = -1;  // their profile entries, completed after the publication
      if (!list_mode) {
        // the sweep's claims go to cnt[nc]
        if (!spec_taken) launch_sweep(nc, nullptr, &sweep_prof);
        if (fw_levels_pull) (*fw_levels_pull)++;
        L = 0;
        batch = true;
        swept = true;
        surv_n = list_max;  // at most
      }
      const uint32_t list_in = surv_n;
      if (spec_taken) {
        L += LIST_BATCH;  // launched with the sweep
        KSPN_TRY(publish_collect(c, R.s, spec_seq, 1));
        spec_seq = 0;
      } else {
        if (batch) launch_batch(N, surv_n, &list_prof);
        KSPN_TRY(publish(c, R.s, nullptr, nullptr, 0, nullptr, &c->cnt[fc]));
      }
      // a sweep's list batch ran iff its remainder was short
      const uint64_t left = c->h_u64[2];
      if (swept && left > list_max) batch = false;
      list_level = batch;
      if (c->profile) {
        // algorithmic bytes: per scanned edge an id and a bitmap bit; per sweep candidate (claimed or
        // left) its two row offsets; per list entry its id, offsets, out entry
        if (sweep_prof >= 0)
          c->pending[sweep_prof].bytes += 4.125 * (double)c->h_u64[1] + 16.0 * ((double)c->h_cnt[nc] + (double)left);
        if (list_prof >= 0)
          c->pending[list_prof].bytes += 4.125 * (double)c->h_u64[3] + (batch ? 24.0 * (swept ? left : list_in) : 0.0);
      }
      if (batch) {
        if (fw_levels_pull) *fw_levels_pull += LIST_BATCH;
        surv_n = c->h_cnt[lcnt[L % 4]];
        list_mode = true;
      } else {
        surv_n = (uint32_t)left;
      }
    }
    if (!pull) {
      // the device evaluates the rule before the publication resets the
      // level's counters; the sweep and batch run behind it, with the
      // frontier buffers and counter slots as they will be after the swap
      // below (claims to cnt[fc], claimed list to F, reset of cnt[nc])
      const bool spec = speculate && first_level && !list_mode;
      SpecGate sg{};
      if (spec) sg = SpecGate{&c->cnt[C_SPEC], &c->cnt[nc], c->scan + f, m_rest, alpha, &c->cnt[C_SURV0]};
      const uint32_t seq0 = publish_launch(c, R.s, c->scan + f, nullptr, 0, nullptr, &c->cnt[fc], nullptr, 0, nullptr, sg);
      if (spec) {
        launch_sweep(fc, &c->cnt[C_SPEC], nullptr);
        L = 0;
        launch_batch(F, list_max, nullptr);
        spec_seq = publish_launch(c, R.s, nullptr, nullptr, 0, nullptr, &c->cnt[nc], nullptr, 1, &c->cnt[C_SPEC]);
      }
      KSPN_TRY(publish_collect(c, R.s, seq0));
      if (c->h_cnt[C_SERR]) {
        note_error("k_deg_scan", "frontier count or ticket out of step with the launch", __FILE__, __LINE__);
        return KSPN_E_HIP;
      }
    }
    first_level = false;
    if (!pull) {
      edges = c->h_u64[0];  // exact total of the level just expanded
      m_rest -= (double)edges;
      if (R.st) R.st->edges_scanned += edges;
      if (c->profile && !c->pending.empty()) c->pending.back().bytes = 20.0 * f + 4.125 * (double)edges;
    } else {
      if (R.st) R.st->edges_scanned += c->h_u64[1] + c->h_u64[3];  // the sweep's and the list levels'
    }
    if (R.live_pending) {
      R.trim_decided = c->h_cnt[C_DECIDED];
      R.first_pivot = c->h_cnt[C_FLAG];
      R.live_n = R.n - R.trim_decided;
      R.live_pending = false;
    }
    // a pull ending in list levels: what the last one claimed (in N)
    uint32_t nf = pull && list_level ? c->h_cnt[C_LREC] : c->h_cnt[nc];
    edges = c->h_u64[2];
    maxd = pull ? 0xFFFFFFFFu : (uint32_t)std::min<uint64_t>(c->h_u64[3], 0xFFFFFFFFull);
    // after a pull level the frontier's out-edges are unknown until a push
    // level sums them; the remaining-edge estimate keeps them until expanded
    known = !pull;
    if (pull && nf > 0 && (double)nf * beta < (double)R.live_n) {
      // back to push: the frontier is what the last pull level claimed. A
      // list level appended it to N (its count goes to cnt[nc]); a sweep left
      // it in the bitmap fnew (its size nf is in cnt[nc]; the sweep zeroed
      // the cursor)
      pull = false;
      list_mode = false;
      if (list_level) {
        k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[nc], &c->cnt[C_LREC]);
      } else {
        Prof p(c, R.s, K_LIVE_LIST, 4.0 * R.words + 4.0 * nf);
        k_bitmap_to_queue<<<b2q_grid(R.words), WG, 0, R.s>>>(c->fnew, R.words, N, &c->cnt[C_CURSOR]);
      }
    }
    std::swap(F, N);
    std::swap(fc, nc);
    f = nf;
  }
  if (spec_seq != 0) c->pub_seq = spec_seq - 1;  // nothing claimed: the gated publication did not happen
  return KSPN_OK;
}

// With publish = false the decided count stays on the device
// (cnt[C_DECIDED]) for the caller's next publication.
// With `tail`, the vertices left live are listed in qa (count cnt[C_LIVE],
// zero since the first trim: one such call per solve) and the tail solver's
// counters are zeroed, for tail_spec.
int fwbw(Run& R, uint32_t pivot, uint32_t* decided, bool publish = true, bool tail = false) {
  Ctx* c = R.c;
  // After the first trim of the plain single-pivot path, k_pivot_reduce
  // seeded both directions on the device (R.live_pending)
  const bool seeded = R.live_pending;
  uint32_t lp = 0, lq = 0;
  if (!seeded) {
    Prof p(c, R.s, K_BITMAP, 4.0 * R.words);
    k_seed<<<stream_grid(c, R.words), WG, 0, R.s>>>(c->fwv, R.words, pivot, nullptr, c->qa, &c->cnt[C_Q0]);
  }
  KSPN_TRY(reach(R, 0, 1, nullptr, &lp, &lq, seeded));
  if (seeded) {
    k_reach_start<<<1, 64, 0, R.s>>>(&c->cnt[C_FLAG], c->qa, &c->cnt[C_Q0], &c->cnt[C_Q1], c->cnt64, R.g->bw_off,
                                     c->scan);
  } else {
    Prof p(c, R.s, K_BITMAP, 4.0 * R.words);
    k_seed<<<stream_grid(c, R.words), WG, 0, R.s>>>(c->bwv, R.words, pivot, nullptr, c->qa, &c->cnt[C_Q0]);
  }
  KSPN_TRY(reach(R, 1, 1, nullptr, &lp, &lq, seeded));
  // the pivot's SCC: clear it from live (listing the rest for the tail
  // solver when it follows), then label it with its smallest member
  const uint32_t nparts = finish_grid(R.words);
  {
    Prof p(c, R.s, K_BITMAP, 12.0 * R.words);
    Zero3 z{};
    if (tail) z = Zero3{{&c->cnt[C_RES_N], &c->cnt[C_ROOM], nullptr}};
    k_fwbw_finish<<<nparts, WG, 0, R.s>>>(c->live, c->bwv, R.words, c->part, tail ? c->qa : nullptr,
                                          &c->cnt[C_LIVE], z);
  }
  if (tail) {
    // tail_spec runs it beside the tail's degree pass (k_label_prep)
    R.label_pending = true;
    R.label_nparts = nparts;
    R.label_dense = seeded;
  } else {
    Prof p(c, R.s, K_LABEL, seeded ? R.n / 4.0 + 4.0 * R.n : R.n / 8.0);
    k_label_bits<<<stream_grid(c, R.n / 8 + 1), WG, 0, R.s>>>(R.n, c->bwv, c->part, nparts, &c->cnt[C_MIN],
                                                              &c->cnt[C_DECIDED], R.labels, seeded ? c->live : nullptr);
  }
  if (R.st) {
    R.st->fwbw_levels_push += lp;
    R.st->fwbw_levels_pull += lq;
  }
  *decided = 0;
  if (!publish) return KSPN_OK;
  KSPN_TRY(read_counters(c, R.s, C_DECIDED + 1));
  *decided = c->h_cnt[C_DECIDED];
  R.live_n -= *decided;
  return KSPN_OK;
}

// ---------------------------------------------------------------- partitioned FW-BW rounds

// Pivot keys of the next round over the S sparse partition ids and the
// "used" flag of every id (pflags)
// the random rule's seed of a round (0: not the random rule)
inline uint64_t part_seed(const Run& R, uint32_t round) {
  return R.o.pivot == KSPN_PIVOT_RANDOM ? ((uint64_t)R.o.seed * 0x10001ull + round + 1) : 0;
}

void part_keys(Run& R, uint32_t S, uint32_t round) {
  Ctx* c = R.c;
  Prof p(c, R.s, K_PIVOT, R.n / 8.0 + 8.0 * R.live_n + 12.0 * S);
  k_fill_u64<<<stream_grid(c, S), WG, 0, R.s>>>((unsigned long long*)c->pbest, S, 0ull);
  k_part_key<<<stream_grid(c, R.n), WG, 0, R.s>>>(R.n, c->live, c->pid, c->vprio, part_seed(R, round),
                                                  (unsigned long long*)c->pbest);
  k_part_used<<<stream_grid(c, (uint64_t)S + 1), WG, 0, R.s>>>(S, (unsigned long long*)c->pbest, c->pflags);
}

// compact index of every used id (newid = c->pnew), their count into cnt[C_AUX]
int part_compact(Run& R, uint32_t S) {
  Ctx* c = R.c;
  KSPN_TRY(exclusive_scan_u32(c, c->pflags, c->pnew, (uint64_t)S + 1, R.s));
  k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_AUX], c->pnew + S);
  return KSPN_OK;
}

// live vertices scored for the learned rule: at least live / kScoreTopDiv, those of highest in*out
constexpr uint32_t kScoreTopDiv = 8;

int pfwbw_rounds(Run& R, uint32_t K, uint32_t np, uint32_t* decided_total) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  const uint64_t n = R.n;
  if (!c->pid) KSPN_TRY(dalloc(&c->pid, (uint64_t)c->cap_n + 64));
  if (!c->pbest) KSPN_TRY(dalloc(&c->pbest, 3ull * c->cap_n + 64));
  if (!c->pmin) KSPN_TRY(dalloc(&c->pmin, (uint64_t)c->cap_n + 64));
  if (!c->pflags) KSPN_TRY(dalloc(&c->pflags, 3ull * c->cap_n + 64));
  if (!c->pnew) KSPN_TRY(dalloc(&c->pnew, 3ull * c->cap_n + 64));
  // sparse ids of the first round: 0 for every vertex, or the preset ids
  // (the graphs of a batch; those the first trim emptied are not used)
  uint32_t S = R.pid_preset ? np : 1u;
  k_pid_init<<<stream_grid(c, n), WG, 0, R.s>>>(R.n, c->live, c->pid, R.pid_preset);
  // the reaches of the rounds test partition ids instead of the live bitmap
  R.dead_pid = n <= DEAD_PID_MAX_N;

  const float* score = nullptr;
  if (R.o.pivot == KSPN_PIVOT_LEARNED) {
    // MLP scores, fixed for the whole solve, of the live vertices with the
    // highest in*out (the only ones likely to be a partition's candidates)
    if (!c->vscore) KSPN_TRY(dalloc(&c->vscore, (uint64_t)c->cap_n + 64));
    if (!c->vfeat) KSPN_TRY(dalloc(&c->vfeat, 7ull * c->cap_n + 64));
    float* feat = c->vfeat;
    float* sc = c->vfeat + 6ull * c->cap_n;
    {
      Prof p(c, R.s, K_PIVOT, 0.0);
      uint32_t scored = 0;
      KSPN_TRY(mlp_score_top(c, g, R.o.mlp_params, R.live_n, kScoreTopDiv, c->ulist, &c->cnt[C_AUX], feat, sc,
                             c->vscore, R.s, &scored));
      // histogram and list passes over the vertices, features and scores of the scored ones
      p.add_bytes(36.0 * R.n + 64.0 * scored);
    }
    score = c->vscore;
  }
  // every vertex's pivot priority, once for the solve
  if (!c->vprio) KSPN_TRY(dalloc(&c->vprio, (uint64_t)c->cap_n + 64));
  {
    Prof p(c, R.s, K_PIVOT, 36.0 * n);
    k_part_prio<<<stream_grid(c, n), WG, 0, R.s>>>(R.n, g->fw_off, g->bw_off, score,
                                                   R.o.pivot != KSPN_PIVOT_PLAIN, c->vprio);
  }
  uint32_t* pcount = nullptr;
  if (R.o.pivot_log) {
    if (!c->pcnt) KSPN_TRY(dalloc(&c->pcnt, 4ull * c->cap_n + 64));
    pcount = c->pcnt;
  }
  // the first round's keys and compaction
  part_keys(R, S, 0);
  KSPN_TRY(part_compact(R, S));
  KSPN_TRY(publish(c, R.s, nullptr));
  np = c->h_cnt[C_AUX];
  uint32_t total = 0;
  for (uint32_t round = 0; round < K && R.live_n > 0 && np > 0; ++round) {
    if (R.o.residual_threshold && R.live_n <= R.o.residual_threshold) break;
    const uint32_t gridn = stream_grid(c, n);
    {
      Prof p(c, R.s, K_PIVOT, 8.0 * np);
      k_fill_u32<<<stream_grid(c, R.words), WG, 0, R.s>>>(c->fwv, R.words, 0u);
      k_fill_u32<<<stream_grid(c, R.words), WG, 0, R.s>>>(c->bwv, R.words, 0u);
      k_part_emit<<<stream_grid(c, S), WG, 0, R.s>>>(S, R.n, (unsigned long long*)c->pbest, c->pnew, c->qa, c->fwv,
                                                     c->bwv, &c->cnt[C_ERR]);
      k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_Q0], np);
    }
    uint32_t lp = 0, lq = 0;
    KSPN_TRY(reach(R, 0, np, c->pid, &lp, &lq));
    {
      Prof p(c, R.s, K_PIVOT, 8.0 * np);
      k_part_emit<<<stream_grid(c, S), WG, 0, R.s>>>(S, R.n, (unsigned long long*)c->pbest, c->pnew, c->qa, c->fwv,
                                                     c->bwv, &c->cnt[C_ERR]);
      k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_Q0], np);
    }
    KSPN_TRY(reach(R, 1, np, c->pid, &lp, &lq));
    const uint32_t S_next = 3u * np;
    const bool more = round + 1 < K;
    if (more && score && R.o.learned_rounds && round + 1 == R.o.learned_rounds) {
      // the next round's keys (formed below) take the plain rule's priorities
      Prof p(c, R.s, K_PIVOT, 36.0 * n);
      k_part_prio<<<stream_grid(c, n), WG, 0, R.s>>>(R.n, g->fw_off, g->bw_off, nullptr, false, c->vprio);
    }
    if (!pcount) {
      // classification, relabelling and the next round's keys in one pass,
      // then the labels of the decided vertices from their list
      {
        Prof p(c, R.s, K_BITMAP, 0.5 * R.n + 16.0 * R.live_n + 12.0 * np + 8.0 * S_next);
        k_fill_u32<<<stream_grid(c, np), WG, 0, R.s>>>(c->pmin, np, 0xFFFFFFFFu);
        if (more) k_fill_u64<<<stream_grid(c, S_next), WG, 0, R.s>>>((unsigned long long*)c->pbest, S_next, 0ull);
        k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_DLIST], 0u);
        // the decided list goes to qb, free between the backward reach and the next round's
        k_part_split<<<gridn, WG, 0, R.s>>>(R.n, c->live, c->fwv, c->bwv, c->pid, c->pnew, c->pmin, c->vprio,
                                            part_seed(R, round + 1), more, (unsigned long long*)c->pbest, c->part,
                                            c->qb, &c->cnt[C_DLIST]);
      }
      {
        Prof p(c, R.s, K_LABEL, 16.0 * np);
        k_part_label<<<gridn, WG, 0, R.s>>>(c->qb, &c->cnt[C_DLIST], c->pid, c->pnew, c->pmin, R.labels);
        if (more)
          k_part_used<<<stream_grid(c, (uint64_t)S_next + 1), WG, 0, R.s>>>(S_next, (unsigned long long*)c->pbest,
                                                                             c->pflags);
      }
    } else {
      Prof p(c, R.s, K_BITMAP, 0.375 * R.n + 12.0 * np);
      k_fill_u32<<<stream_grid(c, np), WG, 0, R.s>>>(c->pmin, np, 0xFFFFFFFFu);
      k_fill_u32<<<stream_grid(c, 4ull * np), WG, 0, R.s>>>(pcount, 4ull * np, 0u);
      k_part_classify<<<gridn, WG, 0, R.s>>>(R.n, c->live, c->fwv, c->bwv, c->pid, c->pnew, c->pmin, pcount);
      if (pcount && R.log_count < R.o.pivot_log_cap) {
        k_part_log<<<stream_grid(c, S), WG, 0, R.s>>>(S, (unsigned long long*)c->pbest, c->pnew, pcount, round,
                                                      R.o.pivot_log, R.log_count, R.o.pivot_log_cap);
        R.log_count = std::min<uint64_t>((uint64_t)R.log_count + np, R.o.pivot_log_cap);
        if (R.st) R.st->pivot_log_count = R.log_count;
      }
    }
    if (pcount) {
      {
        Prof p(c, R.s, K_LABEL, 16.0 * R.live_n);
        k_part_apply<<<gridn, WG, 0, R.s>>>(R.n, c->live, c->fwv, c->bwv, c->pid, c->pmin, c->pnew, R.labels, c->part);
      }
      if (more) part_keys(R, S_next, round + 1);
    }
    // the next round's compaction rides on this round's publication
    k_reduce3<<<1, WG, 0, R.s>>>(c->part, gridn, &c->cnt[C_DECIDED], c->cnt64);
    if (more) KSPN_TRY(part_compact(R, S_next));
    KSPN_TRY(publish(c, R.s, nullptr));
    if (c->h_cnt[C_ERR]) return KSPN_E_INVALID;  // a pivot key that names no vertex
    const uint32_t d = c->h_cnt[C_DECIDED];
    R.live_n -= d;
    total += d;
    if (R.st) {
      R.st->fwbw_levels_push += lp;
      R.st->fwbw_levels_pull += lq;
      R.st->fwbw_rounds_done++;
      if (np > R.st->partitions_max) R.st->partitions_max = np;
    }
    if (!more || d == 0) break;
    S = S_next;
    np = c->h_cnt[C_AUX];
  }
  *decided_total = total;
  R.dead_pid = false;
  return KSPN_OK;
}

// ---------------------------------------------------------------- trims
// exhaustive trim-1 to a fixpoint: sweep a candidate queue, then the live
// neighbours (both directions) of what was decided become the candidates
int trim_exhaustive(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  uint32_t q = 0;
  KSPN_TRY(build_live_list(c, R.n, c->live, c->qa, &q, R.s));
  k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_Q0], &c->cnt[C_AUX]);
  uint32_t* Q = c->qa;
  uint32_t* T = c->qb;
  uint32_t* NQ = c->qc;
  bool first = true;
  uint32_t total = 0;
  OpArgs a{c->live, c->fwv, c->bwv, c->inq, c->col, R.labels, 0};
  while (q > 0) {
    KSPN_HIP(hipMemsetAsync(&c->cnt[C_Q1], 0, sizeof(uint32_t), R.s));
    {
      Prof p(c, R.s, K_TRIM_SWEEP, 4.0 * q + 16.0 * q);
      k_trim_sweep<TRIM_ONE><<<stream_grid(c, q), WG, 0, R.s>>>(Q, &c->cnt[C_Q0], c->live, first ? nullptr : c->inq,
                                                                g->fw_off, g->fw_adj, g->bw_off, g->bw_adj, R.labels, T,
                                                                &c->cnt[C_Q1]);
    }
    first = false;
    KSPN_TRY(read_counters(c, R.s, 2));
    uint32_t t = c->h_cnt[C_Q1];
    total += t;
    if (t == 0) break;
    // candidates: live neighbours of the trimmed vertices, deduplicated by inq
    KSPN_HIP(hipMemsetAsync(&c->cnt[C_Q0], 0, sizeof(uint32_t), R.s));
    uint64_t e1 = 0, e2 = 0;
    KSPN_TRY(expand<OP_NEIGH>(R, T, &c->cnt[C_Q1], t, g->fw_off, g->fw_adj, a, NQ, &c->cnt[C_Q0], K_EXPAND_NEIGH, 0.25, &e1));
    KSPN_TRY(expand<OP_NEIGH>(R, T, &c->cnt[C_Q1], t, g->bw_off, g->bw_adj, a, NQ, &c->cnt[C_Q0], K_EXPAND_NEIGH, 0.25, &e2));
    KSPN_TRY(read_counters(c, R.s, 1));
    q = c->h_cnt[C_Q0];
    std::swap(Q, NQ);
  }
  // leave no inq bit set: clear the last candidate queue
  if (q > 0) {
    k_bm_clear_list<<<stream_grid(c, q), WG, 0, R.s>>>(c->inq, Q, &c->cnt[C_Q0]);
  }
  *decided = total;
  R.live_n -= total;
  return KSPN_OK;
}

// one trim_1_normal-style sweep over the live list
int trim_normal(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  // candidates: every live vertex, in any order (count known on the host)
  const uint32_t q = R.live_n;
  k_zero2<<<1, 64, 0, R.s>>>(&c->cnt[C_Q0], &c->cnt[C_Q1]);
  {
    Prof p(c, R.s, K_LIVE_LIST, 4.0 * R.words + 4.0 * q);
    k_bitmap_to_queue<<<b2q_grid(R.words), WG, 0, R.s>>>(c->live, R.words, c->qa, &c->cnt[C_Q0]);
  }
  if (q) {
    Prof p(c, R.s, K_TRIM_SWEEP, 20.0 * q);
    k_trim_sweep<TRIM_NORMAL><<<stream_grid(c, q), WG, 0, R.s>>>(c->qa, &c->cnt[C_Q0], c->live, nullptr, g->fw_off,
                                                                 g->fw_adj, g->bw_off, g->bw_adj, R.labels, c->qb,
                                                                 &c->cnt[C_Q1]);
  }
  KSPN_TRY(read_counters(c, R.s, 2));
  *decided = c->h_cnt[C_Q1];
  R.live_n -= *decided;
  return KSPN_OK;
}

// trim-2 over all live vertices, in-direction launch then out-direction
int trim_two(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  uint32_t q = 0;
  KSPN_TRY(build_live_list(c, R.n, c->live, c->qa, &q, R.s));
  if (q == 0) {
    *decided = 0;
    return KSPN_OK;
  }
  k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_Q0], &c->cnt[C_AUX]);
  KSPN_HIP(hipMemsetAsync(&c->cnt[C_Q1], 0, sizeof(uint32_t), R.s));
  {
    Prof p(c, R.s, K_TRIM2, 40.0 * q);
    k_trim2<0><<<stream_grid(c, q), WG, 0, R.s>>>(c->qa, &c->cnt[C_Q0], c->live, g->bw_off, g->bw_adj, R.labels, c->qb,
                                                  &c->cnt[C_Q1]);
  }
  {
    Prof p(c, R.s, K_TRIM2, 40.0 * q);
    k_trim2<1><<<stream_grid(c, q), WG, 0, R.s>>>(c->qa, &c->cnt[C_Q0], c->live, g->fw_off, g->fw_adj, R.labels, c->qb,
                                                  &c->cnt[C_Q1]);
  }
  KSPN_TRY(read_counters(c, R.s, 2));
  *decided = c->h_cnt[C_Q1];
  R.live_n -= *decided;
  return KSPN_OK;
}

// ---------------------------------------------------------------- colour
// One round: keys ~rotr(v, r) on live vertices, forward max-propagation to
// a fixpoint (pull sweep, then push from changed vertices), roots where the
// colour is the own key, backward closure over in-edges within a colour.
int color_round(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  uint32_t r = ++R.rotation;
  // keys on live vertices, 0 elsewhere: a full pass the first time, then
  // only over the previous round's live list (everything else is already 0)
  const uint32_t* pid = R.pid_active ? c->pid : nullptr;
  if (!c->cpw) {
    KSPN_TRY(dalloc(&c->cpw, (uint64_t)c->cap_n + 64));
    R.col_valid = false;
  }
  if (!R.col_valid) {
    Prof p(c, R.s, K_COLOR_INIT, 8.0 * R.n + R.n / 8.0);
    k_color_init_all<<<stream_grid(c, R.n), WG, 0, R.s>>>(R.n, c->live, pid, c->cpw, r);
  } else if (R.prev_u) {
    Prof p(c, R.s, K_COLOR_INIT, 16.0 * R.prev_u);
    k_color_init_list<<<stream_grid(c, R.prev_u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], c->live, pid, c->cpw, r);
  }
  uint32_t u = 0;
  KSPN_TRY(build_live_list(c, R.n, c->live, c->ulist, &u, R.s));
  k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_RES_N], &c->cnt[C_AUX]);
  R.prev_u = u;
  OpArgs a{c->live, c->fwv, c->bwv, c->inq, c->col, R.labels, r, nullptr, nullptr, pid, nullptr, 0u, c->cpw};
  // chunked push levels publish no edge count: they sum it into cnt64[3]
  const bool chunk_stats = R.st && (R.o.flags & KSPN_FLAG_CHUNKED);
  if (chunk_stats || c->profile) k_fill_u64<<<1, 64, 0, R.s>>>(c->cnt64 + 3, 1, 0ull);
  // forward: pull sweep over all live vertices
  KSPN_HIP(hipMemsetAsync(&c->cnt[C_Q0], 0, sizeof(uint32_t), R.s));
  {
    Prof p(c, R.s, K_COLOR_PULL, 24.0 * u);
    k_color_pull<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->bw_off, g->bw_adj, c->cpw, c->qa,
                                                    &c->cnt[C_Q0], R.prune ? c->prune_bw : nullptr);
  }
  KSPN_TRY(read_counters(c, R.s, 1));
  uint32_t f = c->h_cnt[C_Q0];
  uint32_t* F = c->qa;
  uint32_t* N = c->qb;
  int fc = C_Q0, nc = C_Q1;
  uint32_t iters = 1;
  // Chunked push levels alternate between the queue bitmaps inq and inq2:
  // a level queues into one and clears its frontier's bits in the other as
  // it reads the frontier, so no clearing launch is needed; inq_f is the
  // bitmap holding F's bits (nullptr after a pull sweep, whose output sets
  // none). Otherwise F's bits are cleared before each level.
  const bool chunk = R.o.flags & KSPN_FLAG_CHUNKED;
  uint32_t* inq_f = nullptr;
  if (chunk) k_fill_u32<<<stream_grid(c, R.words), WG, 0, R.s>>>(c->inq2, R.words, 0u);
  while (f > 0) {
    const bool sweep = (double)f * 2.0 > (double)u;
    // drop F's queue bits so its members can be queued again
    if (!chunk) k_bm_clear_list<<<stream_grid(c, f), WG, 0, R.s>>>(c->inq, F, &c->cnt[fc]);
    else if (sweep && inq_f) k_bm_clear_list<<<stream_grid(c, f), WG, 0, R.s>>>(inq_f, F, &c->cnt[fc]);
    KSPN_HIP(hipMemsetAsync(&c->cnt[nc], 0, sizeof(uint32_t), R.s));
    if (sweep) {
      // large frontier: another pull sweep over every live vertex (half the
      // live list was faster than a quarter or all of it on planted 5e7)
      inq_f = nullptr;
      Prof p(c, R.s, K_COLOR_PULL, 24.0 * u);
      k_color_pull<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->bw_off, g->bw_adj, c->cpw, N,
                                                      &c->cnt[nc], R.prune ? c->prune_bw : nullptr);
    } else if (chunk) {
      // up to kColorChain push levels chained on the device, one round trip
      ColorChain* ch = reinterpret_cast<ColorChain*>(c->cchain);
      uint32_t* inq0 = inq_f == c->inq ? c->inq2 : c->inq;  // level 0 queues into the bitmap without F's bits
      uint32_t* inq1 = inq0 == c->inq ? c->inq2 : c->inq;
      OpArgs af = a;
      if (R.prune) af.prune = c->prune_fw;
      unsigned long long* esum = R.st || c->profile ? c->cnt64 + 3 : nullptr;
      unsigned long long e_before = 0, e_after = 0;
      if (c->profile) {
        KSPN_HIP(hipMemcpyAsync(&e_before, c->cnt64 + 3, sizeof(e_before), hipMemcpyDeviceToHost, R.s));
        KSPN_HIP(hipStreamSynchronize(R.s));
      }
      k_color_chain_init<<<1, 64, 0, R.s>>>(ch, &c->cnt[fc]);
      {
        Prof p(c, R.s, K_EXPAND_COLOR, 20.0 * f);
        const uint32_t cgrid = (uint32_t)c->num_cu * 8;
        for (uint32_t L = 0; L < kColorChain; ++L)
          k_color_chain<<<cgrid, WG, 0, R.s>>>(F, N, g->fw_off, g->fw_adj, af, inq0, inq1, inq_f != nullptr, ch, u, L,
                                               &c->cnt[C_ERR], esum);
      }
      if (c->profile) {
        KSPN_HIP(hipMemcpyAsync(&e_after, c->cnt64 + 3, sizeof(e_after), hipMemcpyDeviceToHost, R.s));
        KSPN_HIP(hipStreamSynchronize(R.s));
        if (!c->pending.empty()) c->pending.back().bytes = 20.0 * f + 12.0 * (double)(e_after - e_before);
      }
      k_color_chain_end<<<1, 64, 0, R.s>>>(ch, kColorChain, &c->cnt[fc], &c->cnt[C_AUX]);
      KSPN_TRY(read_counters(c, R.s, 2));
      const uint32_t E = c->h_cnt[C_AUX];
      if (E == 0 || E > kColorChain) {
        note_error("k_color_chain", "a chain of colour levels ran no level", __FILE__, __LINE__);
        return KSPN_E_HIP;
      }
      f = c->h_cnt[fc];
      if (E & 1) std::swap(F, N);  // the frontier of level E and its count (now in cnt[fc])
      inq_f = ((E - 1) & 1) ? inq1 : inq0;
      iters += E;
      continue;
    } else {
      uint64_t e = 0;
      OpArgs af = a;
      if (R.prune) af.prune = c->prune_fw;
      KSPN_TRY(expand<OP_COLOR>(R, F, &c->cnt[fc], f, g->fw_off, g->fw_adj, af, N, &c->cnt[nc], K_EXPAND_COLOR, 8.0, &e));
    }
    KSPN_TRY(read_counters(c, R.s, 2));
    f = c->h_cnt[nc];
    std::swap(F, N);
    std::swap(fc, nc);
    ++iters;
  }
  // queue bits of the last (empty) frontier are clear; pull-sweep outputs
  // never set inq, push outputs were cleared before each expansion
  // roots start the decided list D (qc); each closure level appends to D
  // and the next level expands the segment just appended
  uint32_t* D = c->qc;
  KSPN_HIP(hipMemsetAsync(&c->cnt[C_DTOT], 0, sizeof(uint32_t), R.s));
  {
    Prof p(c, R.s, K_COLOR_ROOTS, 12.0 * u);
    k_color_roots<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], c->live, c->cpw, R.labels, r, D,
                                                     &c->cnt[C_DTOT]);
  }
  KSPN_TRY(read_counters(c, R.s, C_DTOT + 1));
  const uint32_t roots = c->h_cnt[C_DTOT];
  uint32_t seg_beg = 0, seg_end = roots;
  if (R.prune) a.prune = c->prune_bw;
  while (seg_end > seg_beg) {
    uint32_t fseg = seg_end - seg_beg;
    k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_FSEG], fseg);
    uint64_t e = 0;
    KSPN_TRY(expand<OP_COLOR_BW>(R, D + seg_beg, &c->cnt[C_FSEG], fseg, g->bw_off, g->bw_adj, a, D, &c->cnt[C_DTOT],
                                 K_EXPAND_COLOR_BW, 8.0, &e));
    KSPN_TRY(read_counters(c, R.s, C_DTOT + 1));
    seg_beg = seg_end;
    seg_end = c->h_cnt[C_DTOT];
  }
  const uint32_t total = seg_end;
  if (chunk_stats) {
    unsigned long long e = 0;
    KSPN_HIP(hipMemcpyAsync(&e, c->cnt64 + 3, sizeof(e), hipMemcpyDeviceToHost, R.s));
    KSPN_HIP(hipStreamSynchronize(R.s));
    R.st->edges_scanned += e;
  }
  if (pid) {
    Prof p(c, R.s, K_COLOR_INIT, 12.0 * u);
    k_refine_pid<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], c->live, c->cpw, r, c->pid);
  }
  // canonical ids: every decision of this round carries its root's id;
  // replace it by the smallest member (col[root] is the min table)
  if (total) {
    Prof p(c, R.s, K_CANON, 4.0 * roots + 16.0 * total);
    k_fill_list<<<stream_grid(c, roots), WG, 0, R.s>>>(D, roots, c->col);
    k_canon_min<<<stream_grid(c, total), WG, 0, R.s>>>(R.n, D, total, R.labels, c->col);
    k_canon_apply<<<stream_grid(c, total), WG, 0, R.s>>>(R.n, D, total, R.labels, c->col);
  }
  *decided = total;
  R.live_n -= total;
  R.col_valid = true;  // vertices decided from now on are in ulist and get re-keyed
  if (R.st) {
    R.st->color_rounds++;
    R.st->color_iters += iters;
  }
  return KSPN_OK;
}

// ---------------------------------------------------------------- residual
int residual(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  uint32_t u = 0;
  KSPN_TRY(build_live_list(c, R.n, c->live, c->ulist, &u, R.s));
  if (u == 0) {
    *decided = 0;
    return KSPN_OK;
  }
  k_copy_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_RES_N], &c->cnt[C_AUX]);
  uint32_t* subid = c->col;
  {
    Prof p(c, R.s, K_RESIDUAL, 8.0 * u);
    k_sub_ids<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], subid);
    k_sub_deg<<<stream_grid(c, (uint64_t)u + 1), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->fw_off, g->fw_adj, c->live,
                                                               nullptr, c->deg);
  }
  KSPN_TRY(exclusive_scan_u64(c, c->deg, c->scan, (uint64_t)u + 1, R.s));
  KSPN_TRY(publish(c, R.s, c->scan + u));
  uint64_t sm = c->h_u64[0];
  if (sm + 1 > c->cap_sub_m) {
    KSPN_TRY(dalloc(&c->sub_adj, sm + 1));
    c->cap_sub_m = sm + 1;
  }
  if (!c->tj) KSPN_TRY(dalloc(&c->tj, 6ull * c->cap_n + 64));
  {
    Prof p(c, R.s, K_RESIDUAL, 16.0 * u + 8.0 * sm);
    k_sub_fill<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->fw_off, g->fw_adj, c->live, nullptr,
                                                  subid, c->scan, c->sub_adj);
  }
  uint32_t* sublab = c->qb;
  {
    Prof p(c, R.s, K_TARJAN, 24.0 * u + 4.0 * sm);
    k_tarjan<<<1, 64, 0, R.s>>>(u, c->scan, c->sub_adj, c->tj, sublab);
  }
  k_res_apply<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], sublab, R.labels);
  KSPN_HIP(hipMemsetAsync(c->live, 0, R.words * sizeof(uint32_t), R.s));
  if (R.st) {
    R.st->residual_n = u;
    R.st->residual_m = sm;
  }
  *decided = u;
  R.live_n -= u;
  R.col_valid = false;  // col held sub ids
  return KSPN_OK;
}

// ---------------------------------------------------------------- tail
// compact the live subgraph (both directions) and colour it in one workgroup
// the sub-CSR's adjacency and the solver, for u live vertices whose compact
// ids, offsets (scan, sub_off) and live list (ulist) are built; publishes
// A tail kernel met a listed id >= n (cnt[C_BADID], published): the lists
// that feed the tail are broken, so the solve stops with an error instead
// of indexing with the id. The flag is cleared for the next solve.
int tail_check_ids(Run& R) {
  if (!R.c->h_cnt[C_BADID]) return KSPN_OK;
  k_set_u32<<<1, 64, 0, R.s>>>(&R.c->cnt[C_BADID], 0u);
  note_error("tail", "a listed vertex id is not below n", __FILE__, __LINE__);
  return KSPN_E_INVALID;
}

int tail_finish(Run& R, uint32_t u) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  uint32_t* subid = c->col;
  {
    Prof p(c, R.s, K_RESIDUAL, 32.0 * u);
    k_tail_fill<<<stream_grid(c, (uint64_t)u * WAVE), WG, 0, R.s>>>(R.n, c->ulist, &c->cnt[C_RES_N], g->fw_off,
                                                                   g->fw_adj, g->bw_off, g->bw_adj, c->live, subid,
                                                                   c->scan, c->sub_off, c->sub_adj, c->cap_sub_m,
                                                                   &c->cnt[C_ERR], &c->cnt[C_ROOM], &c->cnt[C_BADID]);
  }
  uint32_t cap = R.o.max_rounds ? R.o.max_rounds - R.rounds : 0;  // > 0: checked by the caller
  {
    Prof p(c, R.s, K_TAIL, 24.0 * u);
    k_tail_solve<<<1, 1024, 0, R.s>>>(&c->cnt[C_RES_N], c->ulist, subid, c->scan, c->sub_off, c->sub_adj, c->cap_sub_m,
                                      c->qa, c->qb, c->qc, R.labels, R.rotation, cap, &c->cnt[C_FLAG], c->live,
                                      &c->cnt[C_ERR], &c->cnt[C_ROOM]);
  }
  return publish(c, R.s, nullptr);
}

int tail_solve(Run& R, uint32_t* decided) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  const uint32_t u = R.live_n;  // the live vertices, in any order
  if (u == 0) {
    *decided = 0;
    return KSPN_OK;
  }
  uint32_t* subid = c->col;
  if (!c->sub_off) KSPN_TRY(dalloc(&c->sub_off, (uint64_t)c->cap_n + 64));
  k_zero3<<<1, 64, 0, R.s>>>(&c->cnt[C_RES_N], &c->cnt[C_ISO], &c->cnt[C_ROOM]);
  {
    Prof p(c, R.s, K_LIVE_LIST, 4.0 * R.words + 4.0 * u);
    k_bitmap_to_queue<<<b2q_grid(R.words), WG, 0, R.s>>>(c->live, R.words, c->ulist, &c->cnt[C_RES_N]);
  }
  if (u < 4096) {
    // compact ids, degrees, both scans and the solver's flag: two launches
    Prof p(c, R.s, K_RESIDUAL, 24.0 * u);
    k_tail_prep<<<stream_grid(c, (uint64_t)u * WAVE), WG, 0, R.s>>>(R.n, c->ulist, &c->cnt[C_RES_N], 4095u,
                                                                   g->fw_off, g->fw_adj, g->bw_off, g->bw_adj,
                                                                   c->live, subid, c->scan, c->sub_off,
                                                                   &c->cnt[C_ISO], &c->cnt[C_BADID]);
    k_tail_scan<<<1, 1024, 0, R.s>>>(&c->cnt[C_RES_N], 4095u, c->scan, c->sub_off, &c->cnt[C_FLAG], &c->cnt[C_ERR]);
  } else {
    {
      Prof p(c, R.s, K_RESIDUAL, 8.0 * u);
      k_sub_ids<<<stream_grid(c, u), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], subid);
      k_sub_deg<<<stream_grid(c, (uint64_t)u + 1), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->fw_off, g->fw_adj,
                                                                 c->live, nullptr, c->deg);
    }
    KSPN_TRY(exclusive_scan_u64(c, c->deg, c->scan, (uint64_t)u + 1, R.s));
    {
      Prof p(c, R.s, K_RESIDUAL, 8.0 * u);
      k_sub_deg<<<stream_grid(c, (uint64_t)u + 1), WG, 0, R.s>>>(c->ulist, &c->cnt[C_RES_N], g->bw_off, g->bw_adj,
                                                                 c->live, nullptr, c->deg);
    }
    KSPN_TRY(exclusive_scan_u64(c, c->deg, c->sub_off, (uint64_t)u + 1, R.s));
    k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_FLAG], 0u);
  }
  KSPN_TRY(publish(c, R.s, c->scan + u));
  KSPN_TRY(tail_check_ids(R));
  // the compaction's own count (the host's live count should agree)
  if (c->h_cnt[C_ERR] || c->h_cnt[C_RES_N] != u) {
    note_error("tail_solve", "live count differs from the live bitmap", __FILE__, __LINE__);
    return KSPN_E_INVALID;
  }
  uint64_t sm = c->h_u64[0];  // the two directions have the same edge count
  if (2 * sm + 2 > c->cap_sub_m) {
    KSPN_TRY(dalloc(&c->sub_adj, 2 * sm + 2));
    c->cap_sub_m = 2 * sm + 2;
  }
  KSPN_TRY(tail_finish(R, u));
  KSPN_TRY(tail_check_ids(R));
  if (c->h_cnt[C_FLAG]) return KSPN_E_ITER_CAP;
  if (R.st) {
    R.st->residual_n = u;
    R.st->residual_m = sm;
  }
  *decided = u;
  R.live_n -= u;
  R.col_valid = false;  // col held sub ids
  return KSPN_OK;
}

// trim_1_normal and a tail right after the last FW-BW round, launched
// before the host knows how many vertices are left: the live lists, the
// sub-CSR and the solver run with every size read on the device, and one
// publication reports the FW-BW's decisions (cnt[C_DECIDED]), the trim's
// (*iso) and the tail's (*decided). When more than `limit` vertices are
// left after the trim the tail kernels do nothing (*done = false) and the
// caller goes on with colour rounds; a sub-CSR larger than its buffer is
// built again once the buffer has grown.
int tail_spec(Run& R, uint32_t limit, uint32_t* fwbw_decided, uint32_t* decided, uint32_t* iso, bool* done) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  *done = false;
  *decided = *iso = 0;
  uint32_t* subid = c->col;
  int prof_split = -1, prof_prep = -1, prof_fill = -1, prof_tail = -1;
  if (!c->sub_off) KSPN_TRY(dalloc(&c->sub_off, (uint64_t)c->cap_n + 64));
  if (c->cap_sub_m < 2) {
    KSPN_TRY(dalloc(&c->sub_adj, 1u << 16));
    c->cap_sub_m = 1u << 16;
  }
  {
    // every live vertex (qa, listed by fwbw's k_fwbw_finish, which also
    // zeroed the counters below), then trim_1_normal keeps those with a live
    // neighbour (ulist)
    // 20 bytes per listed vertex (id, two offset pairs); set once the list's length is known
    Prof p(c, R.s, K_TRIM_SWEEP, 0.0);
    prof_split = p.idx;
    k_tail_split<<<(uint32_t)c->num_cu * 8, WG, 0, R.s>>>(R.n, c->qa, &c->cnt[C_LIVE], g->fw_off, g->fw_adj,
                                                          g->bw_off, g->bw_adj, c->live, R.labels, c->ulist,
                                                          &c->cnt[C_RES_N], &c->cnt[C_BADID]);
  }
  {
    if (!R.label_pending) {
      note_error("tail_spec", "called without fwbw's deferred label pass", __FILE__, __LINE__);
      return KSPN_E_INVALID;
    }
    R.label_pending = false;
    // fwbw's label pass (its SCC's labels, and with label_dense every other
    // label) in the same launch as the tail's degree pass
    const uint32_t lblocks = stream_grid(c, R.n / 8 + 1);
    const uint32_t pblocks = stream_grid(c, (uint64_t)limit * WAVE);
    Prof p(c, R.s, K_LABEL, R.label_dense ? R.n / 4.0 + 4.0 * R.n : R.n / 8.0);
    prof_prep = p.idx;  // + 24 bytes per tail vertex, added below
    k_label_prep<<<lblocks + pblocks, WG, 0, R.s>>>(lblocks, R.n, c->bwv, c->part, R.label_nparts, &c->cnt[C_MIN],
                                                    &c->cnt[C_DECIDED], R.labels, R.label_dense ? c->live : nullptr,
                                                    c->ulist, &c->cnt[C_RES_N], limit, g->fw_off, g->fw_adj,
                                                    g->bw_off, g->bw_adj, c->live, subid, c->scan, c->sub_off,
                                                    &c->cnt[C_BADID]);
  }
  {
    Prof p(c, R.s, K_RESIDUAL, 0.0);
    prof_fill = p.idx;
    k_tail_scan<<<1, 1024, 0, R.s>>>(&c->cnt[C_RES_N], limit, c->scan, c->sub_off, &c->cnt[C_FLAG], &c->cnt[C_ERR]);
    k_tail_fill<<<stream_grid(c, (uint64_t)limit * WAVE), WG, 0, R.s>>>(R.n, c->ulist, &c->cnt[C_RES_N], g->fw_off,
                                                                       g->fw_adj, g->bw_off, g->bw_adj, c->live,
                                                                       subid, c->scan, c->sub_off, c->sub_adj,
                                                                       c->cap_sub_m, &c->cnt[C_ERR], &c->cnt[C_ROOM],
                                                                       &c->cnt[C_BADID]);
  }
  const uint32_t cap = R.o.max_rounds ? R.o.max_rounds - R.rounds : 0;
  {
    Prof p(c, R.s, K_TAIL, 0.0);
    prof_tail = p.idx;
    k_tail_solve<<<1, 1024, 0, R.s>>>(&c->cnt[C_RES_N], c->ulist, subid, c->scan, c->sub_off, c->sub_adj, c->cap_sub_m,
                                      c->qa, c->qb, c->qc, R.labels, R.rotation, cap, &c->cnt[C_FLAG], c->live,
                                      &c->cnt[C_ERR], &c->cnt[C_ROOM]);
  }
  KSPN_TRY(read_counters(c, R.s, C_COUNT_));
  KSPN_TRY(tail_check_ids(R));
  *fwbw_decided = c->h_cnt[C_DECIDED];
  R.live_n -= *fwbw_decided;
  *iso = c->h_cnt[C_LIVE] - c->h_cnt[C_RES_N];  // the split decides every listed vertex it does not keep
  R.live_n -= *iso;
  const uint32_t u = c->h_cnt[C_RES_N];
  if (c->profile) {
    // algorithmic bytes once the sizes are known: the split reads each listed vertex's id and offsets;
    // the degree pass each tail vertex's id and offsets and writes its sub id; the scans read and write
    // two arrays of u+1 offsets and the fill reads them again; the solver reads each vertex's state
    const double listed = c->h_cnt[C_LIVE], tail = c->h_cnt[C_ERR] ? 0.0 : (double)u;
    if (prof_split >= 0) c->pending[prof_split].bytes += 20.0 * listed;
    if (prof_prep >= 0) c->pending[prof_prep].bytes += 24.0 * tail;
    if (prof_fill >= 0) c->pending[prof_fill].bytes += 48.0 * (tail + 1.0);
    if (prof_tail >= 0) c->pending[prof_tail].bytes += 24.0 * tail;
  }
  if (c->h_cnt[C_ERR]) {
    // more than `limit` left after trim_1_normal: nothing else was solved
    k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_ERR], 0u);
    return KSPN_OK;
  }
  if (u != R.live_n) {
    note_error("tail_spec", "live count differs from the live bitmap", __FILE__, __LINE__);
    return KSPN_E_INVALID;
  }
  if (c->h_cnt[C_ROOM]) {
    // grow the sub-CSR's buffer, then build and solve it again
    KSPN_TRY(publish(c, R.s, c->scan + u));
    const uint64_t sm = c->h_u64[0];
    KSPN_TRY(dalloc(&c->sub_adj, 2 * sm + 2));
    c->cap_sub_m = 2 * sm + 2;
    k_set_u32<<<1, 64, 0, R.s>>>(&c->cnt[C_ROOM], 0u);
    KSPN_TRY(tail_finish(R, u));
    KSPN_TRY(tail_check_ids(R));
  }
  if (c->h_cnt[C_FLAG]) return KSPN_E_ITER_CAP;
  if (R.st) R.st->residual_n = u;
  *decided = u;
  R.live_n -= u;
  R.col_valid = false;  // col held sub ids
  *done = true;
  return KSPN_OK;
}

// ---------------------------------------------------------------- canonical ids
int canonicalize(Ctx* c, uint32_t n, uint32_t* labels, hipStream_t s) {
  Prof p(c, s, K_CANON, 16.0 * n);
  k_fill_u32<<<stream_grid(c, n), WG, 0, s>>>(c->col, n, 0xFFFFFFFFu);
  k_canon_min<<<stream_grid(c, n), WG, 0, s>>>(n, nullptr, 0, labels, c->col);
  k_canon_apply<<<stream_grid(c, n), WG, 0, s>>>(n, nullptr, 0, labels, c->col);
  return KSPN_OK;
}

int select_pivot_live(Run& R, uint32_t* pivot) {
  Ctx* c = R.c;
  uint32_t u = c->h_cnt[C_AUX];
  uint32_t nparts = std::min<uint32_t>(stream_grid(c, u), (uint32_t)c->num_cu * 16);
  {
    Prof p(c, R.s, K_PIVOT, 20.0 * u);
    if (R.pivot_live_deg)
      k_pivot_live<true><<<nparts, WG, 0, R.s>>>(c->ulist, &c->cnt[C_AUX], R.g->fw_off, R.g->fw_adj, R.g->bw_off,
                                                 R.g->bw_adj, c->live, c->part);
    else
      k_pivot_live<false><<<nparts, WG, 0, R.s>>>(c->ulist, &c->cnt[C_AUX], R.g->fw_off, R.g->fw_adj, R.g->bw_off,
                                                  R.g->bw_adj, c->live, c->part);
    k_pivot_reduce<<<1, WG, 0, R.s>>>(c->part, nparts, &c->cnt[C_FLAG], nullptr);
  }
  KSPN_TRY(read_counters(c, R.s, C_FLAG + 1));
  *pivot = c->h_cnt[C_FLAG];
  if (*pivot == 0xFFFFFFFFu) *pivot = R.n;
  return KSPN_OK;
}

// Statistics scopes of the reference (debug/statistic.hpp:124-205), timed
// with HIP events on the solve's stream: begin() closes the scope before it,
// end() closes the last one, finish() (after the stream is synchronised)
// turns consecutive marks into scope times. Events come from the context's
// pool and go back to it.
struct Phases {
  Ctx* c;
  kspn_stats* st;
  hipStream_t s;
  bool memory;  // record device memory in use at the end of scopes that carry it
  std::vector<std::pair<int, hipEvent_t>> marks;
  int cur = -1;
  Phases(Ctx* c_, kspn_stats* st_, hipStream_t s_, bool mem) : c(c_), st(st_), s(s_), memory(mem) {}
  ~Phases() {
    for (auto& m : marks) c->free_events.push_back(m.second);
  }
  int mark(int idx) {
    hipEvent_t e = nullptr;
    if (!c->free_events.empty()) {
      e = c->free_events.back();
      c->free_events.pop_back();
    } else {
      KSPN_HIP(hipEventCreate(&e));
    }
    marks.push_back({idx, e});
    KSPN_HIP(hipEventRecord(e, s));
    return KSPN_OK;
  }
  void close_scope() {
    if (cur >= 0 && memory && st->phase[cur].has_memory) {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess) st->phase[cur].memory = (uint64_t)(tot - fr);
    }
  }
  int find(const char* name) const {
    for (uint32_t i = 0; i < st->n_phase; ++i)
      if (std::strcmp(st->phase[i].name, name) == 0) return (int)i;
    return -1;
  }
  // enters scope `name` (nothing happens when it is the open one)
  int begin(const char* name, bool has_memory) {
    if (!st) return KSPN_OK;
    if (cur >= 0 && std::strcmp(st->phase[cur].name, name) == 0) return KSPN_OK;
    close_scope();
    int idx = find(name);
    if (idx < 0 && st->n_phase < KSPN_PHASE_MAX) {
      idx = (int)st->n_phase++;
      kspn_phase& ph = st->phase[idx];
      std::memset(&ph, 0, sizeof(ph));
      std::snprintf(ph.name, sizeof(ph.name), "%s", name);
      ph.has_memory = has_memory ? 1u : 0u;
    }
    if (idx >= 0) st->phase[idx].entries++;
    cur = idx;
    c->cur_phase = idx;
    return mark(idx);
  }
  void decided(uint64_t d) {
    if (st && cur >= 0) st->phase[cur].decided += d;
  }
  // decisions of scope `name`, made while another scope was open
  void credit(const char* name, uint64_t d) {
    if (!st) return;
    const int i = find(name);
    if (i >= 0) st->phase[i].decided += d;
  }
  // forgets the open scope when it was entered once, decided nothing and
  // is the last one (a stage the variant would not have entered); its time
  // goes to the scope before it
  void retract_empty() {
    if (!st || cur < 0 || cur + 1 != (int)st->n_phase) return;
    kspn_phase& ph = st->phase[cur];
    if (ph.entries != 1 || ph.decided != 0) return;
    int prev = -1;  // the scope open before it
    for (size_t i = marks.size(); i-- > 0;)
      if (marks[i].first != cur) {
        prev = marks[i].first;
        break;
      }
    for (auto& m : marks)
      if (m.first == cur) m.first = prev;
    for (auto& ep : c->pending)
      if (ep.phase == cur) ep.phase = prev;
    --st->n_phase;
    cur = prev;
    c->cur_phase = prev;
  }
  // closes the open scope; `hidden` starts an interval outside every scope
  // (the batch's label localisation), timed into ms_canon
  int end(bool hidden = false) {
    if (!st) return KSPN_OK;
    close_scope();
    cur = -1;
    c->cur_phase = -1;
    return mark(hidden ? -2 : -1);
  }
  // scope times, and the coarse ms_* fields as sums of scopes
  void finish() {
    if (!st) return;
    double canon = 0;
    for (size_t i = 0; i + 1 < marks.size(); ++i) {
      const double ms = elapsed_ms(marks[i].second, marks[i + 1].second);
      if (marks[i].first >= 0) st->phase[marks[i].first].ms += ms;
      else if (marks[i].first == -2) canon += ms;
    }
    st->ms_canon = canon;
    st->ms_total = marks.size() > 1 ? elapsed_ms(marks.front().second, marks.back().second) : 0.0;
    for (uint32_t i = 0; i < st->n_phase; ++i) {
      const kspn_phase& ph = st->phase[i];
      const char* nm = ph.name;
      if (!std::strcmp(nm, "trim_1_first") || !std::strcmp(nm, "trim_ex_first") || !std::strcmp(nm, "trim_2"))
        st->ms_trim += ph.ms;
      else if (!std::strcmp(nm, "color") || !std::strcmp(nm, "color_trim_ex"))
        st->ms_color += ph.ms;
      else if (!std::strcmp(nm, "residual"))
        st->ms_residual += ph.ms;
      else
        st->ms_fwbw += ph.ms;  // select_pivot, forward_backward_search, trim_1_normal, trim_ex
    }
  }
};

// The stage sequence of each variant at one rank. Scope names follow the
// variant's own statistics (scc.hpp:33-86, variant/scc_*.hpp).
struct Plan {
  bool ex_first;       // exhaustive trim-1 first (trim_1_exhaustive_first) instead of one trim_1_first pass
  bool trim_ex;        // scc_trim_ex: trim-2, and exhaustive trims after FW-BW and after every colour round
  bool fwbw;           // pivot + FW-BW; the residual variants skip it at one rank, where their
                       // decided threshold n - n/1 is 0 (scc_trim_ex_residual.hpp:47, light :36)
  int after;           // after FW-BW: 0 nothing, 1 trim_1_normal, 2 two trim_1_normal passes, 3 exhaustive
  bool residual_rest;  // everything left after the trims is the "residual" scope
  bool light;          // scc_trim_ex_light_residual: trim over the decided stack (empty here) first
};

Plan plan_for(uint32_t method) {
  switch (method) {
    case KSPN_METHOD_TRIM_EX: return {true, true, true, 3, false, false};
    case KSPN_METHOD_HPC_LIKE: return {false, false, true, 0, false, false};
    case KSPN_METHOD_ISPAN_LIKE: return {false, false, true, 2, true, false};
    case KSPN_METHOD_HPC_TRIM_EX: return {true, false, true, 0, false, false};
    case KSPN_METHOD_TRIM_EX_RESIDUAL: return {true, false, false, 0, true, false};
    case KSPN_METHOD_TRIM_EX_LIGHT_RESIDUAL: return {true, false, false, 0, true, true};
    default: return {false, false, true, 1, false, false};  // KSPN_METHOD_KASPAN
  }
}

int run_pipeline(Run& R) {
  Ctx* c = R.c;
  const kspn_graph* g = R.g;
  kspn_stats* st = R.st;
  const Plan plan = plan_for(R.o.method);
  R.pivot_live_deg = plan.ex_first;
  Phases P(c, st, R.s, (R.o.flags & KSPN_FLAG_STATS_MEMORY) != 0);
  const char* first_name = plan.ex_first ? "trim_ex_first" : "trim_1_first";
  const char* after_name = plan.after == 3 ? "trim_ex" : "trim_1_normal";
  const char* rest_name = plan.residual_rest ? "residual" : (plan.trim_ex ? "color_trim_ex" : "color");
  // ---- trim-1 first pass + pivot
  KSPN_TRY(P.begin(first_name, plan.ex_first));
  c->h_u64[1] = 0;
  // The plain single-pivot path seeds FW-BW with the pivot on the device and
  // learns the trim's decisions from the first publication of the forward
  // reach, one host round trip fewer
  const uint32_t rounds_cfg = plan.fwbw ? (R.o.fwbw_rounds ? R.o.fwbw_rounds : 1) : 0;
  const bool defer = rounds_cfg == 1 && !plan.ex_first && !R.pid_preset && R.o.pivot == KSPN_PIVOT_PLAIN &&
                     !R.o.residual_threshold && R.n > 0;
  // the trim also zeroes the counters and the inq, fwv and bwv bitmaps
  uint32_t nparts = std::min<uint32_t>(stream_grid(c, R.n), (uint32_t)c->num_cu * 16);
  {
    Prof p(c, R.s, K_TRIM_FIRST, (defer ? 16.0 : 20.0) * R.n + R.n / 2.0);
    // the seeded FW-BW writes every label after it (k_label_bits)
    k_trim_first<<<nparts, WG, 0, R.s>>>(R.n, g->fw_off, g->bw_off, c->live, defer ? nullptr : R.labels, c->part,
                                         c->inq, c->fwv, c->bwv, c->cnt, C_COUNT_, c->cnt64);
  }
  {
    Prof p(c, R.s, K_PIVOT, 24.0 * nparts);
    FwbwSeed seed{};
    if (defer) seed = FwbwSeed{c->fwv, c->bwv, c->qa, &c->cnt[C_Q0], g->fw_off, c->scan};
    k_pivot_reduce<<<1, WG, 0, R.s>>>(c->part, nparts, &c->cnt[C_FLAG], &c->cnt[C_DECIDED], seed);
  }
  uint32_t decided = 0, pivot = 0;
  if (defer) {
    R.live_pending = true;
    R.live_n = R.n;  // an upper bound until then
  } else {
    KSPN_TRY(read_counters(c, R.s, C_FLAG + 1));
    decided = c->h_cnt[C_DECIDED];
    pivot = c->h_cnt[C_FLAG];
    R.live_n = R.n - decided;
    P.decided(decided);
    if (st) {
      st->decided_trim = decided;
      st->pivot = pivot;
    }
  }
  if (plan.ex_first && R.live_n > 0) {
    uint32_t d = 0;
    KSPN_TRY(trim_exhaustive(R, &d));
    if (st) st->decided_trim += d;
    P.decided(d);
    if (plan.trim_ex && !(R.o.flags & KSPN_FLAG_NO_TRIM2) && R.live_n > 0) {
      KSPN_TRY(P.begin("trim_2", false));
      KSPN_TRY(trim_two(R, &d));
      if (st) st->decided_trim2 += d;
      P.decided(d);
      if (d && R.live_n > 0) {
        KSPN_TRY(trim_exhaustive(R, &d));
        if (st) st->decided_trim += d;
        P.decided(d);
      }
    }
  }
  // trims after a FW-BW round
  auto after_trim = [&]() -> int {
    if (plan.after == 0 || R.live_n == 0) return KSPN_OK;
    KSPN_TRY(P.begin(after_name, false));
    uint32_t d = 0;
    if (plan.after == 3) {
      KSPN_TRY(trim_exhaustive(R, &d));
    } else {
      KSPN_TRY(trim_normal(R, &d));
      if (plan.after == 2 && R.live_n > 0) {
        uint32_t d2 = 0;
        KSPN_TRY(trim_normal(R, &d2));
        d += d2;
      }
    }
    if (st) st->decided_trim += d;
    P.decided(d);
    return KSPN_OK;
  };
  // ---- FW-BW rounds
  uint32_t rounds = plan.fwbw ? (R.o.fwbw_rounds ? R.o.fwbw_rounds : 1) : 0;
  const uint32_t tail = (R.o.flags & KSPN_FLAG_NO_TAIL) ? 0u : (R.o.tail_threshold ? R.o.tail_threshold : 4096u);
  uint32_t* rest_field = st ? (plan.residual_rest ? &st->decided_residual : &st->decided_color) : nullptr;
  // the single-pivot path may hand what FW-BW leaves straight to the tail
  // solver (trim_1_normal's decisions are the solver's isolated vertices)
  const uint32_t fold_limit = std::min<uint32_t>(tail, 4095u);
  const bool fold_tail = fold_limit > 0 && plan.after != 3 && !R.o.residual_threshold && !R.o.max_rounds &&
                         !R.pid_preset && !(R.o.flags & KSPN_FLAG_NO_REFINE);
  if (plan.fwbw && (rounds > 1 || R.pid_preset) && R.live_n > 0) {
    // multi-pivot: one pivot per partition per round
    uint32_t d = 0;
    KSPN_TRY(P.begin("forward_backward_search", true));
    KSPN_TRY(pfwbw_rounds(R, rounds, R.np0, &d));
    if (st) st->decided_fwbw += d;
    P.decided(d);
    R.pid_active = true;
    KSPN_TRY(after_trim());
    rounds = 0;  // the single-pivot loop below does nothing
  }
  for (uint32_t k = 0; k < rounds && R.live_n > 0; ++k) {
    if (R.o.residual_threshold && R.live_n <= R.o.residual_threshold) break;
    if (R.o.pivot == KSPN_PIVOT_LEARNED) {
      // MLP over the top degree-product candidates (the plain pivot is one of them)
      KSPN_TRY(P.begin("select_pivot", false));
      KSPN_TRY(select_pivot_learned_impl(c, g, R.o.mlp_params, R.s, &pivot));
      if (k == 0 && st) st->pivot = pivot;
    } else if (k > 0 || plan.ex_first) {
      // pivot among live vertices by original degree product
      KSPN_TRY(P.begin("select_pivot", false));
      KSPN_HIP(hipMemsetAsync(c->cnt, 0, sizeof(uint32_t) * C_COUNT_, R.s));
      // reuse k_trim_first's reduction on a live-masked pass
      uint32_t u = 0;
      KSPN_TRY(build_live_list(c, R.n, c->live, c->ulist, &u, R.s));
      (void)u;
      KSPN_TRY(select_pivot_live(R, &pivot));
      if (k == 0 && st) st->pivot = pivot;
    }
    if (pivot >= R.n && !R.live_pending) break;
    uint32_t d = 0;
    KSPN_TRY(P.begin("forward_backward_search", true));
    if (k + 1 == rounds && fold_tail) {
      // the last round: what is left may be small enough for the tail
      // solver, which then runs before the host has seen the FW-BW result
      const bool deferred = R.live_pending;
      KSPN_TRY(fwbw(R, pivot, &d, false, true));
      if (deferred) {
        // the trim's counts arrived with the forward reach
        P.credit(first_name, R.trim_decided);
        if (st) {
          st->decided_trim = R.trim_decided;
          st->pivot = R.first_pivot == 0xFFFFFFFFu ? R.n : R.first_pivot;
        }
      }
      KSPN_TRY(P.begin(plan.after ? after_name : rest_name, !plan.after));
      uint32_t u = 0, iso = 0;
      bool done = false;
      KSPN_TRY(tail_spec(R, fold_limit, &d, &u, &iso, &done));
      if (st) st->decided_fwbw += d;
      P.credit("forward_backward_search", d);
      // trim_1_normal's decisions (live vertices without a live neighbour)
      // belong to that scope; variants without it colour them
      P.decided(iso);
      if (st) {
        if (plan.after) st->decided_trim += iso;
        else if (rest_field) *rest_field += iso;
      }
      if ((uint64_t)R.live_n + iso + (done ? u : 0) == 0) {
        P.retract_empty();  // FW-BW decided everything
        if (R.trim_decided == R.n && deferred) P.retract_empty();  // so did the trim: no FW-BW scope either
        break;
      }
      if (done && u > 0) {
        if (plan.after) KSPN_TRY(P.begin(rest_name, true));
        P.decided(u);
        if (rest_field) *rest_field += u;
      }
      break;  // on to the colour rounds when the tail did not run
    }
    const bool deferred = R.live_pending;
    KSPN_TRY(fwbw(R, pivot, &d));
    if (deferred) {
      P.credit(first_name, R.trim_decided);
      if (st) {
        st->decided_trim = R.trim_decided;
        st->pivot = R.first_pivot == 0xFFFFFFFFu ? R.n : R.first_pivot;
      }
      if (R.trim_decided == R.n) P.retract_empty();  // the trim decided everything: no FW-BW scope
    }
    if (st) st->decided_fwbw += d;
    P.decided(d);
    if (R.live_n == 0) break;
    KSPN_TRY(after_trim());
  }
  // ---- colour rounds (the residual variants' "residual" scope)
  if (plan.light && R.live_n > 0) {
    // trim_1_exhaustive over the decisions since the first trim: at one rank
    // FW-BW and colouring were skipped, the decided stack is empty
    KSPN_TRY(P.begin("trim_ex", false));
  }
  uint32_t cap = R.o.max_rounds;
  const uint32_t stop = plan.residual_rest ? 0u : R.o.residual_threshold;
  if (R.live_n > 0 && !(stop && R.live_n <= stop)) KSPN_TRY(P.begin(rest_name, true));
  if (!(R.o.flags & KSPN_FLAG_NO_REFINE) && !R.pid_active && R.live_n > tail) {
    // colour rounds refine one partition of all live vertices
    if (!c->pid) KSPN_TRY(dalloc(&c->pid, (uint64_t)c->cap_n + 64));
    k_fill_u32<<<stream_grid(c, R.n), WG, 0, R.s>>>(c->pid, R.n, 0u);
    R.pid_active = true;
  }
  if (R.o.flags & KSPN_FLAG_NO_REFINE) R.pid_active = false;
  while (R.live_n > 0) {
    if (stop && R.live_n <= stop) break;
    if (cap && R.rounds >= cap) return KSPN_E_ITER_CAP;
    if (R.live_n <= tail) {
      uint32_t d = 0;
      KSPN_TRY(tail_solve(R, &d));
      if (rest_field) *rest_field += d;
      P.decided(d);
      break;
    }
    uint32_t d = 0;
    KSPN_TRY(color_round(R, &d));
    ++R.rounds;
    if (rest_field) *rest_field += d;
    P.decided(d);
    if (d == 0) return KSPN_E_ITER_CAP;  // cannot happen: the max key is always a root
    if (plan.trim_ex && R.live_n > 0) {
      KSPN_TRY(trim_exhaustive(R, &d));
      if (st) st->decided_trim += d;
      P.decided(d);
    }
  }
  // ---- residual Tarjan (opts.residual_threshold)
  if (R.live_n > 0) {
    uint32_t d = 0;
    KSPN_TRY(P.begin("residual", true));
    KSPN_TRY(residual(R, &d));
    if (st) st->decided_residual += d;
    P.decided(d);
  }
  // every stage writes canonical ids (own id, pair minimum, SCC minimum),
  // so no pass over all labels is needed here; a batch's labels become
  // local to their graph
  if (R.vtx_off) {
    KSPN_TRY(P.end(true));
    Prof p(c, R.s, K_LABEL, 8.0 * R.n);
    k_batch_localize<<<std::min<uint32_t>(R.np0, 16384u), WG, 0, R.s>>>(R.np0, R.vtx_off, R.labels);
  }
  KSPN_TRY(P.end());
  KSPN_HIP(hipStreamSynchronize(R.s));
  P.finish();
  return KSPN_OK;
}

}  // namespace

}  // namespace kspn

// =====================================================================
// C ABI
// =====================================================================

using namespace kspn;

extern "C" {

void kspn_opts_default(kspn_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->method = KSPN_METHOD_KASPAN;
  o->pivot = KSPN_PIVOT_PLAIN;
  o->fwbw_rounds = 1;
  o->alpha = 14.0f;
  o->beta = 24.0f;
}

const char* kspn_last_error(void) { return kspn::g_last_error; }

const char* kspn_strerror(int code) {
  switch (code) {
    case KSPN_OK: return "ok";
    case KSPN_E_INVALID: return "invalid argument";
    case KSPN_E_OOM: return "out of memory";
    case KSPN_E_HIP: return "HIP runtime error";
    case KSPN_E_ITER_CAP: return "iteration cap reached";
    case KSPN_E_UNIMPLEMENTED: return "not implemented";
    case KSPN_E_IO: return "file error";
    default: return "unknown error";
  }
}

int kspn_ctx_create(kspn_ctx** out, int device) {
  if (!out) return KSPN_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return KSPN_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return KSPN_E_HIP;
  kspn_ctx* k = new (std::nothrow) kspn_ctx();
  if (!k) return KSPN_E_OOM;
  Ctx* c = &k->c;
  c->device = device;
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0) c->num_cu = cu;
  void* hm = nullptr;
  if (hipMalloc((void**)&c->cnt, sizeof(uint32_t) * C_COUNT_) != hipSuccess ||
      hipMalloc((void**)&c->cnt64, sizeof(unsigned long long) * 4) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, sizeof(uint32_t) * C_COUNT_, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_u64, sizeof(uint64_t) * 4, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&hm, sizeof(uint32_t) * MAP_BANK * 2, hipHostMallocMapped) != hipSuccess) {
    kspn_ctx_destroy(k);
    return KSPN_E_OOM;
  }
  c->h_map = (volatile uint32_t*)hm;
  if (hipHostGetDevicePointer((void**)&c->d_map, hm, 0) != hipSuccess) {
    kspn_ctx_destroy(k);
    return KSPN_E_HIP;
  }
  std::memset(c->h_cnt, 0, sizeof(uint32_t) * C_COUNT_);
  std::memset(hm, 0, sizeof(uint32_t) * MAP_BANK * 2);
  std::memset(c->h_u64, 0, sizeof(uint64_t) * 4);
  c->kt.resize(K_COUNT);
  for (int i = 0; i < K_COUNT; ++i) c->kt[i].name = kKernNames[i];
  // persistent reachability: one workgroup of COOP_WG threads per CU, all
  // resident at once (the kernel's barrier gives up rather than wait forever)
  {
    int per_cu = 1 << 30;
    for (const void* fn : {reinterpret_cast<const void*>(&k_reach_coop<OP_FW>),
                           reinterpret_cast<const void*>(&k_reach_coop<OP_BW>),
                           reinterpret_cast<const void*>(&k_reach_coop<OP_PBW>)}) {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, COOP_WG, 0) != hipSuccess) b = 0;
      per_cu = std::min(per_cu, b);
    }
    if (per_cu >= 1) {
      if (hipMalloc((void**)&c->coop, sizeof(CoopCtl)) != hipSuccess) {
        kspn_ctx_destroy(k);
        return KSPN_E_OOM;
      }
      if (hipMemset(c->coop, 0, sizeof(CoopCtl)) != hipSuccess) {
        kspn_ctx_destroy(k);
        return KSPN_E_HIP;
      }
      c->coop_grid = (uint32_t)c->num_cu;
    }
  }
  *out = k;
  return KSPN_OK;
}

int kspn_ctx_destroy(kspn_ctx* k) {
  if (!k) return KSPN_OK;
  Ctx* c = &k->c;
  (void)hipSetDevice(c->device);
  graph_release(c);
  if (c->aq) (void)hipFree(c->aq);
  if (c->actl) (void)hipFree(c->actl);
  void* ptrs[] = {c->live, c->fwv, c->bwv, c->inq, c->inq2, c->cchain, c->fnew, c->sbm, c->qa, c->qb, c->qc, c->ulist, c->col, c->deg, c->scan,
                  c->part, c->cnt, c->cub_tmp, c->sub_off, c->sub_adj, c->tj, c->mlp_buf, c->pid, c->pbest,
                  c->pmin, c->pflags, c->pnew, c->vscore, c->vfeat, c->coop, c->pcnt, c->vprio, c->ds_st, c->deg8, c->cpw,
                  c->prune_fw, c->prune_bw, c->tstart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->h_map) (void)hipHostFree((void*)c->h_map);
  if (c->cnt64) (void)hipFree(c->cnt64);
  if (c->h_u64) (void)hipHostFree(c->h_u64);
  for (auto& ep : c->pending) {
    (void)hipEventDestroy(ep.a);
    (void)hipEventDestroy(ep.b);
  }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  delete k;
  return KSPN_OK;
}

int kspn_ctx_reserve(kspn_ctx* k, uint32_t n, uint64_t m) {
  if (!k) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  return ctx_reserve(&k->c, n, m);
}

}  // extern "C"

namespace {

// one solve: a graph, or the union graph of a batch (vtx_off != nullptr)
int solve(kspn_ctx* k, const kspn_graph* g, uint32_t* labels, const kspn_opts* opts, kspn_stats* stats, void* stream,
          const uint32_t* vtx_off, uint32_t count) {
  if (!k || !g || (g->n > 0 && (!labels || !g->fw_off || !g->bw_off))) return KSPN_E_INVALID;
  if (g->m > 0 && (!g->fw_adj || !g->bw_adj)) return KSPN_E_INVALID;
  if (g->n == 0xFFFFFFFFu) return KSPN_E_INVALID;  // 0xFFFFFFFF is the undecided sentinel
  kspn_opts o;
  if (opts) o = *opts;
  else kspn_opts_default(&o);
  if (o.method > KSPN_METHOD_TRIM_EX_LIGHT_RESIDUAL) return KSPN_E_INVALID;
  if (o.pivot > KSPN_PIVOT_RANDOM) return KSPN_E_INVALID;
  if (o.pivot == KSPN_PIVOT_LEARNED && !o.mlp_params) return KSPN_E_INVALID;
  if ((o.pull_ilp && o.pull_ilp != 1 && o.pull_ilp != 2 && o.pull_ilp != 4) ||
      (o.expand_ept && o.expand_ept != 2 && o.expand_ept != 4 && o.expand_ept != 8 && o.expand_ept != 16) ||
      o.sweep_wgs > 64 ||
      o.list_wgs > 64 || o.chain_levels > 256)
    return KSPN_E_INVALID;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (g->n == 0) return KSPN_OK;
  if (vtx_off && (count == 0 || count > g->n)) return KSPN_E_INVALID;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, g->n, g->m));
  hipStream_t s = (hipStream_t)stream;
  c->profile = stats && (o.flags & KSPN_FLAG_PROFILE);
  c->cur_phase = -1;
  if (c->profile) {
    for (auto& t : c->kt) { t.launches = 0; t.ms = 0; t.bytes = 0; }
    for (int i = 0; i < KSPN_PHASE_MAX; ++i) c->phase_bytes[i] = c->phase_kms[i] = 0;
  }
  if (o.flags & KSPN_FLAG_GRAPH) return graph_solve(c, g, labels, o, vtx_off, count, stats, s);
  Run R{c, g, labels, o, stats, s, g->n, bm_words(g->n)};
  R.coop_edges = o.coop_edges ? o.coop_edges : kCoopEdges;
  R.coop_deg = o.coop_degree ? o.coop_degree : kCoopDeg;
  R.coop_wg = std::min<uint32_t>(o.coop_workgroups ? o.coop_workgroups : kCoopWorkgroups, c->coop_grid);
  R.pull_ilp = o.pull_ilp ? o.pull_ilp : (uint32_t)PULL_ILP;
  R.ept = o.expand_ept;  // 0: chosen per level
  if (!(o.flags & KSPN_FLAG_PERSISTENT) || (o.flags & KSPN_FLAG_CHAIN)) {
    // chained launches (the default) need no co-residency: the grid is not capped by it
    R.chain = o.chain_levels ? o.chain_levels : kChainLevels;
    R.coop_wg = o.coop_workgroups ? o.coop_workgroups : kCoopWorkgroups;
  }
  if (o.flags & KSPN_FLAG_NARROW) R.narrow_len = o.chain_levels ? o.chain_levels : kNarrowChain;
  if (vtx_off) {
    // graph i is preset partition i of the multi-pivot FW-BW rounds
    if (!c->pid) KSPN_TRY(dalloc(&c->pid, (uint64_t)c->cap_n + 64));
    k_batch_pid<<<std::min<uint32_t>(count, 16384u), WG, 0, s>>>(count, vtx_off, c->pid);
    KSPN_HIP(hipGetLastError());
    R.pid_preset = true;
    R.np0 = count;
    R.vtx_off = vtx_off;
  }
  if (o.flags & KSPN_FLAG_PRUNE) {
    // one bit per adjacency entry in each direction, all clear: no edge is
    // known yet to join two partitions
    const uint64_t words = R.g->m / 32 + 1;
    if (c->prune_words < words) {
      c->prune_words = 0;
      KSPN_TRY(dalloc(&c->prune_fw, words));
      KSPN_TRY(dalloc(&c->prune_bw, words));
      c->prune_words = words;
    }
    k_fill_u32<<<stream_grid(c, words), WG, 0, s>>>(c->prune_fw, words, 0u);
    k_fill_u32<<<stream_grid(c, words), WG, 0, s>>>(c->prune_bw, words, 0u);
    R.prune = true;
  }
  int rc = run_pipeline(R);
  if (c->profile) {
    (void)hipStreamSynchronize(s);
    prof_collect(c);
    uint32_t nk = 0;
    for (int i = 0; i < K_COUNT && nk < KSPN_KERN_MAX; ++i) {
      if (!c->kt[i].launches) continue;
      kspn_kern_stat& ks = stats->kern[nk++];
      std::snprintf(ks.name, sizeof(ks.name), "%s", c->kt[i].name);
      ks.launches = c->kt[i].launches;
      ks.ms = c->kt[i].ms;
      ks.bytes = c->kt[i].bytes;
    }
    stats->n_kern = nk;
    // algorithmic bytes and launch time per scope, and the coarse bytes_* as sums of scopes
    for (uint32_t i = 0; i < stats->n_phase && i < (uint32_t)KSPN_PHASE_MAX; ++i) {
      kspn_phase& ph = stats->phase[i];
      ph.bytes = c->phase_bytes[i];
      ph.kernel_ms = c->phase_kms[i];
      const char* nm = ph.name;
      if (!std::strcmp(nm, "trim_1_first") || !std::strcmp(nm, "trim_ex_first") || !std::strcmp(nm, "trim_2"))
        stats->bytes_trim += ph.bytes;
      else if (!std::strcmp(nm, "color") || !std::strcmp(nm, "color_trim_ex"))
        stats->bytes_color += ph.bytes;
      else if (!std::strcmp(nm, "residual"))
        stats->bytes_residual += ph.bytes;
      else
        stats->bytes_fwbw += ph.bytes;
    }
    c->profile = false;
  }
  return rc;
}

}  // namespace

extern "C" {

int kspn_scc(kspn_ctx* k, const kspn_graph* g, uint32_t* labels, const kspn_opts* opts, kspn_stats* stats, void* stream) {
  return solve(k, g, labels, opts, stats, stream, nullptr, 0);
}

int kspn_scc_batch(kspn_ctx* k, const kspn_batch* b, uint32_t* labels, const kspn_opts* opts, kspn_stats* stats,
                   void* stream) {
  if (!b || (b->n > 0 && !b->vtx_off)) return KSPN_E_INVALID;
  kspn_graph g{b->n, b->m, b->fw_off, b->fw_adj, b->bw_off, b->bw_adj};
  return solve(k, &g, labels, opts, stats, stream, b->vtx_off, b->count);
}

int kspn_scc_host(kspn_ctx* k, uint32_t n, uint64_t m, const uint64_t* fw_off, const uint32_t* fw_adj,
                  const uint64_t* bw_off, const uint32_t* bw_adj, uint32_t* labels, const kspn_opts* opts,
                  kspn_stats* stats) {
  if (!k || (n > 0 && (!fw_off || !bw_off || !labels)) || (m > 0 && (!fw_adj || !bw_adj))) return KSPN_E_INVALID;
  if (n == 0) return KSPN_OK;
  if (fw_off[0] != 0 || bw_off[0] != 0 || fw_off[n] != m || bw_off[n] != m) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  uint64_t *dfo = nullptr, *dbo = nullptr;
  uint32_t *dfa = nullptr, *dba = nullptr, *dl = nullptr;
  int rc = KSPN_OK;
  if (hipMalloc((void**)&dfo, (n + 1ull) * 8) != hipSuccess || hipMalloc((void**)&dbo, (n + 1ull) * 8) != hipSuccess ||
      hipMalloc((void**)&dfa, (m ? m : 1) * 4) != hipSuccess || hipMalloc((void**)&dba, (m ? m : 1) * 4) != hipSuccess ||
      hipMalloc((void**)&dl, (uint64_t)n * 4) != hipSuccess) {
    rc = KSPN_E_OOM;
  }
  if (rc == KSPN_OK) {
    if (hipMemcpy(dfo, fw_off, (n + 1ull) * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dbo, bw_off, (n + 1ull) * 8, hipMemcpyHostToDevice) != hipSuccess ||
        (m && hipMemcpy(dfa, fw_adj, m * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (m && hipMemcpy(dba, bw_adj, m * 4, hipMemcpyHostToDevice) != hipSuccess))
      rc = KSPN_E_HIP;
  }
  if (rc == KSPN_OK) {
    kspn_graph g{n, m, dfo, dfa, dbo, dba};
    rc = kspn_scc(k, &g, dl, opts, stats, nullptr);
  }
  if (rc == KSPN_OK && hipMemcpy(labels, dl, (uint64_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = KSPN_E_HIP;
  for (void* p : {(void*)dfo, (void*)dbo, (void*)dfa, (void*)dba, (void*)dl})
    if (p) (void)hipFree(p);
  return rc;
}

int kspn_canonicalize(kspn_ctx* k, uint32_t n, uint32_t* labels, void* stream) {
  if (!k || (n > 0 && !labels)) return KSPN_E_INVALID;
  if (n == 0) return KSPN_OK;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, n, 0));
  hipStream_t s = (hipStream_t)stream;
  KSPN_TRY(canonicalize(c, n, labels, s));
  KSPN_HIP(hipStreamSynchronize(s));
  return KSPN_OK;
}

int kspn_tarjan(kspn_ctx* k, uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* adj, uint32_t* labels,
                void* stream) {
  if (!k || (n > 0 && (!off || !labels)) || (m > 0 && !adj)) return KSPN_E_INVALID;
  if (n == 0) return KSPN_OK;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, n, m));
  if (!c->tj) KSPN_TRY(dalloc(&c->tj, 6ull * c->cap_n + 64));
  hipStream_t s = (hipStream_t)stream;
  k_tarjan<<<1, 64, 0, s>>>(n, off, adj, c->tj, labels);
  KSPN_HIP(hipGetLastError());
  KSPN_HIP(hipStreamSynchronize(s));
  return KSPN_OK;
}

int kspn_subgraph(kspn_ctx* k, uint32_t n, const uint64_t* off, const uint32_t* adj, const uint8_t* mask,
                  uint32_t* super_ids, uint64_t* sub_off, uint32_t* sub_adj, uint32_t* sub_n, uint64_t* sub_m,
                  void* stream) {
  if (!k || !sub_n || !sub_m || (n > 0 && (!off || !mask || !super_ids || !sub_off))) return KSPN_E_INVALID;
  *sub_n = 0;
  *sub_m = 0;
  if (n == 0) return KSPN_OK;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, n, 0));
  hipStream_t s = (hipStream_t)stream;
  // byte mask -> bitmap -> ascending list of members
  uint64_t words = bm_words(n);
  KSPN_HIP(hipMemsetAsync(c->fwv, 0, words * 4, s));
  k_mask_to_bitmap<<<stream_grid(c, n), WG, 0, s>>>(n, mask, c->fwv);
  uint32_t sn = 0;
  KSPN_TRY(build_live_list(c, n, c->fwv, super_ids, &sn, s));
  if (sn == 0) {
    KSPN_HIP(hipMemsetAsync(sub_off, 0, 8, s));
    KSPN_HIP(hipStreamSynchronize(s));
    return KSPN_OK;
  }
  k_copy_u32<<<1, 64, 0, s>>>(&c->cnt[C_RES_N], &c->cnt[C_AUX]);
  k_sub_ids<<<stream_grid(c, sn), WG, 0, s>>>(super_ids, &c->cnt[C_RES_N], c->col);
  k_sub_deg<<<stream_grid(c, (uint64_t)sn + 1), WG, 0, s>>>(super_ids, &c->cnt[C_RES_N], off, adj, nullptr, mask, c->deg);
  KSPN_TRY(exclusive_scan_u64(c, c->deg, sub_off, (uint64_t)sn + 1, s));
  k_sub_fill<<<stream_grid(c, sn), WG, 0, s>>>(super_ids, &c->cnt[C_RES_N], off, adj, nullptr, mask, c->col, sub_off,
                                              sub_adj);
  KSPN_TRY(publish(c, s, sub_off + sn));
  *sub_n = sn;
  *sub_m = c->h_u64[0];
  return KSPN_OK;
}

}  // extern "C"


// ---------------------------------------------------------------- test hook
extern "C" int kspn_test_tail_guard(kspn_ctx* k, const kspn_graph* g, uint32_t n_limit, const uint32_t* list,
                                    uint32_t count, uint32_t* flagged, void* stream) {
  if (!k || !g || !list || !flagged || n_limit > g->n || count == 0 || count > g->n || count > 4095)
    return KSPN_E_INVALID;
  *flagged = 0;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, g->n, g->m));
  hipStream_t s = (hipStream_t)stream;
  if (!c->sub_off) KSPN_TRY(dalloc(&c->sub_off, (uint64_t)c->cap_n + 64));
  if (c->cap_sub_m < 2 * g->m + 2) {
    KSPN_TRY(dalloc(&c->sub_adj, 2 * g->m + 2));
    c->cap_sub_m = 2 * g->m + 2;
  }
  uint32_t* labels = nullptr;
  KSPN_TRY(dalloc(&labels, (uint64_t)g->n));
  const uint64_t words = bm_words(g->n);
  int rc = KSPN_OK;
  // every vertex live; the list in qa with its count in C_LIVE, as the tail split finds them
  auto run = [&]() -> int {
    KSPN_HIP(hipMemsetAsync(c->live, 0xFF, words * 4, s));
    KSPN_HIP(hipMemcpyAsync(c->qa, list, sizeof(uint32_t) * count, hipMemcpyDeviceToDevice, s));
    KSPN_HIP(hipMemcpyAsync(c->ulist, list, sizeof(uint32_t) * count, hipMemcpyDeviceToDevice, s));
    k_set_u32<<<1, 64, 0, s>>>(&c->cnt[C_LIVE], count);
    k_set_u32<<<1, 64, 0, s>>>(&c->cnt[C_AUX], count);
    const uint32_t kernels[3] = {1, 2, 4};
    for (uint32_t which : kernels) {
      k_zero3<<<1, 64, 0, s>>>(&c->cnt[C_BADID], &c->cnt[C_RES_N], &c->cnt[C_ROOM]);
      k_set_u32<<<1, 64, 0, s>>>(&c->cnt[C_ERR], 0u);
      if (which == 1)
        k_tail_split<<<8, WG, 0, s>>>(n_limit, c->qa, &c->cnt[C_LIVE], g->fw_off, g->fw_adj, g->bw_off, g->bw_adj,
                                      c->live, labels, c->qb, &c->cnt[C_RES_N], &c->cnt[C_BADID]);
      else if (which == 2)
        k_tail_prep<<<8, WG, 0, s>>>(n_limit, c->ulist, &c->cnt[C_AUX], 4095u, g->fw_off, g->fw_adj, g->bw_off,
                                     g->bw_adj, c->live, c->col, c->scan, c->sub_off, nullptr, &c->cnt[C_BADID]);
      else {
        KSPN_HIP(hipMemsetAsync(c->scan, 0, sizeof(uint64_t) * (count + 1), s));
        KSPN_HIP(hipMemsetAsync(c->sub_off, 0, sizeof(uint64_t) * (count + 1), s));
        k_tail_fill<<<8, WG, 0, s>>>(n_limit, c->ulist, &c->cnt[C_AUX], g->fw_off, g->fw_adj, g->bw_off, g->bw_adj,
                                     c->live, c->col, c->scan, c->sub_off, c->sub_adj, c->cap_sub_m, &c->cnt[C_ERR],
                                     &c->cnt[C_ROOM], &c->cnt[C_BADID]);
      }
      KSPN_TRY(publish(c, s, nullptr));
      if (c->h_cnt[C_BADID]) *flagged |= which;
    }
    k_set_u32<<<1, 64, 0, s>>>(&c->cnt[C_BADID], 0u);
    KSPN_HIP(hipStreamSynchronize(s));
    return KSPN_OK;
  };
  rc = run();
  (void)hipFree(labels);
  return rc;
}

// ---------------------------------------------------------------- Tarjan without allocation
// kspn_tarjan_noalloc keeps k_tarjan's work area (index, low, stack, frame
// vertices and 64-bit frame iterators: 6 words per vertex and 2 of slack)
// in one file-scope device array instead of context memory. Every device
// that runs it holds its own copy for the life of the process. Its address
// is taken once per device with that device current; host code passes only
// that address, never the array's name (a host-side stand-in).
namespace {
constexpr uint32_t kTarjanStaticMaxN = KSPN_TARJAN_NOALLOC_MAX_N;
constexpr uint64_t kTarjanStaticWords = 6ull * kTarjanStaticMaxN + 64;
__device__ __align__(256) uint32_t scratch[kTarjanStaticWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

// the calling device's copy of scratch (taken once per device)
int scratch_for_device(int dev, uint32_t** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    note_error("kspn_tarjan_noalloc", "device index beyond the library's table", __FILE__, __LINE__);
    return KSPN_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    KSPN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return KSPN_OK;
}
}  // namespace

extern "C" int kspn_tarjan_noalloc(uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* adj,
                                   uint32_t* labels, void* stream) {
  if ((n > 0 && (!off || !labels)) || (m > 0 && !adj)) return KSPN_E_INVALID;
  if (n > kTarjanStaticMaxN) {
    note_error("kspn_tarjan_noalloc", "n exceeds KSPN_TARJAN_NOALLOC_MAX_N", __FILE__, __LINE__);
    return KSPN_E_INVALID;
  }
  if (n == 0) return KSPN_OK;
  int dev = 0;
  KSPN_HIP(hipGetDevice(&dev));
  uint32_t* work = nullptr;
  KSPN_TRY(scratch_for_device(dev, &work));
  // k_tarjan initialises index[] itself and writes every other word before reading it
  k_tarjan<<<1, 64, 0, (hipStream_t)stream>>>(n, off, adj, work, labels);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}
