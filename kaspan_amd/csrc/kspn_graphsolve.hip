// kspn_graphsolve.hip -- the device-driven solve (KSPN_FLAG_GRAPH).
//
// The host-driven solve (kspn_scc.hip) reads counters after most levels and
// decides the next launch on the host. This solve decides everything on
// the device: a fixed sequence of launches, whose persistent kernels loop
// over levels and rounds behind grid barriers, with no host round trip
// until the labels are complete. A fixed sequence can be captured once into
// a HIP graph and replayed, which is how kspn_scc / kspn_scc_batch run it.
//
// Stages (the shape of the reference's hpc_like variant,
// scc/variant/scc_hpc_like.hpp, with one pivot per graph of a batch):
//   k_g_clear    control block, per-graph pivot keys and minima
//   k_g_init     trim_1_first (scc/trim_1_local.hpp:12-46): vertices of in-
//                or out-degree 0 are decided with their own id; the others
//                are live; per graph the live vertex of largest in*out
//                (smallest id on ties) becomes the pivot
//   k_g_seed     the pivots start the forward reach
//   k_g_reach<0> forward reach over live vertices (persistent: push levels
//                over 256-edge row slices, Beamer's switch to pull sweeps)
//   k_g_seed     the pivots start the backward reach
//   k_g_reach<1> backward reach inside the forward sets
//   k_g_scc      FW and BW both reached: the pivot's SCC, labelled with its
//                smallest member (k_g_scc_min, then k_g_scc_label)
//   k_g_colour   colour rounds to completion (persistent): keys
//                ~rotr(v, r) propagated by max along in-edges inside each
//                partition, roots = vertices keyed by their own colour,
//                backward closure within a colour, smallest member as the
//                label; the partition becomes the colour's root
//                (cache_multi_pivot_search_pull.hpp:40-266 for the rounds;
//                the refinement keeps the partition of every SCC)
//   k_g_localize batch labels made local to each graph
//
// Grid barriers follow the HIP memory model: every wave releases its
// writes at agent scope before the workgroup barrier, one thread per
// workgroup arrives with an atomic add and polls, then acquires at agent
// scope before the workgroup continues. Everything written in one phase is
// therefore visible to every workgroup in the next with plain loads. Inside
// a phase, reads of values other workgroups write (colours, visited bits)
// may be stale; every loop that relies on them repeats until a whole
// sweep, started after a barrier, changes nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "kspn_internal.h"

namespace kspn {
namespace gs {

constexpr int GWG = 1024;                 // threads per workgroup of the persistent kernels
constexpr int GWAVES = GWG / WAVE;        // wavefronts per workgroup
constexpr uint32_t SLICE = 256;           // edges per frontier entry (a vertex's row in slices)
constexpr uint32_t SPIN_LIMIT = 1u << 22; // barrier polls before giving up
constexpr uint32_t ERR_BARRIER = 1, ERR_QUEUE = 2;

struct GCtl {
  uint32_t bar[4];                // grid barrier arrivals, one counter per persistent launch
  uint32_t err;                   // ERR_* bits
  uint32_t done;                  // every vertex decided
  uint32_t incomplete;            // the colour rounds hit their cap
  uint32_t rounds;                // colour rounds run
  uint32_t qn[3], qv[3], cur[3];  // frontier ring (level % 3): entries, vertices, entries handed out
  uint32_t ucount;                // live vertices listed by the colour round
  unsigned long long qe[3];       // out-edges of each frontier's vertices
  unsigned long long expanded[2]; // edges expanded by push levels, per direction
  uint32_t levels_push[2], levels_pull[2];
  uint32_t changed[3];            // fixpoint flags (ring)
  uint32_t dec_trim, dec_fwbw, dec_color;
  uint32_t live_after_trim;
};

__device__ inline uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// grid barrier k (1, 2, ...) of a persistent launch over counter *bar
__device__ bool gbar(uint32_t* bar, uint32_t* err, uint32_t k) {
  __shared__ uint32_t s_ok;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this wave's writes, before the workgroup arrives
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    const uint32_t target = k * gridDim.x;
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (ld_agent(bar) < target) {
      if (++spins > SPIN_LIMIT || ld_agent(err)) {
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (!ok) atomicOr(err, ERR_BARRIER);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // what the others wrote before arriving
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__device__ inline uint32_t wscan_incl(uint32_t v) {
  const uint32_t lane = lane_id();
  for (uint32_t o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Concurrent atomics on one address serialise (about 11 ns each), so counters
// and flags every wavefront updates are combined per workgroup first: one
// atomic per workgroup. Every thread of the workgroup calls these.
__device__ inline void block_add(uint32_t v, uint32_t* dst) {
  __shared__ uint32_t s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  v = wave_sum(v);
  if (lane_id() == 0 && v) atomicAdd(&s_sum, v);
  __syncthreads();
  if (threadIdx.x == 0 && s_sum) atomicAdd(dst, s_sum);
}
__device__ inline void block_add2(uint32_t v, unsigned long long w, uint32_t* dst, unsigned long long* dst64) {
  __shared__ uint32_t s_sum;
  __shared__ unsigned long long s_sum64;
  if (threadIdx.x == 0) {
    s_sum = 0;
    s_sum64 = 0;
  }
  __syncthreads();
  v = wave_sum(v);
  w = wave_sum64(w);
  if (lane_id() == 0) {
    if (v) atomicAdd(&s_sum, v);
    if (w) atomicAdd(&s_sum64, w);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_sum) atomicAdd(dst, s_sum);
    if (s_sum64) atomicAdd(dst64, s_sum64);
  }
}

constexpr uint32_t HEAVY = 64;  // rows longer than this are scanned by the whole wavefront

// The lanes with `heavy` set own rows [it, end) longer than HEAVY: each such
// row in turn is scanned by all 64 lanes, `hit(h, j, valid)` being called by
// every lane for edge j of lane h's row (valid: j is inside it). A row's
// scan stops at its first hit. Returns, for each heavy lane, whether its
// row had a hit. Every lane of the wavefront must call it.
template <class F>
__device__ inline bool heavy_rows(bool heavy, uint64_t it, uint64_t end, F&& hit) {
  uint64_t pend = __ballot(heavy);
  const uint32_t lane = lane_id();
  bool found = false;
  while (pend) {
    const uint32_t h = __ffsll((unsigned long long)pend) - 1;
    const uint64_t b = __shfl(it, h), e = __shfl(end, h);
    for (uint64_t j0 = b; j0 < e; j0 += WAVE) {
      const uint64_t j = j0 + lane;
      const bool any = __ballot(hit(h, j, j < e)) != 0;
      if (any) {
        if (lane == h) found = true;
        break;
      }
    }
    pend &= pend - 1;
  }
  return found;
}

__device__ inline uint32_t key_of(uint32_t v, uint32_t r) { return ~rotr32(v, r); }
__device__ inline uint32_t root_of(uint32_t c, uint32_t r) { return rotl32(~c, r); }

// graph index of vertex v (vtx_off null: one graph) by binary search in [lo, hi]
__device__ inline uint32_t graph_of(const uint32_t* vtx_off, uint32_t lo, uint32_t hi, uint32_t v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (vtx_off[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_g_clear(GCtl* ctl, unsigned long long* pkey, uint32_t* gmin, uint32_t count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < sizeof(GCtl) / 4) reinterpret_cast<uint32_t*>(ctl)[t] = 0;
  for (uint32_t i = t; i < count; i += gridDim.x * blockDim.x) {
    pkey[i] = 0;
    gmin[i] = 0xFFFFFFFFu;
  }
}

// trim_1_first and the pivot keys: one wavefront per 64 consecutive
// vertices (it writes their two bitmap words), one vertex per lane.
// pc[v] = graph index << 32 (colour 0).
__global__ __launch_bounds__(WG) void k_g_init(uint32_t n, const uint64_t* __restrict__ fw_off,
                                               const uint64_t* __restrict__ bw_off, const uint32_t* __restrict__ vtx_off,
                                               uint32_t count, uint32_t* __restrict__ live, uint32_t* __restrict__ fwv,
                                               uint32_t* __restrict__ bwv, uint32_t* __restrict__ labels,
                                               unsigned long long* __restrict__ pc, unsigned long long* pkey,
                                               GCtl* ctl) {
  __shared__ uint32_t s_g0, s_g1;
  __shared__ uint32_t s_rg[WG / WAVE];
  __shared__ unsigned long long s_rm[WG / WAVE];
  const uint32_t lane = lane_id();
  const uint64_t words64 = ((uint64_t)n + 63) / 64;  // 64-vertex groups
  uint32_t dead = 0, alive = 0;
  // the wavefront's running maximum key of one graph (wave-uniform), written
  // out when the graph changes and, combined per workgroup, at the end
  uint32_t run_g = 0xFFFFFFFFu;
  unsigned long long run_m = 0;
  for (uint64_t gbase = (uint64_t)blockIdx.x * (WG / WAVE); gbase < words64; gbase += (uint64_t)gridDim.x * (WG / WAVE)) {
    // graphs covering this workgroup's 256 vertices
    const uint64_t first = gbase * 64, last = std::min<uint64_t>(first + WG, n) - 1;
    if (threadIdx.x == 0) {
      s_g0 = vtx_off ? graph_of(vtx_off, 0, count - 1, (uint32_t)first) : 0u;
      s_g1 = vtx_off ? graph_of(vtx_off, 0, count - 1, (uint32_t)last) : 0u;
    }
    __syncthreads();
    const uint64_t grp = gbase + threadIdx.x / WAVE;
    const uint64_t v64 = grp * 64 + lane;
    bool is_live = false;
    uint32_t g = 0;
    unsigned long long key = 0;
    if (grp < words64 && v64 < n) {
      const uint32_t v = (uint32_t)v64;
      const uint64_t out = fw_off[v + 1] - fw_off[v], in = bw_off[v + 1] - bw_off[v];
      is_live = in != 0 && out != 0;
      g = s_g0 == s_g1 ? s_g0 : graph_of(vtx_off, s_g0, s_g1, v);
      labels[v] = is_live ? UNDEC : v;
      pc[v] = (unsigned long long)g << 32;
      if (is_live) {
        const uint64_t p = in * out;
        key = ((unsigned long long)std::min<uint64_t>(p, 0xFFFFFFFFull) << 32) | (~v & 0xFFFFFFFFu);
      }
      dead += !is_live;
      alive += is_live;
    }
    if (grp < words64) {
      const uint64_t mask = __ballot(is_live);
      if (lane == 0 || lane == 32) {
        const uint64_t w = grp * 2 + (lane >> 5);
        live[w] = (uint32_t)(mask >> lane);
        fwv[w] = 0;
        bwv[w] = 0;
      }
      // the wavefront's largest key per graph: one atomic when all its lanes share a graph
      const uint32_t g0 = __shfl(g, 0);
      const bool same = __ballot(grp < words64 && v64 < n && g != g0) == 0;
      if (same) {
        unsigned long long m = key;
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long t = __shfl_xor(m, o);
          m = t > m ? t : m;
        }
        if (g0 != run_g) {
          if (lane == 0 && run_m) atomicMax(&pkey[run_g], run_m);
          run_g = g0;
          run_m = 0;
        }
        run_m = m > run_m ? m : run_m;
      } else if (key) {
        atomicMax(&pkey[g], key);
      }
    }
    __syncthreads();
  }
  const uint32_t wid = threadIdx.x / WAVE;
  if (lane == 0) {
    s_rg[wid] = run_m ? run_g : 0xFFFFFFFFu;
    s_rm[wid] = run_m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // the waves' maxima, one atomic per distinct graph
    for (uint32_t a = 0; a < WG / WAVE; ++a) {
      if (s_rg[a] == 0xFFFFFFFFu) continue;
      unsigned long long m = s_rm[a];
      for (uint32_t b = a + 1; b < WG / WAVE; ++b)
        if (s_rg[b] == s_rg[a]) {
          m = s_rm[b] > m ? s_rm[b] : m;
          s_rg[b] = 0xFFFFFFFFu;
        }
      atomicMax(&pkey[s_rg[a]], m);
    }
  }
  block_add(dead, &ctl->dec_trim);
  block_add(alive, &ctl->live_after_trim);
}

// appends the row slices of v (deg edges) to queue q: every lane with `take`
// calls together; one atomic per wavefront. Entries are v << 32 | slice.
__device__ inline void append_slices(bool take, uint32_t v, uint64_t deg, unsigned long long* q, uint32_t* qn,
                                     uint64_t qcap, uint32_t* err) {
  if (!__ballot(take && deg)) return;
  const uint32_t lane = lane_id();
  const uint32_t k = take ? (uint32_t)((deg + SLICE - 1) / SLICE) : 0u;
  const uint32_t incl = wscan_incl(k);
  const uint32_t total = __shfl(incl, WAVE - 1);
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(qn, total);
  base = __shfl(base, 0);
  if (take) {
    const uint64_t b = (uint64_t)base + incl - k;
    if (b + k > qcap) {
      atomicOr(err, ERR_QUEUE);
    } else {
      for (uint32_t s = 0; s < k; ++s) q[b + s] = ((unsigned long long)v << 32) | s;
    }
  }
}

// A wavefront-private LDS buffer of queue entries: appends collect there and
// go out with one atomic per flush, not one per call (every wave appending
// to one counter would serialise on it).
constexpr uint32_t STAGE = 256;
struct Stage {
  unsigned long long* buf;  // STAGE entries (LDS)
  uint32_t n;               // entries held (uniform across the wavefront)
};

__device__ inline void stage_flush(Stage& st, unsigned long long* q, uint32_t* qn, uint64_t qcap, uint32_t* err) {
  if (st.n == 0) return;
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(qn, st.n);
  base = __shfl(base, 0);
  if ((uint64_t)base + st.n > qcap) {
    if (lane == 0) atomicOr(err, ERR_QUEUE);
  } else {
    for (uint32_t i = lane; i < st.n; i += WAVE) q[(uint64_t)base + i] = st.buf[i];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  st.n = 0;
}

// append_slices through the stage: every lane calls together
__device__ inline void stage_slices(Stage& st, bool take, uint32_t v, uint64_t deg, unsigned long long* q, uint32_t* qn,
                                    uint64_t qcap, uint32_t* err) {
  if (!__ballot(take && deg)) return;
  const uint32_t k = take ? (uint32_t)((deg + SLICE - 1) / SLICE) : 0u;
  const uint32_t incl = wscan_incl(k);
  const uint32_t total = __shfl(incl, WAVE - 1);
  if (total > STAGE) {  // a hub's many slices: straight to the queue
    stage_flush(st, q, qn, qcap, err);
    append_slices(take, v, deg, q, qn, qcap, err);
    return;
  }
  if (st.n + total > STAGE) stage_flush(st, q, qn, qcap, err);
  const uint32_t b = st.n + incl - k;
  for (uint32_t s = 0; s < k; ++s) st.buf[b + s] = ((unsigned long long)v << 32) | s;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  st.n += total;
}

// the pivots of every graph start a reach: their bit in vis, their slices in q
// (ring slot 0), their count and out-degree sum in slot 0 of the counters
template <int DIR>
__global__ void k_g_seed(uint32_t count, const unsigned long long* __restrict__ pkey,
                         const uint64_t* __restrict__ noff, uint32_t* vis, unsigned long long* q, uint64_t qcap,
                         GCtl* ctl) {
  for (uint32_t g0 = blockIdx.x * WG; g0 < count; g0 += gridDim.x * WG) {
    const uint32_t g = g0 + threadIdx.x;
    bool take = false;
    uint32_t p = 0;
    uint64_t deg = 0;
    if (g < count && pkey[g]) {
      p = ~(uint32_t)pkey[g];
      atomicOr(&vis[p >> 5], 1u << (p & 31));
      deg = noff[p + 1] - noff[p];
      take = true;
      atomicAdd(&ctl->qv[0], 1u);
      atomicAdd(&ctl->qe[0], (unsigned long long)deg);
    }
    append_slices(take, p, deg, q, &ctl->qn[0], qcap, &ctl->err);
  }
}

// Persistent reach. DIR 0: forward over fw rows from live vertices (cand =
// live), visited = fwv; DIR 1: backward over bw rows inside the forward sets
// (cand = fwv), visited = bwv. Frontier entries are 256-edge row slices in
// two queues (level parity); counters in the GCtl ring (level % 3). A
// level is push (wavefronts take 64 entries from a cursor) or pull (every
// candidate scans its reverse row for a visited neighbour); the switch is
// Beamer's rule on the frontier's out-edges and size, evaluated by every
// workgroup on the same counters. Claimed vertices go to the next queue with
// their slices. Ends when a level claims nothing.
template <int DIR>
__global__ __launch_bounds__(GWG) void k_g_reach(uint32_t n, uint64_t m, const uint64_t* __restrict__ off,
                                                 const uint32_t* __restrict__ adj, const uint64_t* __restrict__ roff,
                                                 const uint32_t* __restrict__ radj, const uint32_t* __restrict__ cand,
                                                 uint32_t* vis, unsigned long long* q0, unsigned long long* q1,
                                                 uint64_t qcap, GCtl* ctl, float alpha, float beta) {
  __shared__ uint64_t s_beg[GWAVES][WAVE];
  __shared__ uint32_t s_ex[GWAVES][WAVE];
  __shared__ unsigned long long s_stage[GWAVES][STAGE];
  uint32_t* bar = &ctl->bar[DIR];
  uint32_t k = 0;  // barriers passed
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t gw = blockIdx.x * GWAVES + wid, waves = gridDim.x * GWAVES;
  const uint64_t words = ((uint64_t)n + 31) / 32;
  const uint32_t live_total = ctl->live_after_trim;  // written before the launch
  bool pull = false;
  unsigned long long expanded = 0;
  Stage st{s_stage[wid], 0};
  for (uint32_t L = 0;; ++L) {
    const uint32_t s0 = L % 3, s1 = (L + 1) % 3, s2 = (L + 2) % 3;
    unsigned long long* Q = (L & 1) ? q1 : q0;
    unsigned long long* N = (L & 1) ? q0 : q1;
    const uint32_t fv = ctl->qv[s0];
    // entries past the queue were never written (ERR_QUEUE is set): never read them
    const uint32_t fn = (uint32_t)std::min<uint64_t>(ctl->qn[s0], qcap);
    const unsigned long long fe = ctl->qe[s0];
    if (fv == 0) break;
    // Beamer's rule (iSpan's alpha and beta), on values every workgroup reads alike
    const double rest = (double)m - (double)expanded;
    if (!pull && (double)fe * alpha > rest && fv > 64) pull = true;
    else if (pull && (double)fv * beta < (double)live_total) pull = false;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      ctl->qn[s2] = 0;
      ctl->qv[s2] = 0;
      ctl->qe[s2] = 0;
      ctl->cur[s2] = 0;
      if (pull) ++ctl->levels_pull[DIR]; else ++ctl->levels_push[DIR];
    }
    uint32_t claimed = 0;
    unsigned long long cdeg = 0;
    if (!pull) {
      expanded += fe;
      // push: a wavefront takes 64 entries (one per lane), scans their edge
      // counts and walks the concatenated edges, four per lane per step with
      // the loads of the four issued together; each edge finds its entry by
      // a binary search over the wavefront's scan in LDS
      uint64_t* sb = s_beg[wid];
      uint32_t* sx = s_ex[wid];
      // 64-entry chunks in a fixed order (entries are row slices of at most
      // SLICE edges, so chunks cost about the same)
      for (uint64_t u0 = (uint64_t)gw * WAVE; u0 < fn; u0 += (uint64_t)waves * WAVE) {
        const uint64_t i = u0 + lane;
        uint64_t b = 0;
        uint32_t cnt = 0;
        if (i < fn) {
          const unsigned long long e = Q[i];
          const uint32_t v = (uint32_t)(e >> 32), sl = (uint32_t)e;
          const uint64_t re = off[v + 1];
          b = off[v] + (uint64_t)sl * SLICE;
          cnt = (uint32_t)(std::min<uint64_t>(b + SLICE, re) - b);
        }
        const uint32_t incl = wscan_incl(cnt);
        const uint32_t total = __shfl(incl, WAVE - 1);
        sb[lane] = b;
        sx[lane] = incl - cnt;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t base = 0; base < total; base += 4 * WAVE) {
          uint32_t w[4];
          bool ok[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t e = base + q * WAVE + lane;
            ok[q] = e < total;
            w[q] = 0;
            if (ok[q]) {
              uint32_t lo = 0, hi = WAVE;  // last entry whose scan is <= e
              while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (sx[mid] <= e) lo = mid; else hi = mid;
              }
              w[q] = adj[sb[lo] + (e - sx[lo])];
            }
          }
          uint32_t cw[4], vw[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            cw[q] = ok[q] ? cand[w[q] >> 5] : 0u;
            vw[q] = ok[q] ? vis[w[q] >> 5] : 0xFFFFFFFFu;
          }
          bool take[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t bit = 1u << (w[q] & 31);
            take[q] = (cw[q] & bit) && !(vw[q] & bit);
            if (take[q]) take[q] = (atomicOr(&vis[w[q] >> 5], bit) & bit) == 0;
          }
          uint64_t d0[4], d1[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            d0[q] = take[q] ? off[w[q]] : 0;
            d1[q] = take[q] ? off[w[q] + 1] : 0;
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint64_t dw = d1[q] - d0[q];
            claimed += take[q];
            cdeg += dw;
            stage_slices(st, take[q], w[q], dw, N, &ctl->qn[s1], qcap, &ctl->err);
          }
        }
        __builtin_amdgcn_wave_barrier();  // sb/sx are rewritten by the next grab
      }
    } else {
      // pull: 256 vertices per wavefront and step, four per lane (one in each
      // 64-vertex group, whose two visited words the wavefront owns), their
      // loads issued together: candidate words, row offsets, first in-neighbours
      constexpr int PG = 4;
      for (uint64_t blk = gw; blk * 2 * PG < words; blk += waves) {
        bool cnd[PG], take[PG];
        uint32_t vv[PG], u0[PG];
        uint64_t rb[PG], re[PG];
#pragma unroll
        for (int g = 0; g < PG; ++g) {
          const uint64_t v64 = (blk * PG + g) * 64 + lane;
          vv[g] = (uint32_t)v64;
          const uint32_t bit = 1u << (v64 & 31);
          cnd[g] = v64 < n && (cand[v64 >> 5] & bit) && !(vis[v64 >> 5] & bit);
          take[g] = false;
        }
#pragma unroll
        for (int g = 0; g < PG; ++g) {
          rb[g] = cnd[g] ? roff[vv[g]] : 0;
          re[g] = cnd[g] ? roff[vv[g] + 1] : 0;
        }
#pragma unroll
        for (int g = 0; g < PG; ++g) u0[g] = rb[g] < re[g] ? radj[rb[g]] : 0u;
#pragma unroll
        for (int g = 0; g < PG; ++g) {
          if (rb[g] < re[g]) {
            take[g] = (vis[u0[g] >> 5] >> (u0[g] & 31)) & 1u;
            ++rb[g];
          }
        }
#pragma unroll
        for (int g = 0; g < PG; ++g) {
          const bool heavy = cnd[g] && !take[g] && re[g] - rb[g] > HEAVY;
          if (cnd[g] && !take[g] && !heavy) {
            for (uint64_t it = rb[g]; it < re[g]; ++it) {
              const uint32_t u = radj[it];
              if (vis[u >> 5] & (1u << (u & 31))) {
                take[g] = true;
                break;
              }
            }
          }
          // long rows: the wavefront scans them one at a time
          const bool got = heavy_rows(heavy, rb[g], re[g], [&](uint32_t, uint64_t j, bool valid) {
            if (!valid) return false;
            const uint32_t u = radj[j];
            return (vis[u >> 5] & (1u << (u & 31))) != 0;
          });
          take[g] = take[g] || got;
        }
#pragma unroll
        for (int g = 0; g < PG; ++g) {
          const uint64_t mask = __ballot(take[g]);
          if (mask && (lane == 0 || lane == 32)) {
            const uint32_t bits = (uint32_t)(mask >> lane);
            if (bits) atomicOr(&vis[(blk * PG + g) * 2 + (lane >> 5)], bits);
          }
          const uint64_t dv = take[g] ? off[vv[g] + 1] - off[vv[g]] : 0;
          claimed += take[g];
          cdeg += dv;
          stage_slices(st, take[g], vv[g], dv, N, &ctl->qn[s1], qcap, &ctl->err);
        }
      }
    }
    stage_flush(st, N, &ctl->qn[s1], qcap, &ctl->err);
    block_add2(claimed, cdeg, &ctl->qv[s1], &ctl->qe[s1]);
    if (!gbar(bar, &ctl->err, ++k)) break;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->expanded[DIR] = expanded;
}

// the pivots' SCCs: fwv & bwv. Smallest member per graph first, then labels.
__global__ __launch_bounds__(WG) void k_g_scc_min(uint32_t n, const uint32_t* __restrict__ fwv,
                                                  const uint32_t* __restrict__ bwv,
                                                  const unsigned long long* __restrict__ pc, uint32_t* gmin) {
  // lanes of one graph take their minimum first, and a wavefront keeps a
  // running minimum while its graph stays the same: one atomic per wavefront
  // and graph change
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  const uint32_t lane = lane_id();
  uint32_t run_g = 0xFFFFFFFFu, run_m = 0xFFFFFFFFu;  // wave-uniform
  for (uint64_t v64 = blockIdx.x * (uint64_t)WG + threadIdx.x; v64 < pad; v64 += (uint64_t)gridDim.x * WG) {
    const uint32_t v = (uint32_t)v64, bit = 1u << (v & 31);
    bool in = v64 < n && (fwv[v >> 5] & bwv[v >> 5] & bit) != 0;
    const uint32_t g = in ? (uint32_t)(pc[v] >> 32) : 0u;
    uint64_t pend = __ballot(in);
    while (pend) {
      const uint32_t leader = __ffsll((unsigned long long)pend) - 1;
      const uint32_t gl = __shfl(g, leader);
      const bool mine = in && g == gl;
      const uint32_t m = __shfl(wave_min(mine ? v : 0xFFFFFFFFu), 0);
      if (gl != run_g) {
        if (lane == 0 && run_m != 0xFFFFFFFFu) atomicMin(&gmin[run_g], run_m);
        run_g = gl;
        run_m = 0xFFFFFFFFu;
      }
      run_m = m < run_m ? m : run_m;
      pend &= ~__ballot(mine);
      in = in && !mine;
    }
  }
  if (lane == 0 && run_m != 0xFFFFFFFFu) atomicMin(&gmin[run_g], run_m);
}

__global__ __launch_bounds__(WG) void k_g_scc_label(uint32_t n, const uint32_t* __restrict__ fwv,
                                                    const uint32_t* __restrict__ bwv, uint32_t* live,
                                                    unsigned long long* pc, const uint32_t* __restrict__ gmin,
                                                    uint32_t* __restrict__ labels, GCtl* ctl) {
  const uint32_t lane = lane_id();
  const uint64_t groups = ((uint64_t)n + 63) / 64;
  uint32_t dec = 0;
  for (uint64_t grp = (blockIdx.x * (uint64_t)WG + threadIdx.x) / WAVE; grp < groups;
       grp += (uint64_t)gridDim.x * (WG / WAVE)) {
    const uint64_t v64 = grp * 64 + lane;
    bool in = false;
    if (v64 < n) {
      const uint32_t v = (uint32_t)v64, bit = 1u << (v & 31);
      in = (fwv[v >> 5] & bwv[v >> 5] & bit) != 0;
      if (in) {
        labels[v] = gmin[(uint32_t)(pc[v] >> 32)];
        pc[v] &= ~0xFFFFFFFFull;  // colour 0: no longer propagates
      }
    }
    const uint64_t mask = __ballot(in);
    if (mask && (lane == 0 || lane == 32)) {
      const uint32_t bits = (uint32_t)(mask >> lane);
      if (bits) live[grp * 2 + (lane >> 5)] &= ~bits;  // the wavefront owns these words
    }
    dec += in;
  }
  block_add(dec, &ctl->dec_fwbw);
}

// Persistent colour rounds until no vertex is live (or max_rounds). pc[v]
// holds partition << 32 | colour; dec (a bitmap) marks the round's decided
// vertices; U lists the round's live vertices; minid[root] collects the
// smallest member of each root's SCC.
__global__ __launch_bounds__(GWG) void k_g_colour(uint32_t n, const uint64_t* __restrict__ fw_off,
                                                  const uint32_t* __restrict__ fw_adj,
                                                  const uint64_t* __restrict__ bw_off,
                                                  const uint32_t* __restrict__ bw_adj, uint32_t* live, uint32_t* dec,
                                                  unsigned long long* pc, uint32_t* minid, uint32_t* U,
                                                  uint32_t* __restrict__ labels, GCtl* ctl, uint32_t max_rounds) {
  uint32_t* bar = &ctl->bar[2];
  uint32_t k = 0;
  const uint32_t lane = lane_id();
  const uint64_t T = (uint64_t)gridDim.x * GWG;
  const uint64_t tid = (uint64_t)blockIdx.x * GWG + threadIdx.x;
  const uint64_t groups = ((uint64_t)n + 63) / 64;
  const uint32_t gw = blockIdx.x * GWAVES + threadIdx.x / WAVE, waves = gridDim.x * GWAVES;
  uint32_t fix = 0;  // fixpoint flags used so far (ring of three)
  for (uint32_t r = 1;; ++r) {
    // (a) list the live vertices (each wavefront a contiguous run of 64-vertex
    // groups: count, one atomic, then write), clear dec
    if (tid == 0) ctl->ucount = 0;
    if (!gbar(bar, &ctl->err, ++k)) return;
    {
      const uint64_t per = (groups + waves - 1) / waves;
      const uint64_t g0 = (uint64_t)gw * per, g1 = std::min<uint64_t>(g0 + per, groups);
      uint32_t cnt = 0;
      for (uint64_t grp = g0 + lane; grp < g1; grp += WAVE) cnt += __popc(live[grp * 2]) + __popc(live[grp * 2 + 1]);
      cnt = wave_sum(cnt);
      uint32_t base = 0;
      if (lane == 0 && cnt) base = atomicAdd(&ctl->ucount, cnt);
      base = __shfl(base, 0);
      for (uint64_t gb = g0; gb < g1; gb += WAVE) {
        const uint64_t grp = gb + lane;
        uint64_t bits = 0;
        if (grp < g1) {
          bits = ((uint64_t)live[grp * 2 + 1] << 32) | live[grp * 2];
          dec[grp * 2] = 0;
          dec[grp * 2 + 1] = 0;
        }
        const uint32_t c = (uint32_t)__popcll(bits);
        const uint32_t incl = wscan_incl(c);
        uint32_t pos = base + incl - c;
        while (bits) {
          const uint32_t b = __ffsll((unsigned long long)bits) - 1;
          U[pos++] = (uint32_t)(grp * 64 + b);
          bits &= bits - 1;
        }
        base += __shfl(incl, WAVE - 1);
      }
    }
    if (!gbar(bar, &ctl->err, ++k)) return;
    const uint32_t u = ctl->ucount;
    if (u == 0) {
      if (tid == 0) {
        ctl->done = 1;
        ctl->rounds = r - 1;
      }
      return;
    }
    if (max_rounds && r > max_rounds) {
      if (tid == 0) {
        ctl->incomplete = 1;
        ctl->rounds = r - 1;
      }
      return;
    }
    // (b) keys
    for (uint64_t i = tid; i < u; i += T) {
      const uint32_t v = U[i];
      pc[v] = (pc[v] & ~0xFFFFFFFFull) | key_of(v, r);
    }
    // (c) max colour along in-edges inside the partition, to a fixpoint
    for (;;) {
      uint32_t* fl = &ctl->changed[fix % 3];
      if (tid == 0) ctl->changed[(fix + 1) % 3] = 0;
      if (!gbar(bar, &ctl->err, ++k)) return;
      bool ch = false;
      for (uint64_t i0 = (uint64_t)gw * WAVE; i0 < u; i0 += (uint64_t)waves * WAVE) {
        const uint64_t i = i0 + lane;
        uint32_t v = 0, part = 0, c0 = 0, c = 0;
        uint64_t rb = 0, re = 0;
        bool heavy = false;
        if (i < u) {
          v = U[i];
          const unsigned long long p = pc[v];
          part = (uint32_t)(p >> 32);
          c0 = c = (uint32_t)p;
          rb = bw_off[v];
          re = bw_off[v + 1];
          heavy = re - rb > HEAVY;
          if (!heavy)
            for (uint64_t it = rb; it < re; ++it) {
              const unsigned long long x = pc[bw_adj[it]];
              if ((uint32_t)(x >> 32) == part) c = max(c, (uint32_t)x);
            }
        }
        uint64_t pend = __ballot(heavy);
        while (pend) {
          const uint32_t h = __ffsll((unsigned long long)pend) - 1;
          const uint64_t b = __shfl(rb, h), e = __shfl(re, h);
          const uint32_t hp = __shfl(part, h);
          uint32_t m = 0;
          for (uint64_t j = b + lane; j < e; j += WAVE) {
            const unsigned long long x = pc[bw_adj[j]];
            if ((uint32_t)(x >> 32) == hp) m = max(m, (uint32_t)x);
          }
          for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o));
          if (lane == h) c = max(c, m);
          pend &= pend - 1;
        }
        if (i < u && c > c0) {
          pc[v] = ((unsigned long long)part << 32) | c;  // only v's lane writes pc[v]
          ch = true;
        }
      }
      if (__syncthreads_or(ch) && threadIdx.x == 0) atomicOr(fl, 1u);
      if (!gbar(bar, &ctl->err, ++k)) return;
      const bool again = *fl != 0;
      ++fix;
      if (!again) break;
    }
    // (d) roots
    for (uint64_t i = tid; i < u; i += T) {
      const uint32_t v = U[i];
      if ((uint32_t)pc[v] == key_of(v, r)) {
        atomicOr(&dec[v >> 5], 1u << (v & 31));
        minid[v] = v;
      }
    }
    // (e) backward closure within a colour over out-edges, to a fixpoint
    for (;;) {
      uint32_t* fl = &ctl->changed[fix % 3];
      if (tid == 0) ctl->changed[(fix + 1) % 3] = 0;
      if (!gbar(bar, &ctl->err, ++k)) return;
      bool ch = false;
      for (uint64_t i0 = (uint64_t)gw * WAVE; i0 < u; i0 += (uint64_t)waves * WAVE) {
        const uint64_t i = i0 + lane;
        uint32_t v = 0, c = 0;
        uint64_t rb = 0, re = 0;
        bool heavy = false, hit = false;
        if (i < u) {
          v = U[i];
          if (!(dec[v >> 5] & (1u << (v & 31)))) {
            c = (uint32_t)pc[v];
            rb = fw_off[v];
            re = fw_off[v + 1];
            heavy = re - rb > HEAVY;
            if (!heavy)
              for (uint64_t it = rb; it < re; ++it) {
                const uint32_t w = fw_adj[it];
                if ((dec[w >> 5] & (1u << (w & 31))) && (uint32_t)pc[w] == c) {
                  hit = true;
                  break;
                }
              }
          }
        }
        const bool got = heavy_rows(heavy, rb, re, [&](uint32_t h, uint64_t j, bool valid) {
          const uint32_t hc = __shfl(c, h);  // every lane calls: the shuffle is uniform
          if (!valid) return false;
          const uint32_t w = fw_adj[j];
          return (dec[w >> 5] & (1u << (w & 31))) && (uint32_t)pc[w] == hc;
        });
        hit = hit || got;
        if (hit) {
          atomicOr(&dec[v >> 5], 1u << (v & 31));
          ch = true;
        }
      }
      if (__syncthreads_or(ch) && threadIdx.x == 0) atomicOr(fl, 1u);
      if (!gbar(bar, &ctl->err, ++k)) return;
      const bool again = *fl != 0;
      ++fix;
      if (!again) break;
    }
    // (f) smallest member per root
    for (uint64_t i = tid; i < u; i += T) {
      const uint32_t v = U[i];
      if (dec[v >> 5] & (1u << (v & 31))) atomicMin(&minid[root_of((uint32_t)pc[v], r)], v);
    }
    if (!gbar(bar, &ctl->err, ++k)) return;
    // (g) labels, and the refined partition (the colour's root) of the rest
    uint32_t d = 0;
    for (uint64_t i = tid; i < u; i += T) {
      const uint32_t v = U[i];
      const uint32_t root = root_of((uint32_t)pc[v], r);
      if (dec[v >> 5] & (1u << (v & 31))) {
        labels[v] = minid[root];
        pc[v] = 0;
        atomicAnd(&live[v >> 5], ~(1u << (v & 31)));
        ++d;
      } else {
        pc[v] = (unsigned long long)root << 32;
      }
    }
    block_add(d, &ctl->dec_color);
  }
}

// the ring counters of a reach back to zero, before the backward reach is seeded
__global__ void k_g_clear_ring(GCtl* ctl) {
  if (threadIdx.x < 3) {
    ctl->qn[threadIdx.x] = 0;
    ctl->qv[threadIdx.x] = 0;
    ctl->cur[threadIdx.x] = 0;
    ctl->qe[threadIdx.x] = 0;
  }
}

__global__ void k_g_localize(uint32_t count, const uint32_t* __restrict__ vtx_off, uint32_t* __restrict__ labels) {
  for (uint32_t g = blockIdx.x; g < count; g += gridDim.x) {
    const uint32_t b = vtx_off[g], e = vtx_off[g + 1];
    for (uint32_t v = b + threadIdx.x; v < e; v += WG) labels[v] -= b;
  }
}

}  // namespace gs

// ---------------------------------------------------------------- host
namespace {

template <class T>
int galloc(T** p, uint64_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc((void**)p, std::max<uint64_t>(count, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    note_error("graph solve", "hipMalloc failed", __FILE__, __LINE__);
    return KSPN_E_OOM;
  }
  return KSPN_OK;
}

uint32_t sgrid(Ctx* c, uint64_t items, uint32_t per = WG) {
  const uint64_t g = (items + per - 1) / per;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)c->num_cu * 16));
}

// the launch sequence; nothing in it reads device memory on the host
int g_launch(Ctx* c, const kspn_graph* g, uint32_t* labels, const uint32_t* vtx_off, uint32_t count,
             const kspn_opts& o, hipStream_t s) {
  GraphMode& G = c->gm;
  const uint32_t n = g->n;
  gs::GCtl* ctl = reinterpret_cast<gs::GCtl*>(G.ctl);
  const float alpha = o.alpha > 0 ? o.alpha : 14.0f, beta = o.beta > 0 ? o.beta : 24.0f;
  // a barrier costs more with more workgroups: small graphs get one per 8192 vertices (at least 8)
  const uint32_t pgrid = (uint32_t)std::min<uint64_t>(G.grid, std::max<uint64_t>(8, ((uint64_t)n + 8191) / 8192));
  gs::k_g_clear<<<sgrid(c, std::max<uint64_t>(count, 64)), WG, 0, s>>>(ctl, G.pkey, G.gmin, count);
  gs::k_g_init<<<sgrid(c, ((uint64_t)n + 63) / 64 * 64), WG, 0, s>>>(n, g->fw_off, g->bw_off, vtx_off, count, c->live,
                                                                      c->fwv, c->bwv, labels, G.pc, G.pkey, ctl);
  gs::k_g_seed<0><<<sgrid(c, count), WG, 0, s>>>(count, G.pkey, g->fw_off, c->fwv, G.q0, G.qcap, ctl);
  gs::k_g_reach<0><<<pgrid, gs::GWG, 0, s>>>(n, g->m, g->fw_off, g->fw_adj, g->bw_off, g->bw_adj, c->live, c->fwv,
                                             G.q0, G.q1, G.qcap, ctl, alpha, beta);
  // the backward reach starts from ring slot 0 again: k_g_seed<1> runs after the forward reach cleared
  // nothing of slot 0 it needs (its counters are reset here)
  gs::k_g_clear_ring<<<1, 64, 0, s>>>(ctl);
  gs::k_g_seed<1><<<sgrid(c, count), WG, 0, s>>>(count, G.pkey, g->bw_off, c->bwv, G.q0, G.qcap, ctl);
  gs::k_g_reach<1><<<pgrid, gs::GWG, 0, s>>>(n, g->m, g->bw_off, g->bw_adj, g->fw_off, g->fw_adj, c->fwv, c->bwv,
                                             G.q0, G.q1, G.qcap, ctl, alpha, beta);
  gs::k_g_scc_min<<<sgrid(c, n), WG, 0, s>>>(n, c->fwv, c->bwv, G.pc, G.gmin);
  gs::k_g_scc_label<<<sgrid(c, ((uint64_t)n + 63) / 64 * 64), WG, 0, s>>>(n, c->fwv, c->bwv, c->live, G.pc, G.gmin,
                                                                          labels, ctl);
  const uint32_t cap = o.max_rounds ? o.max_rounds : 0;
  gs::k_g_colour<<<pgrid, gs::GWG, 0, s>>>(n, g->fw_off, g->fw_adj, g->bw_off, g->bw_adj, c->live, c->fwv, G.pc, c->col,
                                           c->ulist, labels, ctl, cap);
  if (vtx_off) gs::k_g_localize<<<std::min<uint32_t>(count, 16384u), WG, 0, s>>>(count, vtx_off, labels);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}


// buffers for n vertices, m edges and count graphs (grown, never shrunk)
int g_reserve(Ctx* c, uint32_t n, uint64_t m, uint32_t count) {
  GraphMode& G = c->gm;
  if (!G.ctl) {
    KSPN_HIP(hipMalloc(&G.ctl, sizeof(gs::GCtl)));
    KSPN_HIP(hipHostMalloc(&G.h_status, sizeof(gs::GCtl), hipHostMallocDefault));
    KSPN_HIP(hipEventCreate(&G.ev[0]));
    KSPN_HIP(hipEventCreate(&G.ev[1]));
    KSPN_HIP(hipStreamCreateWithFlags(&G.cap_stream, hipStreamNonBlocking));
  }
  if (!G.grid) {
    // one workgroup of GWG threads per CU, all resident together (the barriers give up otherwise)
    int per_cu = 1 << 30;
    for (const void* fn : {reinterpret_cast<const void*>(&gs::k_g_reach<0>),
                           reinterpret_cast<const void*>(&gs::k_g_reach<1>),
                           reinterpret_cast<const void*>(&gs::k_g_colour)}) {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, gs::GWG, 0) != hipSuccess) b = 0;
      per_cu = std::min(per_cu, b);
    }
    if (per_cu < 1) {
      note_error("graph solve", "the persistent kernels cannot be resident together", __FILE__, __LINE__);
      return KSPN_E_UNIMPLEMENTED;
    }
    G.grid = (uint32_t)c->num_cu * (uint32_t)std::min(per_cu, 2);
  }
  if (n > G.cap_n) {
    KSPN_TRY(galloc(&G.pc, (uint64_t)n + 64));
    G.cap_n = n;
  }
  const uint64_t qneed = (uint64_t)n + m / gs::SLICE + count + 64;
  if (qneed > G.qcap || m > G.cap_m) {
    KSPN_TRY(galloc(&G.q0, qneed));
    KSPN_TRY(galloc(&G.q1, qneed));
    G.qcap = qneed;
    G.cap_m = m;
  }
  if (count > G.cap_count) {
    KSPN_TRY(galloc(&G.pkey, count));
    KSPN_TRY(galloc(&G.gmin, count));
    G.cap_count = count;
  }
  return KSPN_OK;
}

}  // namespace

void graph_release(Ctx* c) {
  GraphMode& G = c->gm;
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  for (void* p : {(void*)G.ctl, (void*)G.pc, (void*)G.q0, (void*)G.q1, (void*)G.pkey, (void*)G.gmin})
    if (p) (void)hipFree(p);
  if (G.h_status) (void)hipHostFree(G.h_status);
  for (hipEvent_t e : G.ev)
    if (e) (void)hipEventDestroy(e);
  if (G.cap_stream) (void)hipStreamDestroy(G.cap_stream);
  G = GraphMode{};
}

int graph_solve(Ctx* c, const kspn_graph* g, uint32_t* labels, const kspn_opts& o, const uint32_t* vtx_off,
                uint32_t count, kspn_stats* st, hipStream_t s) {
  if (o.pivot != KSPN_PIVOT_PLAIN || (o.method != KSPN_METHOD_KASPAN && o.method != KSPN_METHOD_HPC_LIKE)) {
    note_error("graph solve", "KSPN_FLAG_GRAPH runs the plain pivot and the kaspan / hpc_like stages", __FILE__,
               __LINE__);
    return KSPN_E_UNIMPLEMENTED;
  }
  if (g->n == 0) return KSPN_OK;
  if (!vtx_off) count = 1;
  KSPN_TRY(g_reserve(c, g->n, g->m, count));
  GraphMode& G = c->gm;
  // everything the captured launches depend on
  float alpha = o.alpha > 0 ? o.alpha : 14.0f, beta = o.beta > 0 ? o.beta : 24.0f;
  uint32_t ab[2];
  std::memcpy(&ab[0], &alpha, 4);
  std::memcpy(&ab[1], &beta, 4);
  const uint64_t key[16] = {(uint64_t)g->n, g->m, (uint64_t)g->fw_off, (uint64_t)g->fw_adj, (uint64_t)g->bw_off,
                            (uint64_t)g->bw_adj, (uint64_t)labels, (uint64_t)vtx_off, count,
                            ((uint64_t)ab[0] << 32) | ab[1], o.max_rounds, (uint64_t)c->live ^ ((uint64_t)c->fwv << 1),
                            (uint64_t)c->bwv ^ ((uint64_t)c->col << 1), (uint64_t)c->ulist ^ ((uint64_t)G.pc << 1),
                            (uint64_t)G.q0 ^ ((uint64_t)G.q1 << 1) ^ G.qcap,
                            (uint64_t)G.pkey ^ ((uint64_t)G.gmin << 1) ^ (uint64_t)G.ctl};
  if (!G.exec || std::memcmp(key, G.key, sizeof(key)) != 0) {
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.exec = nullptr;
    G.graph = nullptr;
    KSPN_HIP(hipStreamBeginCapture(G.cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = g_launch(c, g, labels, vtx_off, count, o, G.cap_stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(G.cap_stream, &graph);
    if (rc != KSPN_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc != KSPN_OK) return rc;
      KSPN_HIP(e);
    }
    G.graph = graph;
    KSPN_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    std::memcpy(G.key, key, sizeof(key));
    G.replays = 0;
  }
  if (st) KSPN_HIP(hipEventRecord(G.ev[0], s));
  KSPN_HIP(hipGraphLaunch(G.exec, s));
  ++G.replays;
  if (st) KSPN_HIP(hipEventRecord(G.ev[1], s));
  KSPN_HIP(hipMemcpyAsync(G.h_status, G.ctl, sizeof(gs::GCtl), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  const gs::GCtl* h = static_cast<const gs::GCtl*>(G.h_status);
  if (h->err) {
    note_error("graph solve", (h->err & gs::ERR_BARRIER) ? "grid barrier timed out" : "frontier queue overflow",
               __FILE__, __LINE__);
    return KSPN_E_HIP;
  }
  if (h->incomplete) return KSPN_E_ITER_CAP;
  if (!h->done) {
    note_error("graph solve", "the colour rounds ended without deciding every vertex", __FILE__, __LINE__);
    return KSPN_E_HIP;
  }
  if (st) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, G.ev[0], G.ev[1]) == hipSuccess) st->ms_total = ms;
    st->decided_trim = h->dec_trim;
    st->decided_fwbw = h->dec_fwbw;
    st->decided_color = h->dec_color;
    st->color_rounds = h->rounds;
    st->fwbw_levels_push = h->levels_push[0] + h->levels_push[1];
    st->fwbw_levels_pull = h->levels_pull[0] + h->levels_pull[1];
    st->fwbw_rounds_done = 1;
    st->partitions_max = count;
    const char* names[3] = {"trim_1_first", "forward_backward_search", "color"};
    const uint32_t dec[3] = {h->dec_trim, h->dec_fwbw, h->dec_color};
    st->n_phase = 3;
    for (int i = 0; i < 3; ++i) {
      kspn_phase& ph = st->phase[i];
      std::memset(&ph, 0, sizeof(ph));
      std::snprintf(ph.name, sizeof(ph.name), "%s", names[i]);
      ph.decided = dec[i];
      ph.entries = 1;
    }
    // one graph launch: its time (ms_total) is not split by scope
  }
  return KSPN_OK;
}

}  // namespace kspn
