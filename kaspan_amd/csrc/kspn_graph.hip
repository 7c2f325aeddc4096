// kspn_graph.hip -- seeded graph generators, CSR construction and the
// reverse-CSR build on the GPU.
//
// Generators are counter based (draw i of SplitMix64 is computed directly),
// so every edge is a pure function of (seed, index) and the CPU oracle
// produces bit-identical edge lists. CSR construction sorts 64-bit keys
// (src << 32 | dst) with a device radix sort, drops self-loops and
// duplicates and derives row offsets from row boundaries. The transpose
// sorts (dst << 32 | src), which yields the reference's row order:
// backward_complement_graph (kaspan/include/kaspan/scc/backward_complement.hpp:22-68)
// scatters sources in ascending order into each reverse row.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "kspn_internal.h"

namespace kspn {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// output i of SplitMix64 seeded with `seed`
__host__ __device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull); }
__host__ __device__ inline uint32_t mulhi32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

// bijection on [0, 2^bits): three rounds of add, odd multiply, xorshift
__host__ __device__ inline uint32_t perm_pow2(uint64_t seed, uint32_t bits, uint32_t x) {
  if (bits == 0) return 0;
  uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  uint32_t sh = (bits + 1) / 2;
  uint64_t y = x & mask;
  for (uint32_t r = 0; r < 3; ++r) {
    uint64_t k = splitmix_at(seed, 1000 + r);
    y = (y + (k >> 32)) & mask;
    y = (y * ((k & 0xFFFFFFFFull) | 1ull)) & mask;
    y ^= y >> sh;
  }
  return (uint32_t)y;
}

inline uint32_t prob32(double p) {
  if (p <= 0.0) return 0u;
  if (p >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(p * 4294967296.0);
}

inline uint32_t bits_for(uint64_t n) {
  uint32_t b = 0;
  while (b < 33 && (1ull << b) <= n) ++b;  // n itself representable (sentinel row)
  return b;
}

// keys: (u << 32 | v); a dropped draw writes the sentinel (n << 32), which sorts last
__global__ void k_gen_gnm(uint64_t seed, uint32_t n, uint64_t draws, uint64_t* keys) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < draws; i += (uint64_t)gridDim.x * WG) {
    uint64_t r = splitmix_at(seed, i);
    uint32_t u = mulhi32((uint32_t)(r >> 32), n);
    uint32_t v = mulhi32((uint32_t)r, n);
    keys[i] = (u != v) ? (((uint64_t)u << 32) | v) : ((uint64_t)n << 32);
  }
}

// R-MAT: level l uses the high (l even) or low (l odd) half of draw i*16 + l/2
__global__ void k_gen_rmat(uint64_t seed, uint32_t scale, uint64_t draws, uint32_t ta, uint32_t tab, uint32_t tabc,
                           int permute, uint64_t pseed, uint64_t* keys) {
  const uint32_t n = 1u << scale;
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < draws; i += (uint64_t)gridDim.x * WG) {
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) w = splitmix_at(seed, i * 16 + (l >> 1));
      uint32_t x = (l & 1) ? (uint32_t)w : (uint32_t)(w >> 32);
      uint32_t bu = x >= tab;
      uint32_t bv = (x >= ta && x < tab) || x >= tabc;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (permute) {
      u = perm_pow2(pseed, scale, u);
      v = perm_pow2(pseed, scale, v);
    }
    keys[i] = (u != v) ? (((uint64_t)u << 32) | v) : ((uint64_t)n << 32);
  }
}

// 3-D lattice, two key slots per lattice edge (second = sentinel unless bidirectional)
__global__ void k_gen_lattice(uint64_t seed, uint32_t side, uint32_t tb, uint64_t* keys) {
  const uint64_t s = side, per_axis = s * s * (s - 1), total = 3 * per_axis;
  const uint32_t n = (uint32_t)(s * s * s);
  for (uint64_t e = blockIdx.x * (uint64_t)WG + threadIdx.x; e < total; e += (uint64_t)gridDim.x * WG) {
    uint64_t axis = e / per_axis, rem = e % per_axis;
    uint64_t x, y, z, stride;
    if (axis == 0) { x = rem % (s - 1); y = (rem / (s - 1)) % s; z = rem / ((s - 1) * s); stride = 1; }
    else if (axis == 1) { x = rem % s; y = (rem / s) % (s - 1); z = rem / (s * (s - 1)); stride = s; }
    else { x = rem % s; y = (rem / s) % s; z = rem / (s * s); stride = s * s; }
    uint32_t lo = (uint32_t)(x + s * (y + s * z));
    uint32_t hi = (uint32_t)(lo + stride);
    uint64_t r = splitmix_at(seed, e);
    uint64_t k0, k1 = (uint64_t)n << 32;
    if ((uint32_t)(r >> 32) < tb) {
      k0 = ((uint64_t)lo << 32) | hi;
      k1 = ((uint64_t)hi << 32) | lo;
    } else if ((uint32_t)r & 1u) {
      k0 = ((uint64_t)lo << 32) | hi;
    } else {
      k0 = ((uint64_t)hi << 32) | lo;
    }
    keys[2 * e] = k0;
    keys[2 * e + 1] = k1;
  }
}

// cycle-walking bijection of [0, n) from the bijection of [0, 2^bits)
__device__ inline uint32_t perm_n(uint64_t pseed, uint32_t bits, uint32_t n, uint32_t x) {
  uint32_t y = perm_pow2(pseed, bits, x);
  while (y >= n) y = perm_pow2(pseed, bits, y);
  return y;
}

// R-MAT on any n >= 1 (graphs of a batch): `bits` = ceil(log2 n) levels of
// k_gen_rmat's quadrant choice, an id x >= n folds to x - n (2^bits < 2n),
// then perm_n scatters the ids
__global__ void k_gen_rmat_n(uint64_t seed, uint32_t n, uint32_t bits, uint64_t draws, uint32_t ta, uint32_t tab,
                             uint32_t tabc, uint64_t pseed, uint64_t* keys) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < draws; i += (uint64_t)gridDim.x * WG) {
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (uint32_t l = 0; l < bits; ++l) {
      if ((l & 1) == 0) w = splitmix_at(seed, i * 16 + (l >> 1));
      uint32_t x = (l & 1) ? (uint32_t)w : (uint32_t)(w >> 32);
      uint32_t bu = x >= tab;
      uint32_t bv = (x >= ta && x < tab) || x >= tabc;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u >= n) u -= n;
    if (v >= n) v -= n;
    u = perm_n(pseed, bits, n, u);
    v = perm_n(pseed, bits, n, v);
    keys[i] = (u != v) ? (((uint64_t)u << 32) | v) : ((uint64_t)n << 32);
  }
}

// Planted power-law SCC graph. Positions [start[c], start[c+1]) form
// component c, components in topological order; vertex id of position p is
// perm_n(p). Per position: a cycle edge to the next position of its
// component (size > 1), `chords` edges to uniform positions of the
// component (draw p*16+j), `inter` edges to uniform positions after the
// component (draw p*16+8+j). Slots per position: 1 + chords + inter;
// absent edges are the sentinel (n << 32).
__global__ void k_gen_planted(uint32_t n, uint64_t ncomp, const uint32_t* __restrict__ start, uint64_t pseed,
                              uint64_t eseed, uint32_t bits, uint32_t chords, uint32_t inter, uint64_t* __restrict__ keys) {
  const uint32_t slots = 1 + chords + inter;
  const uint64_t none = (uint64_t)n << 32;
  for (uint64_t p = blockIdx.x * (uint64_t)WG + threadIdx.x; p < n; p += (uint64_t)gridDim.x * WG) {
    uint64_t lo = 0, hi = ncomp;  // last c with start[c] <= p
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) >> 1;
      if (start[mid] <= p) lo = mid; else hi = mid;
    }
    const uint32_t s0 = start[lo], len = start[lo + 1] - s0, q = (uint32_t)(p - s0);
    const uint32_t u = perm_n(pseed, bits, n, (uint32_t)p);
    uint64_t* out = keys + p * slots;
    uint32_t k = 0;
    if (len > 1) {
      uint32_t nx = perm_n(pseed, bits, n, s0 + (q + 1) % len);
      out[k++] = ((uint64_t)u << 32) | nx;
      for (uint32_t j = 0; j < chords; ++j) {
        uint64_t r = splitmix_at(eseed, p * 16 + j);
        uint32_t t = perm_n(pseed, bits, n, s0 + mulhi32((uint32_t)(r >> 32), len));
        out[k++] = t != u ? (((uint64_t)u << 32) | t) : none;
      }
    } else {
      for (uint32_t j = 0; j <= chords; ++j) out[k++] = none;
    }
    const uint64_t rest = (uint64_t)n - (s0 + len);
    for (uint32_t j = 0; j < inter; ++j) {
      if (rest > 0) {
        uint64_t r = splitmix_at(eseed, p * 16 + 8 + j);
        uint64_t off = ((r >> 32) * rest) >> 32;
        uint32_t t = perm_n(pseed, bits, n, (uint32_t)(s0 + len + off));
        out[k++] = ((uint64_t)u << 32) | t;
      } else {
        out[k++] = none;
      }
    }
  }
}

// planted canonical labels: minimum vertex id per component
__global__ void k_planted_min(uint32_t n, uint64_t ncomp, const uint32_t* __restrict__ start, uint64_t pseed, uint32_t bits,
                              uint32_t* __restrict__ cmin, int apply, uint32_t* __restrict__ labels) {
  for (uint64_t p = blockIdx.x * (uint64_t)WG + threadIdx.x; p < n; p += (uint64_t)gridDim.x * WG) {
    uint64_t lo = 0, hi = ncomp;
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) >> 1;
      if (start[mid] <= p) lo = mid; else hi = mid;
    }
    uint32_t u = perm_n(pseed, bits, n, (uint32_t)p);
    if (!apply) atomicMin(&cmin[lo], u);
    else labels[u] = cmin[lo];
  }
}

__global__ void k_keys_from_pairs(const uint32_t* src, const uint32_t* dst, uint64_t count, uint32_t n, uint64_t* keys) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < count; i += (uint64_t)gridDim.x * WG) {
    uint32_t u = src[i], v = dst[i];
    keys[i] = (u != v && u < n && v < n) ? (((uint64_t)u << 32) | v) : ((uint64_t)n << 32);
  }
}

// keys (dst << 32 | src) for every CSR edge; the source is found by binary
// search over off. Offsets may carry a base (ebase, a batch's edge slot) and
// ids a vertex base (vbase); keys are local to the graph.
__global__ void k_keys_transposed(uint32_t n, uint64_t m, const uint64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                  uint32_t vbase, uint64_t ebase, uint64_t* __restrict__ keys) {
  for (uint64_t e = blockIdx.x * (uint64_t)WG + threadIdx.x; e < m; e += (uint64_t)gridDim.x * WG) {
    uint32_t lo = 0, hi = n;  // last row with off[row] <= e
    while (hi - lo > 1) {
      uint32_t mid = (lo + hi) >> 1;
      if (off[mid] - ebase <= e) lo = mid; else hi = mid;
    }
    keys[e] = ((uint64_t)(adj[e] - vbase) << 32) | lo;
  }
}

// sorted unique keys -> adj (low words) and row offsets from row
// boundaries; ids shift by vbase and offsets by ebase (0 for a lone graph)
__global__ void k_csr_fill(const uint64_t* __restrict__ keys, uint64_t m, uint32_t n, uint32_t vbase, uint64_t ebase,
                           uint64_t* __restrict__ off, uint32_t* __restrict__ adj) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i <= m; i += (uint64_t)gridDim.x * WG) {
    int64_t prev = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> 32);
    int64_t cur = (i == m) ? (int64_t)n : (int64_t)(keys[i] >> 32);
    for (int64_t r = prev + 1; r <= cur; ++r) off[r] = ebase + i;
    if (i < m) adj[i] = vbase + (uint32_t)keys[i];
  }
}

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
      p = nullptr;
      return KSPN_E_OOM;
    }
    return KSPN_OK;
  }
};

uint32_t gen_grid(uint64_t items) {
  uint64_t g = (items + WG - 1) / WG;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 65536));
}

// sort keys[count] (double buffer), keep distinct keys below the sentinel
// row, and write CSR (off[n+1], adj[m]). Returns m. `keys` and `alt` are
// both clobbered.
int keys_to_csr(uint32_t n, uint64_t count, uint64_t* keys, uint64_t* alt, uint64_t* off, uint32_t** adj_out,
                uint64_t* m_out, bool alloc_adj, uint32_t* adj_given, hipStream_t s, uint32_t vbase = 0,
                uint64_t ebase = 0, uint64_t adj_cap = ~0ull) {
  uint32_t end_bit = 32 + bits_for(n);
  if (end_bit > 64) end_bit = 64;
  hipcub::DoubleBuffer<uint64_t> db(keys, alt);
  size_t tmp_bytes = 0;
  KSPN_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, db, (int)count, 0, (int)end_bit, s));
  DevBuf tmp;
  KSPN_TRY(tmp.alloc(tmp_bytes));
  KSPN_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, db, (int)count, 0, (int)end_bit, s));
  uint64_t* sorted = db.Current();
  uint64_t* uniq = db.Alternate();
  DevBuf nsel;
  KSPN_TRY(nsel.alloc(sizeof(uint64_t)));
  size_t tb2 = 0;
  KSPN_HIP(hipcub::DeviceSelect::Unique(nullptr, tb2, sorted, uniq, (uint64_t*)nsel.p, (int)count, s));
  DevBuf tmp2;
  KSPN_TRY(tmp2.alloc(tb2));
  KSPN_HIP(hipcub::DeviceSelect::Unique(tmp2.p, tb2, sorted, uniq, (uint64_t*)nsel.p, (int)count, s));
  uint64_t nu = 0;
  KSPN_HIP(hipMemcpyAsync(&nu, nsel.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  uint64_t m = nu;
  if (nu > 0) {
    uint64_t last = 0;
    KSPN_HIP(hipMemcpy(&last, uniq + nu - 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if ((last >> 32) >= n) m = nu - 1;  // the sentinel row
  }
  uint32_t* adj = adj_given;
  if (!alloc_adj && m > adj_cap) return KSPN_E_INVALID;
  if (alloc_adj) {
    if (hipMalloc((void**)&adj, (m ? m : 1) * sizeof(uint32_t)) != hipSuccess) return KSPN_E_OOM;
  }
  k_csr_fill<<<gen_grid(m + 1), WG, 0, s>>>(uniq, m, n, vbase, ebase, off, adj);
  KSPN_HIP(hipGetLastError());
  KSPN_HIP(hipStreamSynchronize(s));
  if (adj_out) *adj_out = adj;
  *m_out = m;
  return KSPN_OK;
}

// reverse CSR of (n, m, off, adj) into bw_off / bw_adj (caller buffers);
// with bases, (off, adj) is graph slice of a batch and the result is
// written with the same bases
int transpose(uint32_t n, uint64_t m, const uint64_t* off, const uint32_t* adj, uint64_t* bw_off, uint32_t* bw_adj,
              hipStream_t s, uint32_t vbase = 0, uint64_t ebase = 0) {
  DevBuf a, b;
  KSPN_TRY(a.alloc((m ? m : 1) * sizeof(uint64_t)));
  KSPN_TRY(b.alloc((m ? m : 1) * sizeof(uint64_t)));
  if (m) k_keys_transposed<<<gen_grid(m), WG, 0, s>>>(n, m, off, adj, vbase, ebase, (uint64_t*)a.p);
  KSPN_HIP(hipGetLastError());
  uint64_t mm = 0;
  KSPN_TRY(keys_to_csr(n, m, (uint64_t*)a.p, (uint64_t*)b.p, bw_off, nullptr, &mm, false, bw_adj, s, vbase, ebase, m));
  if (mm != m) return KSPN_E_INVALID;  // duplicate or out-of-range edges in the input
  return KSPN_OK;
}

// component starts of the planted graph (host; the same draws as the
// oracle's orc_planted_sizes): size_i = floor((1-u)^(-1/(expo-1))) clamped
// to [1, max_size], u from draw i of the size stream, until the sum is n
int planted_starts(const kspn_gen_params* p, std::vector<uint32_t>* starts) {
  const uint64_t n = p->p0;
  const double expo = p->a;
  const uint32_t max_size = (uint32_t)std::max<uint64_t>(1, p->p1);
  if (!(expo > 1.0)) return KSPN_E_INVALID;
  uint64_t sseed = mix64(p->seed ^ 0x53495A4553ull);
  starts->clear();
  starts->push_back(0);
  uint64_t total = 0, k = 0;
  while (total < n) {
    uint64_t r = splitmix_at(sseed, k++);
    double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
    double x = std::pow(1.0 - u, -1.0 / (expo - 1.0));
    uint32_t sz = (x >= (double)max_size) ? max_size : (uint32_t)x;
    if (sz < 1) sz = 1;
    if (total + sz > n) sz = (uint32_t)(n - total);
    total += sz;
    starts->push_back((uint32_t)total);
  }
  return KSPN_OK;
}

int build_both(uint32_t n, uint64_t count, uint64_t* keys, uint64_t* alt, uint64_t* m, uint64_t** fw_off, uint32_t** fw_adj,
               uint64_t** bw_off, uint32_t** bw_adj, hipStream_t s) {
  uint64_t *fo = nullptr, *bo = nullptr;
  uint32_t *fa = nullptr, *ba = nullptr;
  if (hipMalloc((void**)&fo, (n + 1ull) * 8) != hipSuccess) return KSPN_E_OOM;
  int rc = keys_to_csr(n, count, keys, alt, fo, &fa, m, true, nullptr, s);
  if (rc == KSPN_OK) {
    if (hipMalloc((void**)&bo, (n + 1ull) * 8) != hipSuccess || hipMalloc((void**)&ba, (*m ? *m : 1) * 4) != hipSuccess)
      rc = KSPN_E_OOM;
  }
  if (rc == KSPN_OK) rc = transpose(n, *m, fo, fa, bo, ba, s);
  if (rc != KSPN_OK) {
    for (void* p : {(void*)fo, (void*)fa, (void*)bo, (void*)ba})
      if (p) (void)hipFree(p);
    return rc;
  }
  *fw_off = fo;
  *fw_adj = fa;
  *bw_off = bo;
  *bw_adj = ba;
  return KSPN_OK;
}

// per-graph plan of a batch: draws 4i..4i+3 of the batch stream give graph
// i's size, average out-degree, kind and seed; log-uniform values are
// exp(log lo + u (log hi - log lo)), u = top 53 bits / 2^53, truncated and
// clamped to [lo, hi]; draws = floor(n * degree). The oracle's
// orc_batch_plan evaluates the same expressions in the same order.
int batch_plan(const kspn_batch_gen* p, uint32_t* kind, uint32_t* n, uint64_t* draws, uint64_t* gseed) {
#pragma clang fp contract(off)
  if (!p || p->count == 0 || p->n_min < 1 || p->n_min > p->n_max || p->n_max >= 0xFFFFFFFFu) return KSPN_E_INVALID;
  if (!(p->deg_min > 0.0) || !(p->deg_min <= p->deg_max) || !(p->deg_max < 1e6)) return KSPN_E_INVALID;
  const double inv53 = 1.0 / 9007199254740992.0;
  const double ln0 = std::log((double)p->n_min), ln1 = std::log((double)p->n_max);
  const double ld0 = std::log(p->deg_min), ld1 = std::log(p->deg_max);
  for (uint32_t i = 0; i < p->count; ++i) {
    double u0 = (double)(splitmix_at(p->seed, 4ull * i) >> 11) * inv53;
    double u1 = (double)(splitmix_at(p->seed, 4ull * i + 1) >> 11) * inv53;
    double u2 = (double)(splitmix_at(p->seed, 4ull * i + 2) >> 11) * inv53;
    double t0 = u0 * (ln1 - ln0);
    double x = std::exp(ln0 + t0);
    uint32_t ni = x >= (double)p->n_max ? p->n_max : (uint32_t)x;
    if (ni < p->n_min) ni = p->n_min;
    double t1 = u1 * (ld1 - ld0);
    double d = std::exp(ld0 + t1);
    if (d > p->deg_max) d = p->deg_max;
    if (d < p->deg_min) d = p->deg_min;
    n[i] = ni;
    draws[i] = (uint64_t)((double)ni * d);
    kind[i] = u2 < p->p_rmat ? (uint32_t)KSPN_GEN_RMAT : (uint32_t)KSPN_GEN_GNM;
    gseed[i] = splitmix_at(p->seed, 4ull * i + 3);
    if (draws[i] >= (1ull << 31)) return KSPN_E_INVALID;  // one device sort per graph
  }
  return KSPN_OK;
}

}  // namespace
}  // namespace kspn

using namespace kspn;

extern "C" {

int kspn_transpose(kspn_ctx* ctx, uint32_t n, uint64_t m, const uint64_t* fw_off, const uint32_t* fw_adj,
                   uint64_t* bw_off, uint32_t* bw_adj, void* stream) {
  (void)ctx;
  if (!fw_off || !bw_off || (m && (!fw_adj || !bw_adj)) || n == 0xFFFFFFFFu) return KSPN_E_INVALID;
  return transpose(n, m, fw_off, fw_adj, bw_off, bw_adj, (hipStream_t)stream);
}

int kspn_build_csr(kspn_ctx* ctx, uint32_t n, uint64_t count, const uint32_t* src, const uint32_t* dst, uint64_t* m,
                   uint64_t** fw_off, uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj, void* stream) {
  (void)ctx;
  if (!m || !fw_off || !fw_adj || !bw_off || !bw_adj || (count && (!src || !dst)) || n == 0xFFFFFFFFu) return KSPN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  DevBuf a, b;
  KSPN_TRY(a.alloc((count ? count : 1) * 8));
  KSPN_TRY(b.alloc((count ? count : 1) * 8));
  if (count) k_keys_from_pairs<<<gen_grid(count), WG, 0, s>>>(src, dst, count, n, (uint64_t*)a.p);
  KSPN_HIP(hipGetLastError());
  return build_both(n, count, (uint64_t*)a.p, (uint64_t*)b.p, m, fw_off, fw_adj, bw_off, bw_adj, s);
}

int kspn_generate(kspn_ctx* ctx, const kspn_gen_params* p, uint32_t* n_out, uint64_t* m, uint64_t** fw_off,
                  uint32_t** fw_adj, uint64_t** bw_off, uint32_t** bw_adj, void* stream) {
  (void)ctx;
  if (!p || !n_out || !m || !fw_off || !fw_adj || !bw_off || !bw_adj) return KSPN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  uint64_t n = 0, count = 0;
  if (p->kind == KSPN_GEN_GNM) {
    n = p->p0;
    count = p->p1;
  } else if (p->kind == KSPN_GEN_RMAT) {
    if (p->p0 < 1 || p->p0 > 31) return KSPN_E_INVALID;
    n = 1ull << p->p0;
    count = p->p1;
  } else if (p->kind == KSPN_GEN_LATTICE) {
    if (p->p0 < 2 || p->p0 > 1625) return KSPN_E_INVALID;
    n = p->p0 * p->p0 * p->p0;
    count = 2 * 3 * p->p0 * p->p0 * (p->p0 - 1);
  } else if (p->kind == KSPN_GEN_PLANTED) {
    n = p->p0;
    uint32_t chords = (uint32_t)(p->b - 1.0 + 0.5), inter = (uint32_t)(p->c + 0.5);
    if (p->b < 1.0 || p->c < 0.0 || chords > 8 || inter > 8) return KSPN_E_INVALID;
    count = n * (1 + chords + inter);
  } else {
    return KSPN_E_INVALID;
  }
  if (n == 0 || n >= 0xFFFFFFFFull || count >= (1ull << 31)) return KSPN_E_INVALID;
  DevBuf a, b;
  KSPN_TRY(a.alloc((count ? count : 1) * 8));
  KSPN_TRY(b.alloc((count ? count : 1) * 8));
  uint64_t* keys = (uint64_t*)a.p;
  if (count) {
    if (p->kind == KSPN_GEN_GNM) {
      k_gen_gnm<<<gen_grid(count), WG, 0, s>>>(p->seed, (uint32_t)n, count, keys);
    } else if (p->kind == KSPN_GEN_RMAT) {
      uint64_t pseed = mix64(p->seed ^ 0x5045524D55544531ull);
      k_gen_rmat<<<gen_grid(count), WG, 0, s>>>(p->seed, (uint32_t)p->p0, count, prob32(p->a), prob32(p->a + p->b),
                                                prob32(p->a + p->b + p->c), (int)p->permute, pseed, keys);
    } else if (p->kind == KSPN_GEN_LATTICE) {
      k_gen_lattice<<<gen_grid(count / 2), WG, 0, s>>>(p->seed, (uint32_t)p->p0, prob32(p->a), keys);
    } else {
      std::vector<uint32_t> starts;
      KSPN_TRY(planted_starts(p, &starts));
      DevBuf ds;
      KSPN_TRY(ds.alloc(starts.size() * sizeof(uint32_t)));
      KSPN_HIP(hipMemcpy(ds.p, starts.data(), starts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      uint32_t chords = (uint32_t)(p->b - 1.0 + 0.5), inter = (uint32_t)(p->c + 0.5);
      k_gen_planted<<<gen_grid(n), WG, 0, s>>>((uint32_t)n, starts.size() - 1, (const uint32_t*)ds.p,
                                               mix64(p->seed ^ 0x544F504F4C4F4759ull), mix64(p->seed ^ 0x45444745ull),
                                               bits_for(n - 1), chords, inter, keys);
      KSPN_HIP(hipGetLastError());
      KSPN_HIP(hipStreamSynchronize(s));  // ds is freed at scope exit
    }
    KSPN_HIP(hipGetLastError());
  }
  *n_out = (uint32_t)n;
  return build_both((uint32_t)n, count, keys, (uint64_t*)b.p, m, fw_off, fw_adj, bw_off, bw_adj, s);
}

int kspn_planted_labels(kspn_ctx* ctx, const kspn_gen_params* p, uint32_t* labels, void* stream) {
  (void)ctx;
  if (!p || p->kind != KSPN_GEN_PLANTED || !labels || p->p0 == 0 || p->p0 >= 0xFFFFFFFFull) return KSPN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = (uint32_t)p->p0;
  std::vector<uint32_t> starts;
  KSPN_TRY(planted_starts(p, &starts));
  uint64_t nc = starts.size() - 1;
  DevBuf ds, dm;
  KSPN_TRY(ds.alloc(starts.size() * sizeof(uint32_t)));
  KSPN_TRY(dm.alloc(nc * sizeof(uint32_t)));
  KSPN_HIP(hipMemcpy(ds.p, starts.data(), starts.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  KSPN_HIP(hipMemset(dm.p, 0xFF, nc * sizeof(uint32_t)));
  uint64_t pseed = mix64(p->seed ^ 0x544F504F4C4F4759ull);
  k_planted_min<<<gen_grid(n), WG, 0, s>>>(n, nc, (const uint32_t*)ds.p, pseed, bits_for(n - 1), (uint32_t*)dm.p, 0, labels);
  k_planted_min<<<gen_grid(n), WG, 0, s>>>(n, nc, (const uint32_t*)ds.p, pseed, bits_for(n - 1), (uint32_t*)dm.p, 1, labels);
  KSPN_HIP(hipGetLastError());
  KSPN_HIP(hipStreamSynchronize(s));
  return KSPN_OK;
}

int kspn_batch_plan(const kspn_batch_gen* p, uint32_t* kind, uint32_t* n, uint64_t* draws, uint64_t* seed) {
  if (!kind || !n || !draws || !seed) return KSPN_E_INVALID;
  return batch_plan(p, kind, n, draws, seed);
}

int kspn_generate_batch(kspn_ctx* ctx, const kspn_batch_gen* p, kspn_batch* out, void* stream) {
  (void)ctx;
  if (!p || !out) return KSPN_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t count = p->count;
  std::vector<uint32_t> kind(count), ns(count);
  std::vector<uint64_t> draws(count), seeds(count);
  KSPN_TRY(batch_plan(p, kind.data(), ns.data(), draws.data(), seeds.data()));
  uint64_t total_n = 0, total_draws = 0, max_draws = 1;
  for (uint32_t i = 0; i < count; ++i) {
    total_n += ns[i];
    total_draws += draws[i];
    max_draws = std::max<uint64_t>(max_draws, draws[i]);
  }
  if (total_n >= 0xFFFFFFFFull) return KSPN_E_INVALID;
  const uint32_t N = (uint32_t)total_n;
  // results (owned by the caller on success) and per-graph scratch
  DevBuf fo, bo, fa, ba, vo, eo, keys, alt;
  KSPN_TRY(fo.alloc((N + 1ull) * 8));
  KSPN_TRY(bo.alloc((N + 1ull) * 8));
  KSPN_TRY(fa.alloc(std::max<uint64_t>(total_draws, 1) * 4));
  KSPN_TRY(ba.alloc(std::max<uint64_t>(total_draws, 1) * 4));
  KSPN_TRY(vo.alloc((count + 1ull) * 4));
  KSPN_TRY(eo.alloc((count + 1ull) * 8));
  KSPN_TRY(keys.alloc(max_draws * 8));
  KSPN_TRY(alt.alloc(max_draws * 8));
  uint64_t* fw_off = (uint64_t*)fo.p;
  uint64_t* bw_off = (uint64_t*)bo.p;
  uint32_t* fw_adj = (uint32_t*)fa.p;
  uint32_t* bw_adj = (uint32_t*)ba.p;
  std::vector<uint32_t> h_vo(count + 1, 0);
  std::vector<uint64_t> h_eo(count + 1, 0);
  const uint32_t ta = prob32(p->a), tab = prob32(p->a + p->b), tabc = prob32(p->a + p->b + p->c);
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t vb = h_vo[i], n = ns[i];
    const uint64_t eb = h_eo[i], cnt = draws[i];
    uint64_t* k = (uint64_t*)keys.p;
    if (cnt) {
      if (kind[i] == KSPN_GEN_RMAT) {
        uint32_t bits = 0;
        while (bits < 32 && (1ull << bits) < n) ++bits;
        k_gen_rmat_n<<<gen_grid(cnt), WG, 0, s>>>(seeds[i], n, bits, cnt, ta, tab, tabc,
                                                  mix64(seeds[i] ^ 0x5045524D55544531ull), k);
      } else {
        k_gen_gnm<<<gen_grid(cnt), WG, 0, s>>>(seeds[i], n, cnt, k);
      }
      KSPN_HIP(hipGetLastError());
    }
    uint64_t m = 0;
    KSPN_TRY(keys_to_csr(n, cnt, k, (uint64_t*)alt.p, fw_off + vb, nullptr, &m, false, fw_adj + eb, s, vb, eb,
                         total_draws - eb));
    KSPN_TRY(transpose(n, m, fw_off + vb, fw_adj + eb, bw_off + vb, bw_adj + eb, s, vb, eb));
    h_vo[i + 1] = vb + n;
    h_eo[i + 1] = eb + m;
  }
  KSPN_HIP(hipMemcpy(vo.p, h_vo.data(), (count + 1ull) * 4, hipMemcpyHostToDevice));
  KSPN_HIP(hipMemcpy(eo.p, h_eo.data(), (count + 1ull) * 8, hipMemcpyHostToDevice));
  out->count = count;
  out->n = N;
  out->m = h_eo[count];
  out->vtx_off = (const uint32_t*)vo.p;
  out->edge_off = (const uint64_t*)eo.p;
  out->fw_off = fw_off;
  out->fw_adj = fw_adj;
  out->bw_off = bw_off;
  out->bw_adj = bw_adj;
  for (DevBuf* d : {&fo, &bo, &fa, &ba, &vo, &eo}) d->p = nullptr;  // handed to the caller
  return KSPN_OK;
}

int kspn_free(void* ptr) {
  if (ptr && hipFree(ptr) != hipSuccess) return KSPN_E_HIP;
  return KSPN_OK;
}

}  // extern "C"
