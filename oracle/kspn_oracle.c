/*
 * kspn_oracle.c -- CPU oracle for the SCC path. TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it; the product library never does.
 * It restates, in plain C, the single-rank behaviour of rawbby/KaSpan's
 * SCC pipeline (citations are paths in the KaSpan tree), plus textbook
 * Tarjan, a brute-force mutual-reachability check and the seeded graph
 * generators whose edge streams the GPU generators must reproduce
 * bit-for-bit.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_tarjan          golden vectors (iSpan tiny graph), planted labels,
 *                       brute force on small graphs, closed forms
 *   orc_tarjan_batch    equal to orc_tarjan on the union, made local per graph
 *   orc_planted_labels  equal to the labels orc_gen_planted writes
 *   orc_canonicalize    hand-made cases; equality with Tarjan's min ids
 *   orc_brute_force     closed forms (cycle, DAG, empty graph, n=1)
 *   orc_transpose       golden vector test_allgather_sub_graph.cpp:33-72,
 *                       invariants (edge multiset, ascending rows)
 *   orc_subgraph        golden vector test_allgather_sub_graph.cpp:74-111
 *   orc_trim_1_first    closed-form degree test per vertex + argmax rule
 *   orc_fwbw            decided set == Tarjan SCC of the pivot
 *   orc_color_round     roots/closure == Tarjan components, invariant
 *                       "colour = min rotated id over ancestors"
 *   orc_kaspan_scc, orc_scc_trim_ex[_residual]   canonical labels == Tarjan
 *   generators          planted labels == Tarjan, fixed edge hashes
 *   MLP                 finite differences of the FP64 loss
 */
#include "kspn_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* random numbers                                                           */
/* ======================================================================== */

/* SplitMix64 (Steele, Lea, Flood 2014): output i of a generator seeded with
 * `seed` is the finaliser applied to seed + (i+1)*golden. Counter-based, so
 * a GPU thread can compute any element of the stream directly. */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

/* bijection on [0, 2^bits): three rounds of add-key, multiply by an odd
 * key, xorshift right -- each step is invertible modulo 2^bits. */
uint32_t orc_perm_pow2(uint64_t seed, uint32_t bits, uint32_t x) {
  if (bits == 0) return 0;
  uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  uint32_t sh = (bits + 1) / 2;
  uint64_t y = x & mask;
  for (uint32_t r = 0; r < 3; ++r) {
    uint64_t k = orc_splitmix64_at(seed, 1000 + r);
    y = (y + (k >> 32)) & mask;
    y = (y * ((k & 0xFFFFFFFFull) | 1ull)) & mask;
    y ^= y >> sh;
  }
  return (uint32_t)y;
}

static inline uint32_t mulhi32(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }
static inline uint32_t prob32(double p) {
  if (p <= 0.0) return 0u;
  if (p >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(p * 4294967296.0);
}

/* ======================================================================== */
/* generators                                                               */
/* ======================================================================== */

/* G(n, m): draw i gives u = hi32 * n >> 32, v = lo32 * n >> 32. */
uint64_t orc_gen_gnm(uint64_t seed, uint32_t n, uint64_t draws, uint32_t* src, uint32_t* dst) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < draws; ++i) {
    uint64_t r = orc_splitmix64_at(seed, i);
    uint32_t u = mulhi32((uint32_t)(r >> 32), n);
    uint32_t v = mulhi32((uint32_t)r, n);
    if (u != v) { src[k] = u; dst[k] = v; ++k; }
  }
  return k;
}

/* R-MAT (Chakrabarti, Zhan, Faloutsos 2004): `scale` levels, each a quadrant
 * choice from one 32-bit draw (two draws per 64-bit word; word index
 * i*16 + level/2). Bit `scale-1-level` of (u,v) is set by the quadrant. */
uint64_t orc_gen_rmat(uint64_t seed, uint32_t scale, uint64_t draws, double a, double b, double c,
                      int permute, uint32_t* src, uint32_t* dst) {
  uint32_t ta = prob32(a), tab = prob32(a + b), tabc = prob32(a + b + c);
  uint64_t pseed = mix64(seed ^ 0x5045524D55544531ull);
  uint64_t k = 0;
  for (uint64_t i = 0; i < draws; ++i) {
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) w = orc_splitmix64_at(seed, i * 16 + (l >> 1));
      uint32_t x = (l & 1) ? (uint32_t)w : (uint32_t)(w >> 32);
      uint32_t bu = 0, bv = 0;
      if (x < ta) { bu = 0; bv = 0; }
      else if (x < tab) { bu = 0; bv = 1; }
      else if (x < tabc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (permute) { u = orc_perm_pow2(pseed, scale, u); v = orc_perm_pow2(pseed, scale, v); }
    if (u != v) { src[k] = u; dst[k] = v; ++k; }
  }
  return k;
}

/* R-MAT on an arbitrary vertex count (the graphs of a batch): the levels of
 * orc_gen_rmat over [0, 2^bits) with bits = ceil(log2 n); since 2^bits < 2n
 * an id x >= n folds to x - n; ids are then scattered by perm_n below. */
static uint32_t perm_n(uint64_t pseed, uint32_t bits, uint32_t n, uint32_t x);
uint64_t orc_gen_rmat_n(uint64_t seed, uint32_t n, uint64_t draws, double a, double b, double c, uint32_t* src,
                        uint32_t* dst) {
  uint32_t ta = prob32(a), tab = prob32(a + b), tabc = prob32(a + b + c);
  uint64_t pseed = mix64(seed ^ 0x5045524D55544531ull);
  uint32_t bits = 0;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  uint64_t k = 0;
  for (uint64_t i = 0; i < draws; ++i) {
    uint32_t u = 0, v = 0;
    uint64_t w = 0;
    for (uint32_t l = 0; l < bits; ++l) {
      if ((l & 1) == 0) w = orc_splitmix64_at(seed, i * 16 + (l >> 1));
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
    if (u != v) { src[k] = u; dst[k] = v; ++k; }
  }
  return k;
}

/* Per-graph plan of a batch: draws 4i..4i+3 of the batch stream give graph
 * i's size, average out-degree, kind and seed. Log-uniform values are
 * exp(log lo + u (log hi - log lo)) with u = top 53 bits / 2^53, truncated
 * and clamped to [lo, hi]; draws = floor(n * degree). */
void orc_batch_plan(uint64_t seed, uint32_t count, uint32_t n_min, uint32_t n_max, double deg_min, double deg_max,
                    double p_rmat, uint32_t* kind, uint32_t* n, uint64_t* draws, uint64_t* gseed) {
  const double inv53 = 1.0 / 9007199254740992.0;
  const double ln0 = log((double)n_min), ln1 = log((double)n_max);
  const double ld0 = log(deg_min), ld1 = log(deg_max);
  for (uint32_t i = 0; i < count; ++i) {
    double u0 = (double)(orc_splitmix64_at(seed, 4ull * i) >> 11) * inv53;
    double u1 = (double)(orc_splitmix64_at(seed, 4ull * i + 1) >> 11) * inv53;
    double u2 = (double)(orc_splitmix64_at(seed, 4ull * i + 2) >> 11) * inv53;
    double t0 = u0 * (ln1 - ln0);
    double x = exp(ln0 + t0);
    uint32_t ni = x >= (double)n_max ? n_max : (uint32_t)x;
    if (ni < n_min) ni = n_min;
    double t1 = u1 * (ld1 - ld0);
    double d = exp(ld0 + t1);
    if (d > deg_max) d = deg_max;
    if (d < deg_min) d = deg_min;
    n[i] = ni;
    draws[i] = (uint64_t)((double)ni * d);
    kind[i] = u2 < p_rmat ? 1u : 0u;
    gseed[i] = orc_splitmix64_at(seed, 4ull * i + 3);
  }
}

uint64_t orc_lattice_max_edges(uint32_t side) {
  uint64_t s = side;
  return s < 2 ? 0 : 2 * 3 * s * s * (s - 1);
}

/* 3-D lattice side^3, id = x + side*(y + side*z). Lattice edge e enumerates
 * axis-major; draw e: hi32 < p_bidir keeps both directions, else bit 0 of
 * lo32 picks low->high (1) or high->low (0). */
uint64_t orc_gen_lattice(uint64_t seed, uint32_t side, double p_bidir, uint32_t* src, uint32_t* dst) {
  if (side < 2) return 0;
  uint64_t s = side, per_axis = s * s * (s - 1);
  uint32_t tb = prob32(p_bidir);
  uint64_t k = 0;
  for (uint64_t e = 0; e < 3 * per_axis; ++e) {
    uint64_t axis = e / per_axis, rem = e % per_axis;
    uint64_t x, y, z, stride;
    if (axis == 0) { x = rem % (s - 1); y = (rem / (s - 1)) % s; z = rem / ((s - 1) * s); stride = 1; }
    else if (axis == 1) { x = rem % s; y = (rem / s) % (s - 1); z = rem / (s * (s - 1)); stride = s; }
    else { x = rem % s; y = (rem / s) % s; z = rem / (s * s); stride = s * s; }
    uint32_t lo = (uint32_t)(x + s * (y + s * z));
    uint32_t hi = (uint32_t)(lo + stride);
    uint64_t r = orc_splitmix64_at(seed, e);
    if ((uint32_t)(r >> 32) < tb) {
      src[k] = lo; dst[k] = hi; ++k;
      src[k] = hi; dst[k] = lo; ++k;
    } else if ((uint32_t)r & 1u) {
      src[k] = lo; dst[k] = hi; ++k;
    } else {
      src[k] = hi; dst[k] = lo; ++k;
    }
  }
  return k;
}

/* Planted-SCC fuzzy graph with the structure of fuzzy_global_scc_id_and_graph
 * (test/fuzzy.hpp:81-218): vertex v opens a new component with probability
 * 1/4 (v=0 always), else joins a uniformly chosen existing one; each
 * component of size > 1 gets a random directed cycle (fuzzy.hpp:139-147);
 * extra edges go from a member of component C to a member of C or of an
 * earlier-opened component until m/n >= d (fuzzy.hpp:154-179), so planted
 * label = opener id = minimum member. Sequential SplitMix64 stream, not
 * mt19937, so the exact graphs differ from the reference's while keeping
 * its construction. Returns the number of distinct edges (<= cap). */
typedef struct { uint64_t seed, i; } seqrng;
static inline uint64_t sr_next(seqrng* r) { return orc_splitmix64_at(r->seed, r->i++); }
static inline uint32_t sr_below(seqrng* r, uint32_t n) { return mulhi32((uint32_t)(sr_next(r) >> 32), n); }

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

typedef struct {
  uint64_t* keys;
  uint64_t cap, size;
  uint64_t* table;
  uint64_t tmask;
} edgeset;

static int es_insert(edgeset* s, uint32_t u, uint32_t v) {
  if (u == v) return 0;
  uint64_t key = ((uint64_t)u << 32) | v;
  uint64_t h = mix64(key) & s->tmask;
  while (s->table[h] != ~0ull) {
    if (s->table[h] == key) return 0;
    h = (h + 1) & s->tmask;
  }
  if (s->size >= s->cap) return 0;
  s->table[h] = key;
  s->keys[s->size++] = key;
  return 1;
}

uint64_t orc_gen_fuzzy(uint64_t seed, uint32_t n, double d, uint32_t* src, uint32_t* dst, uint64_t cap,
                       uint32_t* labels) {
  seqrng rng = {seed, 0};
  if (n == 0) return 0;
  if (n < 2) d = -1.0;
  else {
    double min_d = log(fmax(1.0, log((double)n)));
    double max_d = (double)(n - 1) / 2.0;
    if (d == -1.0) d = min_d + (max_d - min_d) * ((double)(sr_next(&rng) >> 11) / 9007199254740992.0);
    else if (d < min_d) d = min_d;
    else if (d > max_d) d = max_d;
  }
  uint32_t* reps = (uint32_t*)malloc(sizeof(uint32_t) * n);
  uint32_t* comp_of = (uint32_t*)malloc(sizeof(uint32_t) * n);   /* component index per vertex */
  uint32_t* comp_size = (uint32_t*)calloc(n, sizeof(uint32_t));
  uint32_t nreps = 0;
  for (uint32_t v = 0; v < n; ++v) {
    if (v == 0 || sr_below(&rng, 4) == 0) {
      reps[nreps] = v; comp_of[v] = nreps; ++nreps;
      labels[v] = v;
    } else {
      uint32_t c = sr_below(&rng, nreps);
      comp_of[v] = c; labels[v] = reps[c];
    }
    comp_size[comp_of[v]]++;
  }
  /* members grouped by component, ascending ids inside (map order in the reference) */
  uint32_t* start = (uint32_t*)calloc((size_t)nreps + 1, sizeof(uint32_t));
  for (uint32_t c = 0; c < nreps; ++c) start[c + 1] = start[c] + comp_size[c];
  uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * (nreps + 1));
  memcpy(fill, start, sizeof(uint32_t) * (nreps + 1));
  uint32_t* members = (uint32_t*)malloc(sizeof(uint32_t) * n);
  for (uint32_t v = 0; v < n; ++v) members[fill[comp_of[v]]++] = v;

  edgeset es;
  es.cap = cap; es.size = 0; es.keys = (uint64_t*)malloc(sizeof(uint64_t) * (cap ? cap : 1));
  uint64_t tsz = 1;
  while (tsz < 2 * cap + 2) tsz <<= 1;
  es.table = (uint64_t*)malloc(sizeof(uint64_t) * tsz);
  memset(es.table, 0xFF, sizeof(uint64_t) * tsz);
  es.tmask = tsz - 1;

  for (uint32_t c = 0; c < nreps; ++c) {
    uint32_t sz = comp_size[c];
    uint32_t* m = members + start[c];
    if (sz > 1) {
      for (uint32_t i = sz - 1; i > 0; --i) { /* Fisher-Yates */
        uint32_t j = sr_below(&rng, i + 1);
        uint32_t t = m[i]; m[i] = m[j]; m[j] = t;
      }
      es_insert(&es, m[0], m[sz - 1]);
      for (uint32_t i = 1; i < sz; ++i) es_insert(&es, m[i], m[i - 1]);
    }
  }
  while (d >= 0.0 && (double)es.size / n < d && es.size < cap) {
    uint64_t before = es.size;
    uint64_t attempts = 0;
    for (uint32_t c = 0; c < nreps; ++c) {
      uint32_t sz = comp_size[c];
      for (uint32_t i = 0; i < sz; ++i) {
        if ((double)es.size / n >= d || es.size >= cap) break;
        uint32_t cv = sr_below(&rng, c + 1); /* earlier-or-same component */
        uint32_t u = members[start[c] + sr_below(&rng, sz)];
        uint32_t v = members[start[cv] + sr_below(&rng, comp_size[cv])];
        es_insert(&es, u, v);
        ++attempts;
      }
    }
    if (es.size == before) break; /* saturated (or no component could add an edge) */
    (void)attempts;
  }
  qsort(es.keys, es.size, sizeof(uint64_t), cmp_u64);
  for (uint64_t i = 0; i < es.size; ++i) { src[i] = (uint32_t)(es.keys[i] >> 32); dst[i] = (uint32_t)es.keys[i]; }
  uint64_t out = es.size;
  free(es.keys); free(es.table); free(reps); free(comp_of); free(comp_size); free(start); free(fill);
  free(members);
  return out;
}

/* Component sizes for the power-law planted graph: size_i = floor(x) with
 * x = (1-u)^(-1/(expo-1)) (continuous Pareto inverse CDF, u from draw i of
 * the size stream), clamped to [1, max_size]; drawn until the running sum
 * reaches n, the last one truncated. Returns the component count. */
uint64_t orc_planted_sizes(uint64_t seed, uint32_t n, double expo, uint32_t max_size, uint32_t* sizes,
                           uint64_t cap) {
  uint64_t sseed = mix64(seed ^ 0x53495A4553ull);
  uint64_t total = 0, k = 0;
  while (total < n && k < cap) {
    uint64_t r = orc_splitmix64_at(sseed, k);
    double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
    double x = pow(1.0 - u, -1.0 / (expo - 1.0));
    uint32_t s = (x >= (double)max_size) ? max_size : (uint32_t)x;
    if (s < 1) s = 1;
    if (total + s > n) s = (uint32_t)(n - total);
    sizes[k++] = s;
    total += s;
  }
  return k;
}

/* Power-law planted SCC graph. Positions 0..n-1 are laid out component by
 * component in topological order; vertex id of position p is a seeded
 * bijection of [0,n) (cycle-walking over the next power of two). Per
 * position p in component c = [s, s+len): a cycle edge p -> next position
 * in c (wrapping), round(intra_deg-1) chords to uniform positions of c
 * (len > 1), and round(inter_deg) edges to uniform positions in [s+len, n).
 * Edge draws use stream index p*16 + j. Labels: minimum id per component. */
static uint32_t perm_n(uint64_t pseed, uint32_t bits, uint32_t n, uint32_t x) {
  uint32_t y = orc_perm_pow2(pseed, bits, x);
  while (y >= n) y = orc_perm_pow2(pseed, bits, y);
  return y;
}

uint64_t orc_gen_planted(uint64_t seed, uint32_t n, double expo, uint32_t max_size, double intra_deg,
                         double inter_deg, uint32_t* src, uint32_t* dst, uint64_t cap, uint32_t* labels) {
  uint32_t* sizes = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t ncomp = orc_planted_sizes(seed, n, expo, max_size, sizes, n);
  uint32_t bits = 0;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  uint64_t pseed = mix64(seed ^ 0x544F504F4C4F4759ull);
  uint64_t eseed = mix64(seed ^ 0x45444745ull);
  uint32_t chords = (uint32_t)(intra_deg - 1.0 + 0.5);
  uint32_t inter = (uint32_t)(inter_deg + 0.5);
  uint64_t k = 0;
  uint64_t s = 0;
  for (uint64_t c = 0; c < ncomp; ++c) {
    uint32_t len = sizes[c];
    uint32_t minid = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < len; ++q) {
      uint32_t id = perm_n(pseed, bits, n, (uint32_t)(s + q));
      if (id < minid) minid = id;
    }
    for (uint32_t q = 0; q < len; ++q) {
      uint64_t p = s + q;
      uint32_t u = perm_n(pseed, bits, n, (uint32_t)p);
      labels[u] = minid;
      if (len > 1) {
        uint32_t nx = perm_n(pseed, bits, n, (uint32_t)(s + (q + 1) % len));
        if (k < cap) { src[k] = u; dst[k] = nx; ++k; }
        for (uint32_t j = 0; j < chords; ++j) {
          uint64_t r = orc_splitmix64_at(eseed, p * 16 + j);
          uint32_t t = perm_n(pseed, bits, n, (uint32_t)(s + mulhi32((uint32_t)(r >> 32), len)));
          if (t != u && k < cap) { src[k] = u; dst[k] = t; ++k; }
        }
      }
      uint64_t rest = (uint64_t)n - (s + len);
      if (rest > 0) {
        for (uint32_t j = 0; j < inter; ++j) {
          uint64_t r = orc_splitmix64_at(eseed, p * 16 + 8 + j);
          uint64_t off = ((r >> 32) * rest) >> 32;
          uint32_t t = perm_n(pseed, bits, n, (uint32_t)(s + len + off));
          if (k < cap) { src[k] = u; dst[k] = t; ++k; }
        }
      }
    }
    s += len;
  }
  free(sizes);
  return k;
}

/* The labels orc_gen_planted writes, without its edges: positions laid out
 * component by component, the same bijection of positions to ids, and the
 * minimum id of each component as its label. */
void orc_planted_labels(uint64_t seed, uint32_t n, double expo, uint32_t max_size, uint32_t* labels) {
  uint32_t* sizes = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t ncomp = orc_planted_sizes(seed, n, expo, max_size, sizes, n);
  uint32_t bits = 0;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  uint64_t pseed = mix64(seed ^ 0x544F504F4C4F4759ull);
  uint64_t s = 0;
  for (uint64_t c = 0; c < ncomp; ++c) {
    uint32_t len = sizes[c], minid = 0xFFFFFFFFu;
    for (uint32_t q = 0; q < len; ++q) {
      uint32_t id = perm_n(pseed, bits, n, (uint32_t)(s + q));
      if (id < minid) minid = id;
    }
    for (uint32_t q = 0; q < len; ++q) labels[perm_n(pseed, bits, n, (uint32_t)(s + q))] = minid;
    s += len;
  }
  free(sizes);
}

/* ======================================================================== */
/* CSR                                                                      */
/* ======================================================================== */

static void radix_sort_u64(uint64_t* a, uint64_t* tmp, uint64_t count, uint32_t keybits) {
  const uint32_t RB = 11;
  uint64_t* srcp = a;
  uint64_t* dstp = tmp;
  uint32_t passes = 0;
  for (uint32_t shift = 0; shift < keybits; shift += RB) {
    uint64_t hist[1 << 11];
    memset(hist, 0, sizeof(hist));
    for (uint64_t i = 0; i < count; ++i) hist[(srcp[i] >> shift) & ((1u << RB) - 1)]++;
    uint64_t acc = 0;
    for (uint32_t d = 0; d < (1u << RB); ++d) { uint64_t h = hist[d]; hist[d] = acc; acc += h; }
    for (uint64_t i = 0; i < count; ++i) dstp[hist[(srcp[i] >> shift) & ((1u << RB) - 1)]++] = srcp[i];
    uint64_t* t = srcp; srcp = dstp; dstp = t;
    ++passes;
  }
  if (passes & 1) memcpy(a, srcp, count * sizeof(uint64_t));
}

uint64_t orc_csr_from_edges(uint32_t n, uint64_t count, const uint32_t* src, const uint32_t* dst,
                            uint64_t* off, uint32_t* adj) {
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (count ? count : 1));
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (count ? count : 1));
  uint64_t k = 0;
  for (uint64_t i = 0; i < count; ++i)
    if (src[i] != dst[i]) keys[k++] = ((uint64_t)src[i] << 32) | dst[i];
  uint32_t bits = 0;
  while (bits < 32 && (1ull << bits) < n) ++bits;
  if (bits == 0) bits = 1;
  radix_sort_u64(keys, tmp, k, 32 + bits);
  free(tmp);
  memset(off, 0, sizeof(uint64_t) * ((size_t)n + 1));
  uint64_t m = 0;
  for (uint64_t i = 0; i < k; ++i) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    adj[m++] = (uint32_t)keys[i];
    off[(keys[i] >> 32) + 1]++;
  }
  for (uint32_t u = 0; u < n; ++u) off[u + 1] += off[u];
  free(keys);
  return m;
}

/* backward_complement_graph (scc/backward_complement.hpp:22-68): counting
 * sort -- in-degree histogram, exclusive scan, scatter in source order, so
 * each reverse row lists its sources in ascending order. */
void orc_transpose(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, uint64_t* bw_off,
                   uint32_t* bw_adj) {
  memset(bw_off, 0, sizeof(uint64_t) * ((size_t)n + 1));
  for (uint64_t it = 0; it < fw_off[n]; ++it) bw_off[fw_adj[it] + 1]++;
  uint64_t acc = 0;
  for (uint32_t u = 0; u < n; ++u) { uint64_t d = bw_off[u + 1]; bw_off[u + 1] = acc; acc += d; }
  for (uint32_t u = 0; u < n; ++u)
    for (uint64_t it = fw_off[u]; it < fw_off[u + 1]; ++it) bw_adj[bw_off[fw_adj[it] + 1]++] = u;
}

/* allgather_fw_sub_graph at one rank (scc/allgather_sub_graph.hpp:26-126,
 * 253-301): sub ids in ascending global id order; an edge is kept iff both
 * ends are in the subgraph, in CSR order. */
uint64_t orc_subgraph(uint32_t n, const uint64_t* off, const uint32_t* adj, const uint8_t* mask,
                      uint32_t* super_ids, uint32_t* sub_n_out, uint64_t* sub_off, uint32_t* sub_adj) {
  uint32_t* sub_of = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t sn = 0;
  for (uint32_t u = 0; u < n; ++u) {
    sub_of[u] = mask[u] ? sn : ORC_UNDECIDED;
    if (mask[u]) super_ids[sn++] = u;
  }
  uint64_t sm = 0;
  sub_off[0] = 0;
  for (uint32_t i = 0; i < sn; ++i) {
    uint32_t u = super_ids[i];
    for (uint64_t it = off[u]; it < off[u + 1]; ++it) {
      uint32_t v = adj[it];
      if (sub_of[v] != ORC_UNDECIDED) sub_adj[sm++] = sub_of[v];
    }
    sub_off[i + 1] = sm;
  }
  *sub_n_out = sn;
  free(sub_of);
  return sm;
}

/* ======================================================================== */
/* Tarjan, canonicaliser, brute force                                       */
/* ======================================================================== */

/* Iterative Tarjan with an explicit frame stack (scc/tarjan.hpp:20-122);
 * each SCC is labelled with its minimum member (the callers' convention,
 * scc_trim_ex_residual.hpp:115-123, test_tarjan_vs_kaspan.cpp:37-45).
 * Returns the number of components. */
/* Tarjan over rows off[0..n] whose entries are ids shifted by `base` (the
 * rows of one graph of a batch union: off may point into the union's
 * offsets, adj is the union's adjacency, and an entry x names vertex
 * x - base). labels are the graph's own ids. */
static uint32_t tarjan_impl(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t base, uint32_t* labels);

uint32_t orc_tarjan(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t* labels) {
  return tarjan_impl(n, off, adj, 0, labels);
}

static uint32_t tarjan_impl(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t base, uint32_t* labels) {
  uint32_t* index = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* low = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t* on_stack = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t* st = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* fu = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t* fit = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) index[i] = ORC_UNDECIDED;
  uint32_t count = 0, sp = 0, fp = 0, comps = 0;
  for (uint32_t root = 0; root < n; ++root) {
    if (index[root] != ORC_UNDECIDED) continue;
    index[root] = low[root] = count++;
    st[sp++] = root; on_stack[root] = 1;
    fu[fp] = root; fit[fp] = off[root]; ++fp;
    while (fp) {
      uint32_t u = fu[fp - 1];
      uint64_t it = fit[fp - 1];
      if (it < off[u + 1]) {
        uint32_t v = adj[it] - base;
        fit[fp - 1] = it + 1;
        if (index[v] == ORC_UNDECIDED) {
          index[v] = low[v] = count++;
          st[sp++] = v; on_stack[v] = 1;
          fu[fp] = v; fit[fp] = off[v]; ++fp;
        } else if (on_stack[v] && index[v] < low[u]) {
          low[u] = index[v];
        }
        continue;
      }
      if (low[u] == index[u]) {
        uint32_t end = sp, beg;
        do { --sp; on_stack[st[sp]] = 0; } while (st[sp] != u);
        beg = sp;
        uint32_t mn = ORC_UNDECIDED;
        for (uint32_t i = beg; i < end; ++i) if (st[i] < mn) mn = st[i];
        for (uint32_t i = beg; i < end; ++i) labels[st[i]] = mn;
        ++comps;
      }
      --fp;
      if (fp) { uint32_t p = fu[fp - 1]; if (low[u] < low[p]) low[p] = low[u]; }
    }
  }
  free(index); free(low); free(on_stack); free(st); free(fu); free(fit);
  return comps;
}

/* normalise_scc_id (test/normalise_scc.hpp:10-25): each id becomes the first
 * (smallest) vertex that carries it. Ids here are vertex ids, so an array
 * replaces the reference's unordered_map; any other id is an error. */
void orc_canonicalize(uint32_t n, uint32_t* labels) {
  uint32_t* first = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) first[i] = ORC_UNDECIDED;
  for (uint32_t u = 0; u < n; ++u) {
    uint32_t l = labels[u];
    if (l >= n) { labels[u] = ORC_UNDECIDED; continue; }
    if (first[l] == ORC_UNDECIDED) first[l] = u;
    labels[u] = first[l];
  }
  free(first);
}

/* Tarjan for every graph of a batch union (graph i owns vertices
 * [vtx_off[i], vtx_off[i+1]) and its rows point inside that range), on
 * `threads` POSIX threads that take graphs from a shared counter. labels[v]
 * is the minimum local id of v's SCC inside its own graph. */
struct batch_job {
  uint32_t count;
  const uint32_t* vtx_off;
  const uint64_t* off;
  const uint32_t* adj;
  uint32_t* labels;
  uint32_t next;
};

static void* batch_worker(void* arg) {
  struct batch_job* j = (struct batch_job*)arg;
  for (;;) {
    uint32_t i = __atomic_fetch_add(&j->next, 1u, __ATOMIC_RELAXED);
    if (i >= j->count) break;
    uint32_t b = j->vtx_off[i], e = j->vtx_off[i + 1];
    tarjan_impl(e - b, j->off + b, j->adj, b, j->labels + b);
  }
  return NULL;
}

void orc_tarjan_batch(uint32_t count, const uint32_t* vtx_off, const uint64_t* off, const uint32_t* adj,
                      uint32_t* labels, uint32_t threads) {
  struct batch_job job = {count, vtx_off, off, adj, labels, 0};
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  uint32_t started = 0;
  for (uint32_t t = 1; t < threads; ++t)
    if (pthread_create(&th[started], NULL, batch_worker, &job) == 0) ++started;
  batch_worker(&job);
  for (uint32_t t = 0; t < started; ++t) pthread_join(th[t], NULL);
}

/* u ~ v iff v in reach(u) and u in reach(v); reach sets by BFS from every
 * vertex as bit rows. O(n*(n+m)); for n <= 4096 only. */
int orc_brute_force(uint32_t n, const uint64_t* off, const uint32_t* adj, uint32_t* labels) {
  if (n > 4096) return -1;
  uint32_t words = (n + 63) / 64;
  uint64_t* reach = (uint64_t*)calloc((size_t)n * words + 1, sizeof(uint64_t));
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t s = 0; s < n; ++s) {
    uint64_t* row = reach + (size_t)s * words;
    uint32_t h = 0, t = 0;
    row[s / 64] |= 1ull << (s % 64);
    q[t++] = s;
    while (h < t) {
      uint32_t u = q[h++];
      for (uint64_t it = off[u]; it < off[u + 1]; ++it) {
        uint32_t v = adj[it];
        if (!(row[v / 64] >> (v % 64) & 1)) { row[v / 64] |= 1ull << (v % 64); q[t++] = v; }
      }
    }
  }
  for (uint32_t u = 0; u < n; ++u) {
    labels[u] = u;
    for (uint32_t v = 0; v < u; ++v) {
      int uv = (reach[(size_t)u * words + v / 64] >> (v % 64)) & 1;
      int vu = (reach[(size_t)v * words + u / 64] >> (u % 64)) & 1;
      if (uv && vu) { labels[u] = v; break; }
    }
  }
  free(reach); free(q);
  return 0;
}

/* ======================================================================== */
/* KaSpan's single-rank pipeline                                            */
/* ======================================================================== */

static inline uint32_t rotr32(uint32_t x, uint32_t r) { r &= 31; return r ? (x >> r) | (x << (32 - r)) : x; }
static inline uint32_t rotl32(uint32_t x, uint32_t r) { r &= 31; return r ? (x << r) | (x >> (32 - r)) : x; }

/* trim_1_first (scc/trim_1_local.hpp:12-46): original in/out degree zero ->
 * decided with its own id; pivot = smallest id with the maximum in*out
 * (strict '<' in an ascending scan), n if every vertex was decided. */
uint32_t orc_trim_1_first(uint32_t n, const uint64_t* fw_off, const uint64_t* bw_off, uint8_t* live,
                          uint32_t* scc_id) {
  uint64_t best = 0;
  uint32_t pivot = n;
  for (uint32_t k = 0; k < n; ++k) {
    uint64_t out = fw_off[k + 1] - fw_off[k], in = bw_off[k + 1] - bw_off[k];
    if (in == 0 || out == 0) { live[k] = 0; scc_id[k] = k; }
    else {
      uint64_t prod = in * out;
      if (best < prod) { best = prod; pivot = k; }
    }
  }
  return pivot;
}

/* forward_backward_search (scc/pivot_search.hpp:15-85): LIFO worklist from
 * the pivot over out-edges restricted to undecided vertices, then over
 * in-edges restricted to reached ones; each backward hit is decided with
 * id = pivot. */
uint32_t orc_fwbw(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                  const uint32_t* bw_adj, uint8_t* live, uint32_t pivot, uint32_t* scc_id) {
  uint8_t* reached = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t sp = 0, decided = 0;
  reached[pivot] = 1;
  stack[sp++] = pivot;
  while (sp) {
    uint32_t u = stack[--sp];
    for (uint64_t it = fw_off[u]; it < fw_off[u + 1]; ++it) {
      uint32_t v = fw_adj[it];
      if (live[v] && !reached[v]) { reached[v] = 1; stack[sp++] = v; }
    }
  }
  reached[pivot] = 0;
  live[pivot] = 0; scc_id[pivot] = pivot; ++decided;
  stack[sp++] = pivot;
  while (sp) {
    uint32_t u = stack[--sp];
    for (uint64_t it = bw_off[u]; it < bw_off[u + 1]; ++it) {
      uint32_t v = bw_adj[it];
      if (reached[v]) {
        reached[v] = 0; live[v] = 0; scc_id[v] = pivot; ++decided;
        stack[sp++] = v;
      }
    }
  }
  free(reached); free(stack);
  return decided;
}

/* trim_1_normal (scc/trim_1_local.hpp:48-73): one ascending pass; a vertex
 * is decided (own id) when no fw or bw neighbour is undecided, seeing the
 * decisions already made earlier in the same pass. */
static uint32_t trim_1_normal(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                              const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                              uint32_t* scc_id) {
  uint32_t decided = 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (!live[k]) continue;
    int has = 0;
    for (uint64_t it = fw_off[k]; it < fw_off[k + 1] && !has; ++it) has = live[fw_adj[it]];
    for (uint64_t it = bw_off[k]; it < bw_off[k + 1] && !has; ++it) has = live[bw_adj[it]];
    if (!has) { live[k] = 0; scc_id[k] = k; ++decided; }
  }
  return decided;
}

/* cache_label_search_pull with map = rotr(., rotation)
 * (scc/cache_multi_pivot_search_pull.hpp:40-244,246-266), single rank: the
 * ghost cache and frontier exchange are inert. Forward: LIFO worklist, a
 * popped undecided k pulls the minimum label over undecided in-neighbours
 * and re-activates its undecided out-neighbours on change. Roots (label ==
 * map(k)) are decided with unmap(label). Backward: an undecided k is
 * decided when an out-neighbour is decided with the same label. */
uint32_t orc_color_round(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                         const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                         uint32_t rotation, uint32_t* label, uint32_t* scc_id) {
  uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t* in_active = (uint8_t*)calloc(n ? n : 1, 1);
  uint8_t* changed = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t sp = 0, decided = 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (live[k]) { label[k] = rotr32(k, rotation); in_active[k] = 1; stack[sp++] = k; }
    else label[k] = ORC_UNDECIDED;
  }
  while (sp) {
    uint32_t k = stack[--sp];
    if (!live[k]) { in_active[k] = 0; continue; }
    uint32_t mn = label[k];
    for (uint64_t it = bw_off[k]; it < bw_off[k + 1]; ++it) {
      uint32_t v = bw_adj[it];
      if (live[v] && label[v] < mn) mn = label[v];
    }
    if (mn < label[k]) {
      label[k] = mn;
      for (uint64_t it = fw_off[k]; it < fw_off[k + 1]; ++it) {
        uint32_t v = fw_adj[it];
        if (live[v] && !in_active[v]) { in_active[v] = 1; stack[sp++] = v; }
      }
    }
    in_active[k] = 0;
  }
  memset(in_active, 0, n);
  for (uint32_t k = 0; k < n; ++k) {
    if (live[k] && label[k] == rotr32(k, rotation)) {
      live[k] = 0; scc_id[k] = rotl32(label[k], rotation); changed[k] = 1; ++decided;
    }
  }
  for (uint32_t k = 0; k < n; ++k) {
    if (!changed[k]) continue;
    for (uint64_t it = bw_off[k]; it < bw_off[k + 1]; ++it) {
      uint32_t v = bw_adj[it];
      if (live[v] && !in_active[v]) { in_active[v] = 1; stack[sp++] = v; }
    }
  }
  while (sp) {
    uint32_t k = stack[--sp];
    if (!live[k]) { in_active[k] = 0; continue; }
    int reached = 0;
    uint32_t lk = label[k];
    for (uint64_t it = fw_off[k]; it < fw_off[k + 1]; ++it) {
      uint32_t v = fw_adj[it];
      if (!live[v] && label[v] == lk) reached = 1;
    }
    if (reached) {
      live[k] = 0; scc_id[k] = rotl32(lk, rotation); ++decided;
      for (uint64_t it = bw_off[k]; it < bw_off[k + 1]; ++it) {
        uint32_t v = bw_adj[it];
        if (live[v] && !in_active[v]) { in_active[v] = 1; stack[sp++] = v; }
      }
    }
    in_active[k] = 0;
  }
  free(stack); free(in_active); free(changed);
  return decided;
}

/* ghost-cache sizes the reference allocates at one rank: hash_set of
 * fw_m + bw_m keys (memory/hash_set.hpp:249-259) and hash_map of the same
 * count (memory/hash_map.hpp:204-214), both rounded to a power of two of
 * 64-byte lines and filled with 0xFF on clear. */
static uint64_t bit_ceil64(uint64_t x) { uint64_t p = 1; while (p < x) p <<= 1; return p; }

int orc_kaspan_scc(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                   const uint32_t* bw_adj, uint32_t* scc_id, int faithful, orc_scc_stats* st) {
  orc_scc_stats s;
  memset(&s, 0, sizeof(s));
  uint8_t* live = (uint8_t*)malloc(n ? n : 1);
  memset(live, 1, n);
  uint32_t decided = 0;
  uint64_t m = fw_off[n];
  void* set_cache = NULL;
  void* map_cache = NULL;
  uint64_t set_bytes = 0, map_bytes = 0;

  uint32_t pivot = orc_trim_1_first(n, fw_off, bw_off, live, scc_id);
  for (uint32_t k = 0; k < n; ++k) decided += !live[k];
  s.decided_trim_first = decided;
  s.pivot = pivot;
  if (decided == n) goto done;

  if (faithful) {
    set_bytes = bit_ceil64((2 * m * 6 + 63) / 64) * 64;
    set_cache = malloc(set_bytes);
    if (set_cache) memset(set_cache, 0xFF, set_bytes);
  }
  s.decided_fwbw = orc_fwbw(n, fw_off, fw_adj, bw_off, bw_adj, live, pivot, scc_id);
  decided += s.decided_fwbw;
  if (decided == n) goto done;

  s.decided_trim_normal = trim_1_normal(n, fw_off, fw_adj, bw_off, bw_adj, live, scc_id);
  decided += s.decided_trim_normal;
  if (decided == n) goto done;

  {
    uint32_t* label = (uint32_t*)malloc(sizeof(uint32_t) * n);
    if (faithful) {
      map_bytes = bit_ceil64((2 * m * 12 + 63) / 64) * 64;
      map_cache = malloc(map_bytes);
      if (map_cache) memset(map_cache, 0xFF, map_bytes);
    }
    uint32_t rotation = 0;
    for (;;) {
      ++rotation;
      uint32_t d = orc_color_round(n, fw_off, fw_adj, bw_off, bw_adj, live, rotation, label, scc_id);
      s.decided_color += d;
      decided += d;
      s.color_rounds++;
      if (decided >= n) break;
      if (d == 0) { free(label); free(live); free(set_cache); free(map_cache); return -4; }
      if (map_cache) memset(map_cache, 0xFF, map_bytes);
    }
    free(label);
  }
done:
  if (st) *st = s;
  free(live); free(set_cache); free(map_cache);
  return 0;
}

/* trim_1_exhaustive (scc/trim_1_exhaustive.hpp:14-69) on one direction:
 * every queued vertex lost one live neighbour in that direction; reaching
 * degree 0 decides it (own id) and queues its successors. */
static uint32_t trim_ex_dir(uint32_t n, const uint64_t* off, const uint32_t* adj, uint8_t* live,
                            uint32_t* degree, uint32_t* queue, uint64_t qlen, uint32_t* scc_id,
                            uint32_t* decided_list, uint32_t* ndl) {
  uint32_t decided = 0;
  /* the reference's queue is LIFO or FIFO by frontier type; the fixpoint
   * is the same, so a simple stack is used */
  uint64_t sp = qlen;
  while (sp) {
    uint32_t k = queue[--sp];
    if (!live[k]) continue;
    if (--degree[k] == 0) {
      live[k] = 0; scc_id[k] = k; ++decided;
      if (decided_list) decided_list[(*ndl)++] = k;
      for (uint64_t it = off[k]; it < off[k + 1]; ++it) queue[sp++] = adj[it];
    }
  }
  return decided;
}

/* trim_1_exhaustive over the decided list (trim_1_exhaustive.hpp:48-69):
 * in-degree pass over out-edges of decided vertices, then out-degree pass
 * over in-edges. The queue holds one entry per edge, so it needs m slots. */
static uint32_t trim_ex_after(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                              const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                              uint32_t* outdeg, uint32_t* indeg, uint32_t* queue, const uint32_t* dl,
                              uint32_t ndl, uint32_t* scc_id) {
  uint32_t decided = 0;
  uint64_t q = 0;
  for (uint32_t i = 0; i < ndl; ++i)
    for (uint64_t it = fw_off[dl[i]]; it < fw_off[dl[i] + 1]; ++it) queue[q++] = fw_adj[it];
  decided += trim_ex_dir(n, fw_off, fw_adj, live, indeg, queue, q, scc_id, NULL, NULL);
  q = 0;
  for (uint32_t i = 0; i < ndl; ++i)
    for (uint64_t it = bw_off[dl[i]]; it < bw_off[dl[i] + 1]; ++it) queue[q++] = bw_adj[it];
  decided += trim_ex_dir(n, bw_off, bw_adj, live, outdeg, queue, q, scc_id, NULL, NULL);
  return decided;
}

/* trim_1_exhaustive_first (scc/trim_1_exhaustive.hpp:74-131) */
static uint32_t trim_ex_first(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                              const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                              uint32_t* outdeg, uint32_t* indeg, uint32_t* queue, uint32_t* scc_id) {
  uint32_t decided = 0;
  uint64_t q = 0;
  for (uint32_t k = 0; k < n; ++k) {
    uint64_t d = bw_off[k + 1] - bw_off[k];
    if (d > 0) indeg[k] = (uint32_t)d;
    else {
      live[k] = 0; scc_id[k] = k; ++decided;
      for (uint64_t it = fw_off[k]; it < fw_off[k + 1]; ++it) queue[q++] = fw_adj[it];
    }
  }
  decided += trim_ex_dir(n, fw_off, fw_adj, live, indeg, queue, q, scc_id, NULL, NULL);
  q = 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (!live[k]) continue;
    uint64_t d = fw_off[k + 1] - fw_off[k];
    if (d > 0) outdeg[k] = (uint32_t)d;
    else {
      live[k] = 0; scc_id[k] = k; ++decided;
      for (uint64_t it = bw_off[k]; it < bw_off[k + 1]; ++it) queue[q++] = bw_adj[it];
    }
  }
  decided += trim_ex_dir(n, bw_off, bw_adj, live, outdeg, queue, q, scc_id, NULL, NULL);
  return decided;
}

int orc_scc_trim_ex_residual(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                             const uint64_t* bw_off, const uint32_t* bw_adj, uint32_t* scc_id,
                             orc_scc_stats* st) {
  orc_scc_stats s;
  memset(&s, 0, sizeof(s));
  uint64_t m = fw_off[n];
  uint8_t* live = (uint8_t*)malloc(n ? n : 1);
  memset(live, 1, n);
  uint32_t* outdeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* indeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* queue = (uint32_t*)malloc(sizeof(uint32_t) * (m + n + 1));
  s.decided_trim_first = trim_ex_first(n, fw_off, fw_adj, bw_off, bw_adj, live, outdeg, indeg, queue, scc_id);
  free(queue); free(outdeg); free(indeg);
  /* one rank: decided_threshold = n - n/1 = 0 (scc_trim_ex_residual.hpp:47),
   * so the pivot/colour branch never runs; the rest goes to Tarjan. */
  if (s.decided_trim_first < n) {
    uint32_t* super_ids = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint64_t* sub_off = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    uint32_t* sub_adj = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    uint32_t* sub_lab = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t sn = 0;
    uint64_t sm = orc_subgraph(n, fw_off, fw_adj, live, super_ids, &sn, sub_off, sub_adj);
    orc_tarjan(sn, sub_off, sub_adj, sub_lab);
    /* min member in sub ids maps to the min global id (ids ascend) */
    for (uint32_t i = 0; i < sn; ++i) scc_id[super_ids[i]] = super_ids[sub_lab[i]];
    s.residual_n = sn; s.residual_m = sm; s.decided_residual = sn;
    free(super_ids); free(sub_off); free(sub_adj); free(sub_lab);
  }
  if (st) *st = s;
  free(live);
  return 0;
}

/* select_pivot_from_degree (scc/pivot.hpp:53-74) at one rank: first
 * (smallest) undecided id with the maximum residual out*in. */
static uint32_t select_pivot_from_degree(uint32_t n, const uint8_t* live, const uint32_t* outdeg,
                                         const uint32_t* indeg) {
  uint64_t best = 0;
  uint32_t pivot = 0;
  int found = 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (!live[k]) continue;
    uint64_t p = (uint64_t)outdeg[k] * indeg[k];
    if (!found || p > best) { best = p; pivot = k; found = 1; }
  }
  return pivot;
}

/* label_search push form with map = rotr (scc/multi_pivot_search.hpp:13-150)
 * at one rank: forward relaxes label_k < label[v] into out-neighbours;
 * roots decided; backward closure over in-edges for equal labels. Decided
 * vertices are appended to dl for the incremental trim. */
static uint32_t rot_label_search_push(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                                      const uint64_t* bw_off, const uint32_t* bw_adj, uint8_t* live,
                                      uint32_t rotation, uint32_t* label, uint32_t* scc_id,
                                      uint32_t* dl, uint32_t* ndl) {
  uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t* in_active = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t sp = 0, decided = 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (live[k]) { label[k] = rotr32(k, rotation); in_active[k] = 1; stack[sp++] = k; }
  }
  while (sp) {
    uint32_t k = stack[--sp];
    uint32_t lk = label[k];
    for (uint64_t it = fw_off[k]; it < fw_off[k + 1]; ++it) {
      uint32_t v = fw_adj[it];
      if (lk < rotr32(v, rotation) && live[v] && lk < label[v]) {
        label[v] = lk;
        if (!in_active[v]) { in_active[v] = 1; stack[sp++] = v; }
      }
    }
    in_active[k] = 0;
  }
  for (uint32_t k = 0; k < n; ++k) {
    if (live[k] && label[k] == rotr32(k, rotation)) {
      live[k] = 0; scc_id[k] = rotl32(label[k], rotation); ++decided;
      dl[(*ndl)++] = k;
      in_active[k] = 1; stack[sp++] = k;
    }
  }
  while (sp) {
    uint32_t k = stack[--sp];
    uint32_t lk = label[k];
    for (uint64_t it = bw_off[k]; it < bw_off[k + 1]; ++it) {
      uint32_t v = bw_adj[it];
      if (lk < rotr32(v, rotation) && live[v] && label[v] == lk) {
        live[v] = 0; scc_id[v] = rotl32(lk, rotation); ++decided;
        dl[(*ndl)++] = v;
        if (!in_active[v]) { in_active[v] = 1; stack[sp++] = v; }
      }
    }
    in_active[k] = 0;
  }
  free(stack); free(in_active);
  return decided;
}

int orc_scc_trim_ex(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                    const uint32_t* bw_adj, uint32_t* scc_id, orc_scc_stats* st) {
  orc_scc_stats s;
  memset(&s, 0, sizeof(s));
  uint64_t m = fw_off[n];
  uint8_t* live = (uint8_t*)malloc(n ? n : 1);
  memset(live, 1, n);
  uint32_t* outdeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* indeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* queue = (uint32_t*)malloc(sizeof(uint32_t) * (m + n + 1));
  uint32_t* dl = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* label = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t decided = trim_ex_first(n, fw_off, fw_adj, bw_off, bw_adj, live, outdeg, indeg, queue, scc_id);
  s.decided_trim_first = decided;
  if (decided < n) {
    uint32_t pivot = select_pivot_from_degree(n, live, outdeg, indeg);
    s.pivot = pivot;
    uint32_t ndl = 0;
    /* forward_backward_search with the decided stack (scc_trim_ex.hpp:50-62) */
    uint8_t* before = (uint8_t*)malloc(n);
    memcpy(before, live, n);
    uint32_t d = orc_fwbw(n, fw_off, fw_adj, bw_off, bw_adj, live, pivot, scc_id);
    for (uint32_t k = 0; k < n; ++k) if (before[k] && !live[k]) dl[ndl++] = k;
    free(before);
    s.decided_fwbw = d;
    decided += d;
    if (decided < n) {
      uint32_t t = trim_ex_after(n, fw_off, fw_adj, bw_off, bw_adj, live, outdeg, indeg, queue, dl, ndl, scc_id);
      s.decided_trim_normal = t;
      decided += t;
      uint32_t rotation = 0;
      while (decided < n) {
        ndl = 0;
        ++rotation;
        uint32_t c = rot_label_search_push(n, fw_off, fw_adj, bw_off, bw_adj, live, rotation, label, scc_id, dl, &ndl);
        s.decided_color += c;
        s.color_rounds++;
        decided += c;
        if (decided >= n) break;
        if (c == 0) { decided = n + 1; break; }
        uint32_t t2 = trim_ex_after(n, fw_off, fw_adj, bw_off, bw_adj, live, outdeg, indeg, queue, dl, ndl, scc_id);
        s.decided_trim_normal += t2;
        decided += t2;
      }
    }
  }
  if (st) *st = s;
  free(live); free(outdeg); free(indeg); free(queue); free(dl); free(label);
  return decided == n ? 0 : -4;
}

/* Tarjan on the live vertices' forward subgraph, min-id labels
 * (allgather_fw_sub_graph + tarjan, scc_ispan_like.hpp:69-84) */
static uint32_t residual_tarjan(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint8_t* live,
                                uint32_t* scc_id, orc_scc_stats* s) {
  uint64_t m = fw_off[n];
  uint32_t* super_ids = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t* sub_off = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  uint32_t* sub_adj = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* sub_lab = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t sn = 0;
  uint64_t sm = orc_subgraph(n, fw_off, fw_adj, live, super_ids, &sn, sub_off, sub_adj);
  orc_tarjan(sn, sub_off, sub_adj, sub_lab);
  for (uint32_t i = 0; i < sn; ++i) scc_id[super_ids[i]] = super_ids[sub_lab[i]];
  s->residual_n = sn;
  s->residual_m = sm;
  s->decided_residual += sn;
  free(super_ids); free(sub_off); free(sub_adj); free(sub_lab);
  return sn;
}

/* The other variants of the driver (tool/bench_kaspan/src/main.cpp:63-71)
 * at one rank, by the method numbers of include/kspn/scc.h:
 *   2 scc_hpc_like (variant/scc_hpc_like.hpp:16-69): trim_1_first, FW-BW,
 *     push colouring (rot_label_search) until done;
 *   3 scc_ispan_like (variant/scc_ispan_like.hpp:18-87): trim_1_first,
 *     FW-BW, trim_1_normal twice, Tarjan on the rest;
 *   4 scc_hpc_trim_ex (variant/scc_hpc_trim_ex.hpp:16-76):
 *     trim_1_exhaustive_first, select_pivot_from_degree, FW-BW, push
 *     colouring until done;
 *   6 scc_trim_ex_light_residual (variant/scc_trim_ex_light_residual.hpp:
 *     16-131): trim_1_exhaustive_first; the decided threshold n - n/1 is 0,
 *     so FW-BW and colouring are skipped and the decided stack stays empty;
 *     Tarjan on the rest.
 * Stage counts: decided_trim_first (first trim), decided_fwbw,
 * decided_trim_normal (trims after FW-BW), decided_color,
 * decided_residual. Raw (non-canonical) ids. */
int orc_scc_variant(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                    const uint32_t* bw_adj, int variant, uint32_t* scc_id, orc_scc_stats* st) {
  if (variant != 2 && variant != 3 && variant != 4 && variant != 6) return -1;
  orc_scc_stats s;
  memset(&s, 0, sizeof(s));
  uint64_t m = fw_off[n];
  uint8_t* live = (uint8_t*)malloc(n ? n : 1);
  memset(live, 1, n);
  uint32_t* outdeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* indeg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint32_t* queue = (uint32_t*)malloc(sizeof(uint32_t) * (m + n + 1));
  uint32_t* label = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* dl = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t decided = 0, pivot = n;
  int rc = 0;
  if (variant == 2 || variant == 3) {
    pivot = orc_trim_1_first(n, fw_off, bw_off, live, scc_id);
    for (uint32_t k = 0; k < n; ++k) decided += !live[k];
  } else {
    decided = trim_ex_first(n, fw_off, fw_adj, bw_off, bw_adj, live, outdeg, indeg, queue, scc_id);
    if (variant == 4 && decided < n) pivot = select_pivot_from_degree(n, live, outdeg, indeg);
  }
  s.decided_trim_first = decided;
  s.pivot = pivot;
  if (decided < n && variant != 6) {
    s.decided_fwbw = orc_fwbw(n, fw_off, fw_adj, bw_off, bw_adj, live, pivot, scc_id);
    decided += s.decided_fwbw;
  }
  if (decided < n && variant == 3) {
    s.decided_trim_normal = trim_1_normal(n, fw_off, fw_adj, bw_off, bw_adj, live, scc_id);
    s.decided_trim_normal += trim_1_normal(n, fw_off, fw_adj, bw_off, bw_adj, live, scc_id);
    decided += s.decided_trim_normal;
  }
  if (decided < n && (variant == 3 || variant == 6)) {
    decided += residual_tarjan(n, fw_off, fw_adj, live, scc_id, &s);
  }
  uint32_t rotation = 0;
  while (decided < n) {
    uint32_t ndl = 0;
    ++rotation;
    uint32_t c = rot_label_search_push(n, fw_off, fw_adj, bw_off, bw_adj, live, rotation, label, scc_id, dl, &ndl);
    s.decided_color += c;
    s.color_rounds++;
    decided += c;
    if (c == 0) { rc = -4; break; }
  }
  if (st) *st = s;
  free(live); free(outdeg); free(indeg); free(queue); free(label); free(dl);
  return rc;
}

/* The fixpoint of trim-1 from the live set live[]: vertices removed by
 * repeatedly deciding any live vertex without a live in-neighbour or without
 * a live out-neighbour (self loops count as neighbours). Unique, whatever
 * the order; live[] is updated. */
uint32_t orc_trim_fixpoint(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj, const uint64_t* bw_off,
                           const uint32_t* bw_adj, uint8_t* live) {
  uint32_t decided = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    for (uint32_t k = 0; k < n; ++k) {
      if (!live[k]) continue;
      int has_in = 0, has_out = 0;
      for (uint64_t it = bw_off[k]; it < bw_off[k + 1] && !has_in; ++it) has_in = live[bw_adj[it]];
      for (uint64_t it = fw_off[k]; it < fw_off[k + 1] && !has_out; ++it) has_out = live[fw_adj[it]];
      if (!has_in || !has_out) { live[k] = 0; ++decided; changed = 1; }
    }
  }
  return decided;
}

/* ======================================================================== */
/* learned pivot scorer: features and an FP64 MLP                            */
/* ======================================================================== */

/* 6 features per vertex: log1p(in), log1p(out), log1p(in)*log1p(out)/8,
 * log1p(sum of out-degrees of out-neighbours), log1p(sum of in-degrees of
 * in-neighbours), log1p(min(in,out)). Original (not residual) degrees. */
void orc_pivot_features(uint32_t n, const uint64_t* fw_off, const uint32_t* fw_adj,
                        const uint64_t* bw_off, const uint32_t* bw_adj, float* feat) {
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t out = fw_off[v + 1] - fw_off[v], in = bw_off[v + 1] - bw_off[v];
    uint64_t s2o = 0, s2i = 0;
    for (uint64_t it = fw_off[v]; it < fw_off[v + 1]; ++it) { uint32_t w = fw_adj[it]; s2o += fw_off[w + 1] - fw_off[w]; }
    for (uint64_t it = bw_off[v]; it < bw_off[v + 1]; ++it) { uint32_t w = bw_adj[it]; s2i += bw_off[w + 1] - bw_off[w]; }
    float li = log1pf((float)in), lo = log1pf((float)out);
    float* f = feat + (size_t)v * 6;
    f[0] = li;
    f[1] = lo;
    f[2] = li * lo * 0.125f;
    f[3] = log1pf((float)s2o);
    f[4] = log1pf((float)s2i);
    f[5] = log1pf((float)(in < out ? in : out));
  }
}

/* MLP 6 -> 32 -> 32 -> 1 with ReLU hidden layers; w1 is [32][6], w2 [32][32],
 * w3 [32] (row-major, output-major). */
void orc_mlp_forward(uint32_t rows, const float* x, const double* w1, const double* b1, const double* w2,
                     const double* b2, const double* w3, double b3, double* out) {
  for (uint32_t r = 0; r < rows; ++r) {
    double h1[32], h2[32];
    for (int j = 0; j < 32; ++j) {
      double a = b1[j];
      for (int i = 0; i < 6; ++i) a += w1[j * 6 + i] * (double)x[(size_t)r * 6 + i];
      h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < 32; ++j) {
      double a = b2[j];
      for (int i = 0; i < 32; ++i) a += w2[j * 32 + i] * h1[i];
      h2[j] = a > 0 ? a : 0;
    }
    double o = b3;
    for (int i = 0; i < 32; ++i) o += w3[i] * h2[i];
    out[r] = o;
  }
}

/* mean squared error 0.5*mean((y - t)^2) and its gradients */
double orc_mlp_loss_grad(uint32_t rows, const float* x, const float* target, const double* w1,
                         const double* b1, const double* w2, const double* b2, const double* w3,
                         double b3, double* gw1, double* gb1, double* gw2, double* gb2, double* gw3,
                         double* gb3) {
  memset(gw1, 0, sizeof(double) * 192); memset(gb1, 0, sizeof(double) * 32);
  memset(gw2, 0, sizeof(double) * 1024); memset(gb2, 0, sizeof(double) * 32);
  memset(gw3, 0, sizeof(double) * 32); *gb3 = 0;
  double loss = 0;
  if (rows == 0) return 0;
  double inv = 1.0 / rows;
  for (uint32_t r = 0; r < rows; ++r) {
    const float* xr = x + (size_t)r * 6;
    double a1[32], h1[32], a2[32], h2[32];
    for (int j = 0; j < 32; ++j) {
      double a = b1[j];
      for (int i = 0; i < 6; ++i) a += w1[j * 6 + i] * (double)xr[i];
      a1[j] = a; h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < 32; ++j) {
      double a = b2[j];
      for (int i = 0; i < 32; ++i) a += w2[j * 32 + i] * h1[i];
      a2[j] = a; h2[j] = a > 0 ? a : 0;
    }
    double o = b3;
    for (int i = 0; i < 32; ++i) o += w3[i] * h2[i];
    double e = o - (double)target[r];
    loss += 0.5 * e * e * inv;
    double go = e * inv;
    *gb3 += go;
    double g2[32], g1[32];
    for (int i = 0; i < 32; ++i) { gw3[i] += go * h2[i]; g2[i] = (a2[i] > 0) ? go * w3[i] : 0; }
    for (int j = 0; j < 32; ++j) {
      gb2[j] += g2[j];
      for (int i = 0; i < 32; ++i) gw2[j * 32 + i] += g2[j] * h1[i];
    }
    for (int i = 0; i < 32; ++i) {
      double a = 0;
      for (int j = 0; j < 32; ++j) a += w2[j * 32 + i] * g2[j];
      g1[i] = (a1[i] > 0) ? a : 0;
    }
    for (int j = 0; j < 32; ++j) {
      gb1[j] += g1[j];
      for (int i = 0; i < 6; ++i) gw1[j * 6 + i] += g1[j] * (double)xr[i];
    }
  }
  return loss;
}
