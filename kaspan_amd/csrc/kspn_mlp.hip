// kspn_mlp.hip -- learned pivot scorer: features, MLP inference and
// training (BF16 MFMA, FP32 accumulation), Adam, outcome logging.
//
// No counterpart in the reference: KaSpan picks the pivot with the plain
// degree-product rule (scc/trim_1_local.hpp:32-45, scc/pivot.hpp:53-74),
// which stays the fallback. The scorer only changes which vertex seeds
// FW-BW, never the labels.
//
// MLP: x[6] -> relu(W1 x + b1)[32] -> relu(W2 h1 + b2)[32] -> w3 . h2 + b3.
// Parameter vector (float, KSPN_MLP_PARAMS = 1313): W1[32][6], b1[32],
// W2[32][32], b2[32], w3[32], b3.
#include <cmath>

#include "kspn_internal.h"

namespace kspn {

constexpr int P_W1 = 0, P_B1 = 192, P_W2 = 224, P_B2 = 1248, P_W3 = 1280, P_B3 = 1312, P_COUNT = 1313;
static_assert(P_COUNT == KSPN_MLP_PARAMS, "parameter layout");

// ---------------------------------------------------------------- features
// In/out degree, their logs, and the 2-hop sums (out-degrees of
// out-neighbours, in-degrees of in-neighbours), computed by k_features.
// The neighbours' degrees come from byte arrays (k_deg8: min(degree, 255)),
// n bytes per direction instead of 8n of offsets, so the random reads stay
// in the Infinity Cache; a saturated entry is read exactly from the offsets.
constexpr uint32_t DEG8_SAT = 255;

__global__ void k_deg8(uint32_t n, const uint64_t* __restrict__ fw_off, const uint64_t* __restrict__ bw_off,
                       uint8_t* __restrict__ out8, uint8_t* __restrict__ in8) {
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG) {
    const uint64_t o = fw_off[v + 1] - fw_off[v], i = bw_off[v + 1] - bw_off[v];
    out8[v] = (uint8_t)(o < DEG8_SAT ? o : DEG8_SAT);
    in8[v] = (uint8_t)(i < DEG8_SAT ? i : DEG8_SAT);
  }
}

__device__ inline uint64_t deg_of(const uint8_t* __restrict__ d8, const uint64_t* __restrict__ off, uint32_t w,
                                  uint32_t d) {
  return d < DEG8_SAT ? (uint64_t)d : off[w + 1] - off[w];
}

// Features of `count` vertices (list[i], or i without a list), 256 per
// workgroup step: each thread loads its vertex's two rows, the rows'
// lengths are scanned in LDS, and the concatenated edges (out-row, then
// in-row, vertex by vertex) are read FEAT_EPT per thread at a time, their
// neighbours' degrees summed into the vertex's LDS slot with an LDS atomic.
// Consecutive lanes read consecutive entries of a row; a long row is
// shared by many lanes instead of one.
constexpr int FEAT_EPT = 8;

__global__ __launch_bounds__(WG) void k_features(uint32_t count, const uint32_t* __restrict__ list,
                                                 const uint64_t* __restrict__ fw_off, const uint32_t* __restrict__ fw_adj,
                                                 const uint64_t* __restrict__ bw_off, const uint32_t* __restrict__ bw_adj,
                                                 const uint8_t* __restrict__ out8, const uint8_t* __restrict__ in8,
                                                 float* __restrict__ feat) {
  __shared__ uint64_t s_fb[WG], s_bb[WG];  // row starts
  __shared__ uint64_t s_od[WG];            // out-row length (the vertex's edges past it are its in-row)
  __shared__ uint64_t s_S[WG + 1];         // exclusive prefix of the two rows' lengths
  __shared__ unsigned long long s_so[WG], s_si[WG];
  __shared__ uint64_t s_w[WG / WAVE];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < count; base += (uint64_t)gridDim.x * WG) {
    const uint64_t i = base + threadIdx.x;
    const bool ok = i < count;
    const uint32_t v = ok ? (list ? list[i] : (uint32_t)i) : 0u;
    uint64_t fb = 0, fe = 0, bb = 0, be = 0;
    if (ok) {
      fb = fw_off[v];
      fe = fw_off[v + 1];
      bb = bw_off[v];
      be = bw_off[v + 1];
    }
    const uint64_t d = (fe - fb) + (be - bb);
    s_fb[threadIdx.x] = fb;
    s_bb[threadIdx.x] = bb;
    s_od[threadIdx.x] = fe - fb;
    s_so[threadIdx.x] = 0;
    s_si[threadIdx.x] = 0;
    const uint64_t incl = wave_incl_scan64(d);
    if (lane == WAVE - 1) s_w[wid] = incl;
    __syncthreads();
    uint64_t woff = 0, T = 0;
    for (uint32_t w = 0; w < WG / WAVE; ++w) {
      woff += w < wid ? s_w[w] : 0;
      T += s_w[w];
    }
    s_S[threadIdx.x] = woff + incl - d;
    if (threadIdx.x == 0) s_S[WG] = T;
    __syncthreads();
    uint32_t lo = 0;  // each thread's edges come in increasing order: the search only moves forward
    for (uint64_t e0 = 0; e0 < T; e0 += (uint64_t)WG * FEAT_EPT) {
      uint32_t w[FEAT_EPT], src[FEAT_EPT];
      bool in_row[FEAT_EPT], has[FEAT_EPT];
#pragma unroll
      for (int r = 0; r < FEAT_EPT; ++r) {
        const uint64_t e = e0 + threadIdx.x + (uint64_t)r * WG;
        has[r] = e < T;
        w[r] = 0;
        src[r] = 0;
        in_row[r] = false;
        if (has[r]) {
          uint32_t hi = WG;
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_S[mid] <= e) lo = mid; else hi = mid;
          }
          const uint64_t k = e - s_S[lo];
          src[r] = lo;
          in_row[r] = k >= s_od[lo];
          w[r] = in_row[r] ? bw_adj[s_bb[lo] + (k - s_od[lo])] : fw_adj[s_fb[lo] + k];
        }
      }
      uint32_t d8[FEAT_EPT];
#pragma unroll
      for (int r = 0; r < FEAT_EPT; ++r) d8[r] = has[r] ? (in_row[r] ? in8[w[r]] : out8[w[r]]) : 0u;
#pragma unroll
      for (int r = 0; r < FEAT_EPT; ++r) {
        if (!has[r]) continue;
        const uint64_t dw = in_row[r] ? deg_of(in8, bw_off, w[r], d8[r]) : deg_of(out8, fw_off, w[r], d8[r]);
        if (dw) atomicAdd(in_row[r] ? &s_si[src[r]] : &s_so[src[r]], (unsigned long long)dw);
      }
    }
    __syncthreads();
    if (ok) {
      const uint64_t out = fe - fb, in = be - bb;
      const uint64_t s2o = s_so[threadIdx.x], s2i = s_si[threadIdx.x];
      const float li = log1pf((float)in), lo_ = log1pf((float)out);
      float* f = feat + i * 6;
      f[0] = li;
      f[1] = lo_;
      f[2] = li * lo_ * 0.125f;
      f[3] = log1pf((float)s2o);
      f[4] = log1pf((float)s2i);
      f[5] = log1pf((float)(in < out ? in : out));
    }
    __syncthreads();  // the next step rewrites the LDS tables
  }
}

// ---------------------------------------------------------------- inference
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ inline short to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// Each wavefront scores 16 rows per step with v_mfma_f32_16x16x16_bf16:
// A (16x16, rows x k) holds in lane l row l%16, k = 4*(l/16)..+3; B (16x16,
// k x n) holds in lane l column l%16, k = 4*(l/16)..+3; D holds in lane l
// column l%16, rows 4*(l/16)..+3. Layer outputs go through LDS to be read
// back in the A layout. Weights are rounded to BF16 once per workgroup.
// A workgroup scores 64 rows per step (4 wavefronts x 16), so every
// wavefront runs the same number of steps and barriers are uniform.
__global__ __launch_bounds__(WG) void k_mlp_forward(uint32_t rows, const float* __restrict__ x,
                                                    const float* __restrict__ params, float* __restrict__ out) {
  __shared__ short s_w1[32 * 16];   // W1 as [n][k], k padded 6 -> 16
  __shared__ short s_w2[32 * 32];   // W2 as [n][k]
  __shared__ float s_b1[32], s_b2[32], s_w3[32];
  __shared__ float s_h[WG / WAVE][16 * 33];  // per wavefront: 16 rows x 32 (+1 pad)
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 32 * 16; i += WG) {
    uint32_t n = i / 16, k = i % 16;
    s_w1[i] = k < 6 ? to_bf16(params[P_W1 + n * 6 + k]) : (short)0;
  }
  for (uint32_t i = t; i < 32 * 32; i += WG) s_w2[i] = to_bf16(params[P_W2 + i]);
  if (t < 32) {
    s_b1[t] = params[P_B1 + t];
    s_b2[t] = params[P_B2 + t];
    s_w3[t] = params[P_W3 + t];
  }
  __syncthreads();
  const float b3 = params[P_B3];
  const uint32_t lane = lane_id(), wv = t / WAVE;
  const uint32_t col = lane % 16, kq = lane / 16;  // column / k-quarter (row-quarter for D)
  float* h = s_h[wv];
  const uint64_t chunks = (rows + 63) / 64;
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint64_t r0 = ch * 64 + wv * 16;
    // layer 1: A = X rows (k < 6), two n-tiles of 16
    v4s a;
    for (int j = 0; j < 4; ++j) {
      uint32_t k = 4 * kq + j;
      uint64_t r = r0 + col;
      a[j] = (k < 6 && r < rows) ? to_bf16(x[r * 6 + k]) : (short)0;
    }
    for (int nt = 0; nt < 2; ++nt) {
      v4s b;
      for (int j = 0; j < 4; ++j) b[j] = s_w1[(nt * 16 + col) * 16 + 4 * kq + j];
      v4f d = {0.f, 0.f, 0.f, 0.f};
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
      for (int j = 0; j < 4; ++j) {
        uint32_t row = 4 * kq + j, n = nt * 16 + col;
        h[row * 33 + n] = fmaxf(d[j] + s_b1[n], 0.f);
      }
    }
    __syncthreads();
    // layer 2: A = H1 (16 x 32), two k-steps, two n-tiles
    v4s a2[2];
    for (int ks = 0; ks < 2; ++ks)
      for (int j = 0; j < 4; ++j) a2[ks][j] = to_bf16(h[col * 33 + ks * 16 + 4 * kq + j]);
    __syncthreads();
    for (int nt = 0; nt < 2; ++nt) {
      v4f d = {0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < 2; ++ks) {
        v4s b;
        for (int j = 0; j < 4; ++j) b[j] = s_w2[(nt * 16 + col) * 32 + ks * 16 + 4 * kq + j];
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[ks], b, d, 0, 0, 0);
      }
      for (int j = 0; j < 4; ++j) {
        uint32_t row = 4 * kq + j, n = nt * 16 + col;
        h[row * 33 + n] = fmaxf(d[j] + s_b2[n], 0.f);
      }
    }
    __syncthreads();
    // layer 3: 32 -> 1 dot product per row (lanes 0..15), FP32
    if (lane < 16) {
      uint64_t r = r0 + lane;
      float acc = b3;
      for (int k = 0; k < 32; ++k) acc += s_w3[k] * h[lane * 33 + k];
      if (r < rows) out[r] = acc;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- training
// Forward and backward for loss = 0.5 * mean((y - t)^2) over `rows`, every
// matrix product on v_mfma_f32_16x16x16bf16_1k: BF16 operands, FP32
// accumulation. Operands rounded to BF16: X, W1, H1 = relu(A1), W2 (as in
// k_mlp_forward), G2 and G1 (the gradients at the pre-activations A2, A1).
// A2, H2 = relu(A2), y and the bias / w3 gradients stay FP32.
//   A1 = X W1^T + b1        A2 = H1 W2^T + b2       y = H2 . w3 + b3
//   dy = (y - t) / rows     G2 = [A2 > 0] * (dy w3^T)
//   dW2 = G2^T H1           G1 = [A1 > 0] * (G2 W2)  dW1 = G1^T X
// Each wavefront takes 16 rows per step (a workgroup 64, all wavefronts in
// step): per row block the weight gradients are MFMAs with the rows as the
// reduction dimension, accumulated in registers across the steps. The
// workgroup sums its wavefronts' gradients in LDS and adds each parameter
// to grad[] once. MFMA lane layouts as in k_mlp_forward.
__global__ __launch_bounds__(WG) void k_mlp_grad(uint32_t rows, const float* __restrict__ x,
                                                 const float* __restrict__ target, const float* __restrict__ params,
                                                 float inv_rows, float* __restrict__ grad, float* loss_sum) {
  constexpr int NW = WG / WAVE;
  __shared__ short s_w1[32 * 16];  // W1 as [n][k], k padded 6 -> 16
  __shared__ short s_w2[32 * 32];  // W2 as [n][k]
  __shared__ float s_b1[32], s_b2[32], s_w3[32];
  __shared__ float s_a1[NW][16 * 33], s_a2[NW][16 * 33];  // per wavefront: 16 rows of pre-activations
  __shared__ float s_g[NW][16 * 33];                      // G2, then G1
  __shared__ float s_x[NW][16 * 8];
  __shared__ float s_dy[NW][16];
  __shared__ float s_sum[P_COUNT + 1];  // the workgroup's gradient, then its loss in the last slot
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 32 * 16; i += WG) {
    const uint32_t n = i / 16, k = i % 16;
    s_w1[i] = k < 6 ? to_bf16(params[P_W1 + n * 6 + k]) : (short)0;
  }
  for (uint32_t i = t; i < 32 * 32; i += WG) s_w2[i] = to_bf16(params[P_W2 + i]);
  if (t < 32) {
    s_b1[t] = params[P_B1 + t];
    s_b2[t] = params[P_B2 + t];
    s_w3[t] = params[P_W3 + t];
  }
  for (uint32_t i = t; i <= P_COUNT; i += WG) s_sum[i] = 0.f;
  __syncthreads();
  const float b3 = params[P_B3];
  const uint32_t lane = lane_id(), wv = t / WAVE;
  const uint32_t col = lane % 16, kq = lane / 16;
  float* a1 = s_a1[wv];
  float* a2 = s_a2[wv];
  float* g = s_g[wv];
  float* xr = s_x[wv];
  float* dy = s_dy[wv];
  v4f acc_w2[2][2], acc_w1[2];
  for (int i = 0; i < 2; ++i) {
    acc_w1[i] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < 2; ++j) acc_w2[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  float db1 = 0.f, db2 = 0.f, dw3 = 0.f, db3 = 0.f, loss = 0.f;  // lane o < 32 owns output o
  const uint64_t chunks = ((uint64_t)rows + 63) / 64;
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint64_t r0 = ch * 64 + wv * 16;
    // inputs of the 16 rows (FP32 copy for dW1's B operand)
    for (uint32_t i = lane; i < 16 * 8; i += WAVE) {
      const uint32_t r = i / 8, k = i % 8;
      xr[i] = (k < 6 && r0 + r < rows) ? x[(r0 + r) * 6 + k] : 0.f;
    }
    __syncthreads();
    // ---- forward
    v4s a;
    for (int j = 0; j < 4; ++j) {
      const uint32_t k = 4 * kq + j;
      a[j] = k < 6 ? to_bf16(xr[col * 8 + k]) : (short)0;
    }
    for (int nt = 0; nt < 2; ++nt) {
      v4s b;
      for (int j = 0; j < 4; ++j) b[j] = s_w1[(nt * 16 + col) * 16 + 4 * kq + j];
      v4f d = {0.f, 0.f, 0.f, 0.f};
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
      for (int j = 0; j < 4; ++j) a1[(4 * kq + j) * 33 + nt * 16 + col] = d[j] + s_b1[nt * 16 + col];
    }
    __syncthreads();
    v4s h1[2];  // H1 in the A layout (row col, k = ks*16 + 4kq + j)
    for (int ks = 0; ks < 2; ++ks)
      for (int j = 0; j < 4; ++j) h1[ks][j] = to_bf16(fmaxf(a1[col * 33 + ks * 16 + 4 * kq + j], 0.f));
    for (int nt = 0; nt < 2; ++nt) {
      v4f d = {0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < 2; ++ks) {
        v4s b;
        for (int j = 0; j < 4; ++j) b[j] = s_w2[(nt * 16 + col) * 32 + ks * 16 + 4 * kq + j];
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h1[ks], b, d, 0, 0, 0);
      }
      for (int j = 0; j < 4; ++j) a2[(4 * kq + j) * 33 + nt * 16 + col] = d[j] + s_b2[nt * 16 + col];
    }
    __syncthreads();
    if (lane < 16) {
      const uint64_t r = r0 + lane;
      float y = b3;
      for (int k = 0; k < 32; ++k) y += s_w3[k] * fmaxf(a2[lane * 33 + k], 0.f);
      const float e = r < rows ? y - target[r] : 0.f;
      dy[lane] = e * inv_rows;
      loss += 0.5f * e * e;
    }
    __syncthreads();
    // ---- backward: G2, and the FP32 gradients of w3, b3, b2
    for (uint32_t i = lane; i < 16 * 32; i += WAVE) {
      const uint32_t r = i / 32, o = i % 32;
      g[r * 33 + o] = a2[r * 33 + o] > 0.f ? dy[r] * s_w3[o] : 0.f;
    }
    __syncthreads();
    if (lane < 32) {
      for (int r = 0; r < 16; ++r) {
        db2 += g[r * 33 + lane];
        dw3 += dy[r] * fmaxf(a2[r * 33 + lane], 0.f);
      }
    }
    if (lane == 0)
      for (int r = 0; r < 16; ++r) db3 += dy[r];
    // dW2[o][k] += sum_r G2[r][o] H1[r][k]: M = o, N = k, K = the 16 rows
    for (int ot = 0; ot < 2; ++ot) {
      v4s ga;
      for (int j = 0; j < 4; ++j) ga[j] = to_bf16(g[(4 * kq + j) * 33 + ot * 16 + col]);
      for (int kt = 0; kt < 2; ++kt) {
        v4s hb;
        for (int j = 0; j < 4; ++j) hb[j] = to_bf16(fmaxf(a1[(4 * kq + j) * 33 + kt * 16 + col], 0.f));
        acc_w2[ot][kt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga, hb, acc_w2[ot][kt], 0, 0, 0);
      }
    }
    // G1 = [A1 > 0] * (G2 W2): M = rows, K = o, N = k
    v4f g1t[2];
    {
      v4s ga2[2];
      for (int ks = 0; ks < 2; ++ks)
        for (int j = 0; j < 4; ++j) ga2[ks][j] = to_bf16(g[col * 33 + ks * 16 + 4 * kq + j]);
      for (int nt = 0; nt < 2; ++nt) {
        v4f d = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 2; ++ks) {
          v4s b;
          for (int j = 0; j < 4; ++j) b[j] = s_w2[(ks * 16 + 4 * kq + j) * 32 + nt * 16 + col];
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga2[ks], b, d, 0, 0, 0);
        }
        g1t[nt] = d;
      }
    }
    __syncthreads();  // every lane has read G2
    for (int nt = 0; nt < 2; ++nt)
      for (int j = 0; j < 4; ++j) {
        const uint32_t r = 4 * kq + j, k = nt * 16 + col;
        g[r * 33 + k] = a1[r * 33 + k] > 0.f ? g1t[nt][j] : 0.f;
      }
    __syncthreads();
    if (lane < 32)
      for (int r = 0; r < 16; ++r) db1 += g[r * 33 + lane];
    // dW1[o][i] += sum_r G1[r][o] X[r][i]: M = o, N = i (6 of 16), K = the 16 rows
    {
      v4s xb;
      for (int j = 0; j < 4; ++j) xb[j] = col < 6 ? to_bf16(xr[(4 * kq + j) * 8 + col]) : (short)0;
      for (int ot = 0; ot < 2; ++ot) {
        v4s ga;
        for (int j = 0; j < 4; ++j) ga[j] = to_bf16(g[(4 * kq + j) * 33 + ot * 16 + col]);
        acc_w1[ot] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga, xb, acc_w1[ot], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // the workgroup's sums in LDS, then one global add per parameter
  for (int ot = 0; ot < 2; ++ot)
    for (int kt = 0; kt < 2; ++kt)
      for (int j = 0; j < 4; ++j)
        atomicAdd(&s_sum[P_W2 + (ot * 16 + 4 * kq + j) * 32 + kt * 16 + col], acc_w2[ot][kt][j]);
  if (col < 6)
    for (int ot = 0; ot < 2; ++ot)
      for (int j = 0; j < 4; ++j) atomicAdd(&s_sum[P_W1 + (ot * 16 + 4 * kq + j) * 6 + col], acc_w1[ot][j]);
  if (lane < 32) {
    atomicAdd(&s_sum[P_B1 + lane], db1);
    atomicAdd(&s_sum[P_B2 + lane], db2);
    atomicAdd(&s_sum[P_W3 + lane], dw3);
  }
  if (lane == 0) atomicAdd(&s_sum[P_B3], db3);
  if (lane < 16) atomicAdd(&s_sum[P_COUNT], loss);
  __syncthreads();
  for (uint32_t i = t; i < P_COUNT; i += WG)
    if (s_sum[i] != 0.f) atomicAdd(&grad[i], s_sum[i]);
  if (t == 0) atomicAdd(loss_sum, s_sum[P_COUNT]);
}

// Adam (Kingma & Ba 2015) on the parameter vector; state = [m | v]
__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ state, uint32_t count,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < count; i += gridDim.x * WG) {
    float g = grad[i];
    float m = b1 * state[i] + (1.f - b1) * g;
    float v = b2 * state[count + i] + (1.f - b2) * g * g;
    state[i] = m;
    state[count + i] = v;
    params[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  }
}

// ---------------------------------------------------------------- outcomes
// SCC size of every vertex from canonical labels: count per representative
// (wavefront-grouped atomics), then gather
__global__ __launch_bounds__(WG) void k_count_labels(uint32_t n, const uint32_t* __restrict__ labels, uint32_t* cnt) {
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  for (uint64_t base = (uint64_t)blockIdx.x * WG; base < pad; base += (uint64_t)gridDim.x * WG) {
    uint64_t v = base + threadIdx.x;
    bool valid = v < n;
    uint32_t rep = valid ? labels[v] : 0u;
    if (rep >= n) valid = false;
    uint64_t pending = __ballot(valid);
    while (pending) {
      uint32_t leader = __ffsll((unsigned long long)pending) - 1;
      uint32_t r = __shfl(rep, leader);
      uint64_t grp = __ballot(valid && rep == r);
      if (lane_id() == leader) atomicAdd(&cnt[r], (uint32_t)__popcll(grp));
      pending &= ~grp;
    }
  }
}
__global__ void k_gather_sizes(uint32_t n, const uint32_t* __restrict__ labels, const uint32_t* __restrict__ cnt,
                               uint32_t* __restrict__ sizes) {
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG) {
    uint32_t r = labels[v];
    sizes[v] = r < n ? cnt[r] : 0u;
  }
}

// ---------------------------------------------------------------- candidates
// log2 histogram of in*out (original degrees) over live vertices: bucket
// b = 64 - clz(prod) (0 for prod 0)
__global__ __launch_bounds__(WG) void k_prod_hist(uint32_t n, const uint32_t* __restrict__ live,
                                                  const uint64_t* __restrict__ fw_off,
                                                  const uint64_t* __restrict__ bw_off, uint32_t* hist) {
  __shared__ uint32_t s_h[65];
  for (uint32_t i = threadIdx.x; i < 65; i += WG) s_h[i] = 0;
  __syncthreads();
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < n; v += (uint64_t)gridDim.x * WG) {
    if (!bm_get(live, (uint32_t)v)) continue;
    uint64_t p = (fw_off[v + 1] - fw_off[v]) * (bw_off[v + 1] - bw_off[v]);
    atomicAdd(&s_h[p ? 64 - __clzll(p) : 0], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 65; i += WG)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// live vertices with bucket >= b_min, at most cap of them (*count counts
// them all). Each wavefront stages its candidates in LDS and reserves list
// slots once per 512 of them: one returning atomic per wavefront and step
// on a single address costs about 11 ns, 9 ms over 5e7 vertices.
constexpr uint32_t CAND_QCAP = 1024;
__global__ __launch_bounds__(WG) void k_collect_cand(uint32_t n, const uint32_t* __restrict__ live,
                                                     const uint64_t* __restrict__ fw_off,
                                                     const uint64_t* __restrict__ bw_off, uint32_t b_min, uint32_t cap,
                                                     uint32_t* __restrict__ list, uint32_t* count) {
  __shared__ uint32_t s_q[WG / WAVE][CAND_QCAP];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  uint32_t* q = s_q[wid];
  uint32_t qn = 0;  // wave-uniform
  auto flush = [&]() {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, qn);
    base = __shfl(base, 0);
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = lane; k < qn; k += WAVE)
      if (base + k < cap) list[base + k] = q[k];
    __builtin_amdgcn_wave_barrier();
    qn = 0;
  };
  const uint64_t pad = ((uint64_t)n + WG - 1) / WG * WG;
  for (uint64_t v = blockIdx.x * (uint64_t)WG + threadIdx.x; v < pad; v += (uint64_t)gridDim.x * WG) {
    bool take = false;
    if (v < n && bm_get(live, (uint32_t)v)) {
      const uint64_t p = (fw_off[v + 1] - fw_off[v]) * (bw_off[v + 1] - bw_off[v]);
      take = (p ? 64 - __clzll(p) : 0) >= b_min;
    }
    const uint64_t mask = __ballot(take);
    if (take) q[qn + (uint32_t)__popcll(mask & (lane ? (~0ull) >> (64 - lane) : 0ull))] = (uint32_t)v;
    qn += (uint32_t)__popcll(mask);
    if (qn > CAND_QCAP - WAVE) flush();
  }
  if (qn) flush();
}

// argmax of score over list (ties: smallest vertex id) -> *out
__global__ __launch_bounds__(1024) void k_argmax_score(uint32_t count, const uint32_t* __restrict__ list,
                                                       const float* __restrict__ score, uint32_t* out) {
  __shared__ float s_s[1024];
  __shared__ uint32_t s_v[1024];
  float best = -INFINITY;
  uint32_t bv = 0xFFFFFFFFu;
  for (uint32_t i = threadIdx.x; i < count; i += 1024) {
    float s = score[i];
    uint32_t v = list[i];
    if (s > best || (s == best && v < bv)) { best = s; bv = v; }
  }
  s_s[threadIdx.x] = best;
  s_v[threadIdx.x] = bv;
  __syncthreads();
  for (uint32_t st = 512; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      float s2 = s_s[threadIdx.x + st];
      uint32_t v2 = s_v[threadIdx.x + st];
      if (s2 > s_s[threadIdx.x] || (s2 == s_s[threadIdx.x] && v2 < s_v[threadIdx.x])) {
        s_s[threadIdx.x] = s2;
        s_v[threadIdx.x] = v2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s_v[0];
}

__global__ void k_fill_score(uint32_t n, float v, float* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) out[i] = v;
}

__global__ void k_scatter_score(const uint32_t* __restrict__ list, const uint32_t* count, const float* __restrict__ sc,
                                float* __restrict__ out) {
  const uint32_t cnt = *count;
  for (uint64_t i = blockIdx.x * (uint64_t)WG + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * WG)
    out[list[i]] = sc[i];
}

// host helpers used by the pipeline (kspn_scc.hip)
int mlp_features(Ctx* c, const kspn_graph* g, const uint32_t* list, uint32_t count, float* feat, hipStream_t s) {
  if (count == 0) return KSPN_OK;
  if (c->deg8_cap < g->n) {
    if (c->deg8) (void)hipFree(c->deg8);
    c->deg8 = nullptr;
    c->deg8_cap = 0;
    KSPN_HIP(hipMalloc((void**)&c->deg8, 2ull * g->n + 64));
    c->deg8_cap = g->n;
  }
  uint8_t* out8 = c->deg8;
  uint8_t* in8 = c->deg8 + g->n;
  k_deg8<<<(uint32_t)std::min<uint64_t>(((uint64_t)g->n + WG - 1) / WG, (uint64_t)c->num_cu * 16), WG, 0, s>>>(
      g->n, g->fw_off, g->bw_off, out8, in8);
  uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)count + WG - 1) / WG, (uint64_t)c->num_cu * 16);
  k_features<<<grid, WG, 0, s>>>(count, list, g->fw_off, g->fw_adj, g->bw_off, g->bw_adj, out8, in8, feat);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}

int mlp_forward(Ctx* c, const float* params, const float* x, uint32_t rows, float* out, hipStream_t s) {
  if (rows == 0) return KSPN_OK;
  uint64_t chunks = (rows + 63) / 64;
  uint32_t grid = (uint32_t)std::min<uint64_t>(chunks, (uint64_t)c->num_cu * 8);
  k_mlp_forward<<<grid, WG, 0, s>>>(rows, x, params, out);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}

// Learned pivot among live vertices: candidates = live vertices in the top
// log2 buckets of in*out (at least `want`, at most `cap`), scored by the
// MLP; the best score wins (smallest id on ties). Returns n if none.
// Scores of the multi-pivot rounds. The learned rule only breaks ties
// among a partition's top half-octave of in*out, so a vertex of low in*out
// rarely needs a score: the live vertices of the highest log2 buckets of
// in*out, at least live_n / div of them, get the MLP score, every other
// vertex -inf (a partition whose top candidates have no score takes the
// smallest id among them, as the plain rule's ties do). `list`, `feat` and
// `sc` hold up to n entries; *scored returns how many were scored.
int mlp_score_top(Ctx* c, const kspn_graph* g, const float* params, uint32_t live_n, uint32_t div, uint32_t* list,
                  uint32_t* count_dev, float* feat, float* sc, float* vscore, hipStream_t s, uint32_t* scored) {
  if (!c->mlp_buf) {
    if (hipMalloc((void**)&c->mlp_buf, sizeof(uint32_t) * (4096 + 128 + 4096 * 6 + 4096)) != hipSuccess) return KSPN_E_OOM;
  }
  uint32_t* hist = c->mlp_buf + 4096;
  KSPN_HIP(hipMemsetAsync(hist, 0, 65 * sizeof(uint32_t), s));
  const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)g->n + WG - 1) / WG, (uint64_t)c->num_cu * 16);
  k_prod_hist<<<grid, WG, 0, s>>>(g->n, c->live, g->fw_off, g->bw_off, hist);
  uint32_t h[65];
  KSPN_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  const uint64_t want = std::max<uint64_t>(1, live_n / std::max<uint32_t>(div, 1u));
  uint32_t b_min = 65;
  uint64_t acc = 0;
  while (b_min > 0 && acc < want) acc += h[--b_min];
  k_fill_score<<<grid, WG, 0, s>>>(g->n, -INFINITY, vscore);
  KSPN_HIP(hipMemsetAsync(count_dev, 0, sizeof(uint32_t), s));
  k_collect_cand<<<grid, WG, 0, s>>>(g->n, c->live, g->fw_off, g->bw_off, b_min, g->n, list, count_dev);
  uint32_t nc = 0;
  KSPN_HIP(hipMemcpyAsync(&nc, count_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  nc = std::min(nc, g->n);
  KSPN_TRY(mlp_features(c, g, list, nc, feat, s));
  KSPN_TRY(mlp_forward(c, params, feat, nc, sc, s));
  k_scatter_score<<<(uint32_t)std::min<uint64_t>(((uint64_t)nc + WG - 1) / WG + 1, (uint64_t)c->num_cu * 16), WG, 0, s>>>(
      list, count_dev, sc, vscore);
  KSPN_HIP(hipGetLastError());
  *scored = nc;
  return KSPN_OK;
}

int select_pivot_learned_impl(Ctx* c, const kspn_graph* g, const float* params, hipStream_t s, uint32_t* pivot) {
  const uint32_t want = 1024, cap = 4096;
  if (!c->mlp_buf) {
    // list[4096] | hist[65] | count + pivot | features[4096*6] | scores[4096]
    if (hipMalloc((void**)&c->mlp_buf, sizeof(uint32_t) * (4096 + 128 + 4096 * 6 + 4096)) != hipSuccess) return KSPN_E_OOM;
  }
  uint32_t* list = c->mlp_buf;
  uint32_t* hist = c->mlp_buf + 4096;
  uint32_t* cnt = c->mlp_buf + 4096 + 96;
  float* feat = reinterpret_cast<float*>(c->mlp_buf + 4096 + 128);
  float* score = feat + 4096 * 6;
  KSPN_HIP(hipMemsetAsync(hist, 0, 65 * sizeof(uint32_t), s));
  uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)g->n + WG - 1) / WG, (uint64_t)c->num_cu * 16);
  k_prod_hist<<<grid, WG, 0, s>>>(g->n, c->live, g->fw_off, g->bw_off, hist);
  uint32_t h[65];
  KSPN_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  uint32_t b_min = 65, acc = 0;
  while (b_min > 0 && acc < want) acc += h[--b_min];
  if (acc == 0) {
    *pivot = g->n;
    return KSPN_OK;
  }
  KSPN_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
  k_collect_cand<<<grid, WG, 0, s>>>(g->n, c->live, g->fw_off, g->bw_off, b_min, cap, list, cnt);
  uint32_t nc = 0;
  KSPN_HIP(hipMemcpyAsync(&nc, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  nc = std::min(nc, cap);
  KSPN_TRY(mlp_features(c, g, list, nc, feat, s));
  KSPN_TRY(mlp_forward(c, params, feat, nc, score, s));
  k_argmax_score<<<1, 1024, 0, s>>>(nc, list, score, cnt + 1);
  uint32_t pv = 0;
  KSPN_HIP(hipMemcpyAsync(&pv, cnt + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  KSPN_HIP(hipStreamSynchronize(s));
  *pivot = pv == 0xFFFFFFFFu ? g->n : pv;
  return KSPN_OK;
}

}  // namespace kspn

using namespace kspn;

extern "C" {

int kspn_pivot_features(kspn_ctx* k, const kspn_graph* g, const uint32_t* vertices, uint32_t count, float* feat,
                        void* stream) {
  if (!k || !g || (count && !feat) || (g->n > 0 && (!g->fw_off || !g->bw_off))) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  return mlp_features(&k->c, g, vertices, count, feat, (hipStream_t)stream);
}

int kspn_mlp_forward(kspn_ctx* k, const float* params, const float* x, uint32_t rows, float* out, void* stream) {
  if (!k || !params || (rows && (!x || !out))) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  return mlp_forward(&k->c, params, x, rows, out, (hipStream_t)stream);
}

int kspn_mlp_grad(kspn_ctx* k, const float* params, const float* x, const float* target, uint32_t rows, float* grad,
                  float* loss, void* stream) {
  if (!k || !params || !grad || !loss || (rows && (!x || !target))) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  hipStream_t s = (hipStream_t)stream;
  KSPN_HIP(hipMemsetAsync(grad, 0, KSPN_MLP_PARAMS * sizeof(float), s));
  KSPN_HIP(hipMemsetAsync(loss, 0, sizeof(float), s));
  if (rows == 0) return KSPN_OK;
  const uint32_t chunks = (rows + 63) / 64;  // 64 rows per workgroup step
  const uint32_t grid = std::min<uint32_t>(chunks, (uint32_t)k->c.num_cu * 4);
  k_mlp_grad<<<grid, WG, 0, s>>>(rows, x, target, params, 1.0f / rows, grad, loss);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}

int kspn_mlp_adam(kspn_ctx* k, float* params, const float* grad, float* state, uint32_t step, float lr, void* stream) {
  if (!k || !params || !grad || !state || step == 0) return KSPN_E_INVALID;
  if (hipSetDevice(k->c.device) != hipSuccess) return KSPN_E_HIP;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  k_adam<<<(KSPN_MLP_PARAMS + WG - 1) / WG, WG, 0, (hipStream_t)stream>>>(params, grad, state, KSPN_MLP_PARAMS, lr, b1, b2,
                                                                          1e-8f, c1, c2);
  KSPN_HIP(hipGetLastError());
  return KSPN_OK;
}

int kspn_scc_sizes(kspn_ctx* k, uint32_t n, const uint32_t* labels, uint32_t* sizes, void* stream) {
  if (!k || (n && (!labels || !sizes))) return KSPN_E_INVALID;
  if (n == 0) return KSPN_OK;
  Ctx* c = &k->c;
  if (hipSetDevice(c->device) != hipSuccess) return KSPN_E_HIP;
  KSPN_TRY(ctx_reserve(c, n, 0));
  hipStream_t s = (hipStream_t)stream;
  KSPN_HIP(hipMemsetAsync(c->col, 0, (uint64_t)n * sizeof(uint32_t), s));
  uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)n + WG - 1) / WG, (uint64_t)c->num_cu * 16);
  k_count_labels<<<grid, WG, 0, s>>>(n, labels, c->col);
  k_gather_sizes<<<grid, WG, 0, s>>>(n, labels, c->col, sizes);
  KSPN_HIP(hipGetLastError());
  KSPN_HIP(hipStreamSynchronize(s));
  return KSPN_OK;
}

}  // extern "C"
