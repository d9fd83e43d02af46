// Learned coupling closure: an 18 -> 64 -> 64 -> 3 MLP (GELU) with BF16
// weights and activations and FP32 accumulation on the gfx950 matrix cores.
//
// No counterpart in the reference: the penalty coupling of fsi/coupling.py
// (333-406) is the method the closure corrects. Training (a fused forward,
// loss and backward pass per batch, then Adam) and inference run on
// v_mfma_f32_16x16x16_bf16. One wavefront owns 16 batch rows at a time.
//
// Operand layouts of the 16x16x16 BF16 matrix instruction (lane l of 64):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + j   (4 bf16 per lane, j = 0..3)
//   B[k][n]: n = l % 16, k = 4 (l / 16) + j
//   D[m][n]: n = l % 16, m = 4 (l / 16) + i   (4 floats per lane)
// The activations are carried transposed (feature x batch): the D block of
// one layer, rounded to BF16, is then exactly the B operand of the next
// layer, with no data movement. The weight gradients reduce over the batch,
// which must be the k index; their operands pass through a 16x16 LDS
// transpose per block.

#include <algorithm>
#include <cmath>

#include "common.h"

namespace fsi {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = FSI_MLP_IN, kInPad = 32, kHid = FSI_MLP_HIDDEN, kOut = FSI_MLP_OUT, kOutPad = 16;
// flat FP32 parameter vector: W1[64][32] (input columns 18..31 zero), b1[64],
// W2[64][64], b2[64], W3[16][64] (rows 3..15 zero), b3[16]
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kHid * kInPad;
constexpr int oW2 = oB1 + kHid;
constexpr int oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid;
constexpr int oB3 = oW3 + kOutPad * kHid;
constexpr int kParams = oB3 + kOutPad;
static_assert(kParams == FSI_MLP_PARAMS, "parameter layout");
static_assert(kIn <= kInPad && kOut <= 4, "layer shapes");

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ short bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_value(short b) { return __uint_as_float((uint32_t)(uint16_t)b << 16); }

// GELU(z) = z Phi(z) and its derivative Phi(z) + z phi(z), with Phi the
// standard normal CDF, Phi(z) = (1 + erf(z / sqrt 2)) / 2, and phi its
// density. erf by Abramowitz and Stegun, Handbook of Mathematical Functions,
// 7.1.26 (|error| <= 1.5e-7, far below the BF16 rounding of the
// activations): erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5))))
// exp(-x^2), t = 1 / (1 + p x), x >= 0, odd in x. exp(-x^2) at x = z/sqrt 2
// is the exp(-z^2/2) of phi, so one exponential serves both.
__device__ __forceinline__ void normal_cdf_pdf(float z, float& cdf, float& pdf) {
  const float x = fabsf(z) * 0.70710678118654752f;
  const float t = 1.0f / (1.0f + 0.3275911f * x);
  const float e = __expf(-0.5f * z * z);
  const float poly =
      t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  const float erf_abs = 1.0f - poly * e;
  cdf = 0.5f * (1.0f + copysignf(erf_abs, z));
  pdf = 0.39894228040143268f * e;
}
__device__ __forceinline__ float gelu(float z) {
  float cdf, pdf;
  normal_cdf_pdf(z, cdf, pdf);
  return z * cdf;
}
__device__ __forceinline__ float gelu_grad(float z) {
  float cdf, pdf;
  normal_cdf_pdf(z, cdf, pdf);
  return cdf + z * pdf;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 to_bf16x4(f32x4 v) {
  bf16x4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = bf16_bits(v[i]);
  return r;
}

// 4 consecutive BF16 words of an LDS row (8-byte aligned)
__device__ __forceinline__ bf16x4 lds4(const short* p) { return *reinterpret_cast<const bf16x4*>(p); }

// Transposes a 16x16 block held in D layout (lane l: rows 4 (l/16) + i,
// column l % 16) into A/B layout (lane l: row l % 16, columns 4 (l/16) + j)
// through the wavefront's own 512-byte LDS scratch. LDS operations of one
// wavefront execute in order; the fences keep the compiler from moving the
// reads above the writes.
__device__ __forceinline__ bf16x4 transpose16(bf16x4 v, short* s, int l) {
  const int g = l >> 4, c = l & 15;
#pragma unroll
  for (int i = 0; i < 4; ++i) s[(4 * g + i) * 16 + c] = v[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const bf16x4 r = lds4(s + c * 16 + 4 * g);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return r;
}

struct MlpLds {
  short w1[kHid * kInPad];     // W1[o][i]: A operand of layer 1
  short w2[kHid * kHid];       // W2[o][i]: A of layer 2
  short w3[kOutPad * kHid];    // W3[o][h]: A of layer 3
  short w2t[kHid * kHid];      // W2^T[i][o]: A of dH1 = W2^T dZ2
  short w3t[kHid * kOutPad];   // W3^T[h][o]: A of dH2 = W3^T dY
  float b1[kHid], b2[kHid], b3[kOutPad];
};

// Weights to BF16 in LDS (both orientations where the backward pass needs
// the transpose), biases in FP32.
__device__ void stage_weights(MlpLds& s, const float* __restrict__ P, bool backward) {
  for (int t = threadIdx.x; t < kHid * kInPad; t += blockDim.x) s.w1[t] = bf16_bits(P[oW1 + t]);
  for (int t = threadIdx.x; t < kHid * kHid; t += blockDim.x) {
    const short w = bf16_bits(P[oW2 + t]);
    s.w2[t] = w;
    if (backward) s.w2t[(t % kHid) * kHid + t / kHid] = w;
  }
  for (int t = threadIdx.x; t < kOutPad * kHid; t += blockDim.x) {
    const short w = bf16_bits(P[oW3 + t]);
    s.w3[t] = w;
    if (backward) s.w3t[(t % kHid) * kOutPad + t / kHid] = w;
  }
  for (int t = threadIdx.x; t < kHid; t += blockDim.x) {
    s.b1[t] = P[oB1 + t];
    s.b2[t] = P[oB2 + t];
  }
  for (int t = threadIdx.x; t < kOutPad; t += blockDim.x) s.b3[t] = P[oB3 + t];
}

// Forward pass of one 16-row tile. xb[kb]: the input as B operand
// (k = features 16 kb + 4 (l/16) + j, n = row l % 16). Keeps the layer-1/2
// pre-activations (D layout) for the backward pass.
struct Fwd {
  f32x4 z1[4], z2[4], y;
  bf16x4 h1[4], h2[4];
};

__device__ __forceinline__ void forward_tile(const MlpLds& s, const bf16x4 xb[2], int l, Fwd& f) {
  const int g = l >> 4, c = l & 15;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = s.b1[16 * mb + 4 * g + i];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mfma(lds4(s.w1 + (16 * mb + c) * kInPad + 16 * kb + 4 * g), xb[kb], acc);
    f.z1[mb] = acc;
    f32x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
    f.h1[mb] = to_bf16x4(h);
  }
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    f32x4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = s.b2[16 * mb + 4 * g + i];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(lds4(s.w2 + (16 * mb + c) * kHid + 16 * kb + 4 * g), f.h1[kb], acc);
    f.z2[mb] = acc;
    f32x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = gelu(acc[i]);
    f.h2[mb] = to_bf16x4(h);
  }
  f32x4 acc;
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = s.b3[4 * g + i];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = mfma(lds4(s.w3 + c * kHid + 16 * kb + 4 * g), f.h2[kb], acc);
  f.y = acc;  // lanes 0..15 hold the 3 outputs (i < 3) of row l % 16
}

// Input tile as the B operand: row n = l % 16 of the tile, features
// 16 kb + 4 (l/16) + j, normalised by (x - mean) * inv_std; padding and
// missing rows are zero. `outside` reports a feature outside [lo, hi]
// (normalised units) for the envelope test of inference.
__device__ __forceinline__ void load_inputs(const float* __restrict__ X, int64_t row, bool valid, int l,
                                            const float* __restrict__ norm, bf16x4 xb[2], bool& outside,
                                            float* xn_out = nullptr) {
  const int g = l >> 4;
  outside = false;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 16 * kb + 4 * g + j;
      float v = 0.0f;
      if (valid && k < kIn) {
        v = X[row * kIn + k];
        if (norm) {
          v = (v - norm[k]) * norm[kIn + k];
          outside |= !(v >= norm[2 * kIn + k] && v <= norm[3 * kIn + k]);
        }
      }
      if (xn_out) xn_out[kb * 4 + j] = v;
      xb[kb][j] = bf16_bits(v);
    }
}

// Training step: forward, mean squared error over the real outputs, and
// backward for the batch rows idx[0..n) (or rows 0..n) of X (normalised
// features, row-major n x 18) against T (normalised targets, n x 3). Each
// wavefront keeps its weight-gradient blocks in registers across its tiles
// and writes them to its row of G (waves x kParams partial sums, reduced by
// k_mlp_reduce); the loss sum is added into *loss.
__global__ __launch_bounds__(kBlock) void k_mlp_train(int n, const float* __restrict__ X, const float* __restrict__ T,
                                                    const int32_t* __restrict__ idx, const float* __restrict__ P,
                                                    float* __restrict__ G, double* __restrict__ loss) {
  __shared__ MlpLds s;
  __shared__ short scratch[kBlock / 64][256];
  stage_weights(s, P, true);
  __syncthreads();
  const int l = threadIdx.x & 63, wid = threadIdx.x >> 6, g = l >> 4, c = l & 15;
  short* sc = scratch[wid];
  const int tiles = (n + 15) / 16;
  const int wave = blockIdx.x * (kBlock / 64) + wid, waves = gridDim.x * (kBlock / 64);
  const float dscale = 2.0f / (3.0f * (float)n);

  f32x4 gw1[4][2], gw2[4][4], gw3[4], gb1[4], gb2[4], gb3;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gw1[a][0] = gw1[a][1] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) gw2[a][b] = zero;
    gw3[a] = gb1[a] = gb2[a] = zero;
  }
  gb3 = zero;
  float lsum = 0.0f;

  for (int t = wave; t < tiles; t += waves) {
    const int r = t * 16 + c;
    const bool valid = r < n;
    const int64_t row = valid ? (idx ? (int64_t)idx[r] : (int64_t)r) : 0;
    bf16x4 xb[2];
    bool outside;
    load_inputs(X, row, valid, l, nullptr, xb, outside);
    Fwd f;
    forward_tile(s, xb, l, f);

    // dL/dy for the 3 real outputs (rows of D held by lanes 0..15)
    f32x4 dy = zero;
    if (g == 0 && valid) {
#pragma unroll
      for (int i = 0; i < kOut; ++i) {
        const float e = f.y[i] - T[row * kOut + i];
        lsum += e * e;
        dy[i] = dscale * e;
      }
    }
    const bf16x4 dyb = to_bf16x4(dy);
    gb3 += dy;

    // dZ2 = (W3^T dY) * GELU'(Z2)
    bf16x4 dz2b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 d = mfma(lds4(s.w3t + (16 * mb + c) * kOutPad + 4 * g), dyb, zero);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] *= gelu_grad(f.z2[mb][i]);
      gb2[mb] += d;
      dz2b[mb] = to_bf16x4(d);
    }
    // dZ1 = (W2^T dZ2) * GELU'(Z1)
    bf16x4 dz1b[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 d = zero;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) d = mfma(lds4(s.w2t + (16 * mb + c) * kHid + 16 * kb + 4 * g), dz2b[kb], d);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[i] *= gelu_grad(f.z1[mb][i]);
      gb1[mb] += d;
      dz1b[mb] = to_bf16x4(d);
    }
    // weight gradients, batch as the reduction index: dW = dZ^T-block x
    // H-block, each operand transposed once (layer 3, then 1, then 2, so
    // few transposed blocks are live at a time)
    const bf16x4 dyt = transpose16(dyb, sc, l);
#pragma unroll
    for (int b = 0; b < 4; ++b) gw3[b] = mfma(dyt, transpose16(f.h2[b], sc, l), gw3[b]);
    bf16x4 xt[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) xt[b] = transpose16(xb[b], sc, l);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 a1 = transpose16(dz1b[mb], sc, l);
#pragma unroll
      for (int b = 0; b < 2; ++b) gw1[mb][b] = mfma(a1, xt[b], gw1[mb][b]);
    }
    bf16x4 h1t[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) h1t[b] = transpose16(f.h1[b], sc, l);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 a2 = transpose16(dz2b[mb], sc, l);
#pragma unroll
      for (int b = 0; b < 4; ++b) gw2[mb][b] = mfma(a2, h1t[b], gw2[mb][b]);
    }
  }

  // bias gradients: sum the 16 batch lanes of each row group
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        gb1[mb][i] += __shfl_xor(gb1[mb][i], off);
        gb2[mb][i] += __shfl_xor(gb2[mb][i], off);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) gb3[i] += __shfl_xor(gb3[i], off);
  }
  // this wavefront's gradient into its own row of the partial sums (every
  // parameter written, padding as zeros); k_mlp_reduce adds the rows in
  // wavefront order. D layout of every gradient block: row 4 g + i, column c
  float* Gw = G + (int64_t)wave * kParams;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 16 * mb + 4 * g + i;
#pragma unroll
      for (int b = 0; b < 2; ++b) Gw[oW1 + o * kInPad + 16 * b + c] = gw1[mb][b][i];
#pragma unroll
      for (int b = 0; b < 4; ++b) Gw[oW2 + o * kHid + 16 * b + c] = gw2[mb][b][i];
      if (c == 0) {
        Gw[oB1 + o] = gb1[mb][i];
        Gw[oB2 + o] = gb2[mb][i];
      }
    }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = 4 * g + i;
#pragma unroll
    for (int b = 0; b < 4; ++b) Gw[oW3 + o * kHid + 16 * b + c] = gw3[b][i];
    if (c == 0) Gw[oB3 + o] = gb3[i];
  }
  // loss: wavefront sum, one atomic
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (l == 0) atomicAdd(loss, (double)lsum);
}

// Gradient = sum of the wavefronts' rows in a fixed order (no atomics: the
// same batch gives the same gradient bit for bit), in two passes: group y
// of kReduceGroups adds the rows w = y, y + groups, ... (coalesced over the
// parameter index), then the groups' sums are added in group order.
constexpr int kReduceGroups = 64;

__global__ __launch_bounds__(kBlock) void k_mlp_reduce_rows(int waves, const float* __restrict__ partial,
                                                          float* __restrict__ sums) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (t >= kParams) return;
  float acc = 0.0f;
  for (int w = y; w < waves; w += kReduceGroups) acc += partial[(int64_t)w * kParams + t];
  sums[(int64_t)y * kParams + t] = acc;
}

__global__ __launch_bounds__(kBlock) void k_mlp_reduce_groups(const float* __restrict__ sums, float* __restrict__ G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kParams) return;
  float acc = 0.0f;
  for (int y = 0; y < kReduceGroups; ++y) acc += sums[(int64_t)y * kParams + t];
  G[t] = acc;
}

// Adam (D. P. Kingma and J. Ba, "Adam: a method for stochastic
// optimization", ICLR 2015) on the FP32 parameters; step t >= 1.
__global__ __launch_bounds__(kBlock) void k_mlp_adam(float* __restrict__ P, const float* __restrict__ G,
                                                   float* __restrict__ M, float* __restrict__ V, float lr, float b1,
                                                   float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kParams) return;
  const float gr = G[t];
  const float m = b1 * M[t] + (1.0f - b1) * gr;
  const float v = b2 * V[t] + (1.0f - b2) * gr * gr;
  M[t] = m;
  V[t] = v;
  P[t] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// Inference: features X (raw, n x 18) -> correction out[a * stride + row]
// = mean_y[a] + std_y[a] * clamp(y[a], y_lo[a], y_hi[a]), or 0 with the row
// counted in *fallbacks when a normalised feature leaves the training
// envelope. norm: mean[18], inv_std[18], lo[18], hi[18], out_mean[3],
// out_std[3], y_lo[3], y_hi[3] (the training targets' range, normalised).
__global__ __launch_bounds__(kBlock) void k_mlp_infer(int n, const float* __restrict__ X, const float* __restrict__ P,
                                                    const float* __restrict__ norm, float* __restrict__ out,
                                                    int64_t stride, int32_t* __restrict__ fallbacks) {
  __shared__ MlpLds s;
  stage_weights(s, P, false);
  __syncthreads();
  const int l = threadIdx.x & 63, wid = threadIdx.x >> 6, g = l >> 4, c = l & 15;
  const int tiles = (n + 15) / 16;
  const int wave = blockIdx.x * (kBlock / 64) + wid, waves = gridDim.x * (kBlock / 64);
  int nfall = 0;
  for (int t = wave; t < tiles; t += waves) {
    const int r = t * 16 + c;
    const bool valid = r < n;
    bf16x4 xb[2];
    bool outside;
    load_inputs(X, valid ? r : 0, valid, l, norm, xb, outside);
    // a row is outside when any of its four lane groups saw a feature outside
    outside |= __shfl_xor((int)outside, 16) != 0;
    outside |= __shfl_xor((int)outside, 32) != 0;
    Fwd f;
    forward_tile(s, xb, l, f);
    if (g == 0 && valid) {
#pragma unroll
      for (int i = 0; i < kOut; ++i)
        out[i * stride + r] =
            outside ? 0.0f
                    : norm[4 * kIn + i] + norm[4 * kIn + kOut + i] * fminf(fmaxf(f.y[i], norm[4 * kIn + 2 * kOut + i]),
                                                                           norm[4 * kIn + 3 * kOut + i]);
      nfall += outside ? 1 : 0;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nfall += __shfl_xor(nfall, off);
  if (l == 0 && nfall) atomicAdd(fallbacks, nfall);
}

// Forward outputs (normalised units, n x 3) for the parity tests.
__global__ __launch_bounds__(kBlock) void k_mlp_forward(int n, const float* __restrict__ X, const float* __restrict__ P,
                                                      float* __restrict__ Y) {
  __shared__ MlpLds s;
  stage_weights(s, P, false);
  __syncthreads();
  const int l = threadIdx.x & 63, wid = threadIdx.x >> 6, g = l >> 4, c = l & 15;
  const int tiles = (n + 15) / 16;
  const int wave = blockIdx.x * (kBlock / 64) + wid, waves = gridDim.x * (kBlock / 64);
  for (int t = wave; t < tiles; t += waves) {
    const int r = t * 16 + c;
    const bool valid = r < n;
    bf16x4 xb[2];
    bool outside;
    load_inputs(X, valid ? r : 0, valid, l, nullptr, xb, outside);
    Fwd f;
    forward_tile(s, xb, l, f);
    if (g == 0 && valid) {
#pragma unroll
      for (int i = 0; i < kOut; ++i) Y[(int64_t)r * kOut + i] = f.y[i];
    }
  }
}

unsigned mlp_grid(int n) {
  const int tiles = (n + 15) / 16;
  // one tile per wavefront (4 per workgroup) up to 1024 workgroups (4 per
  // CU), then several per wavefront: a batch of 16,384 rows spreads over
  // all 256 CUs
  return (unsigned)std::max(1, std::min(1024, (tiles + 3) / 4));
}

}  // namespace
}  // namespace fsi

using namespace fsi;

extern "C" {

int fsi_mlp_train_workspace(int32_t n, int64_t* floats) {
  FSI_REQUIRE(n >= 0 && floats, "mlp_train_workspace: bad arguments");
  *floats = ((int64_t)mlp_grid(std::max(n, 1)) * (kBlock / 64) + kReduceGroups) * kParams;
  return 0;
}

int fsi_mlp_train_step(int32_t n, const float* X, const float* T, const int32_t* idx, const float* params,
                       float* grads, double* loss_sum, float* workspace, int64_t workspace_floats, void* stream) {
  FSI_REQUIRE(n >= 0, "mlp_train_step: negative batch size");
  FSI_REQUIRE(X && T && params && grads && loss_sum && workspace, "mlp_train_step: null buffer");
  const unsigned blocks = mlp_grid(std::max(n, 1));
  const int waves = (int)blocks * (kBlock / 64);
  const int64_t need = ((int64_t)waves + kReduceGroups) * kParams;
  FSI_REQUIRE(workspace_floats >= need, "mlp_train_step: workspace of %lld floats, need %lld",
              (long long)workspace_floats, (long long)need);
  hipStream_t st = (hipStream_t)stream;
  float* sums = workspace + (int64_t)waves * kParams;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(kBlock), 0, st, n, X, T, idx, params, workspace, loss_sum);
  FSI_TRY_LAUNCH("mlp_train_step");
  hipLaunchKernelGGL(k_mlp_reduce_rows, dim3(blocks_for(kParams), kReduceGroups), dim3(kBlock), 0, st, waves,
                     workspace, sums);
  FSI_TRY_LAUNCH("mlp_reduce_rows");
  hipLaunchKernelGGL(k_mlp_reduce_groups, dim3(blocks_for(kParams)), dim3(kBlock), 0, st, sums, grads);
  FSI_TRY_LAUNCH("mlp_reduce_groups");
  return 0;
}

int fsi_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr, float beta1,
                 float beta2, float eps, void* stream) {
  FSI_REQUIRE(params && grads && m && v, "mlp_adam: null buffer");
  FSI_REQUIRE(step >= 1, "mlp_adam: step counts from 1");
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)step)));
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kParams)), dim3(kBlock), 0, (hipStream_t)stream, params, grads, m, v,
                     lr, beta1, beta2, eps, c1, c2);
  FSI_TRY_LAUNCH("mlp_adam");
  return 0;
}

int fsi_mlp_forward(int32_t n, const float* X, const float* params, float* Y, void* stream) {
  FSI_REQUIRE(n >= 0, "mlp_forward: negative row count");
  FSI_REQUIRE(X && params && Y, "mlp_forward: null buffer");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(n)), dim3(kBlock), 0, (hipStream_t)stream, n, X, params, Y);
  FSI_TRY_LAUNCH("mlp_forward");
  return 0;
}

int fsi_mlp_infer(int32_t n, const float* X, const float* params, const float* norm, float* out, int64_t stride,
                  int32_t* fallbacks, void* stream) {
  FSI_REQUIRE(n >= 0, "mlp_infer: negative row count");
  FSI_REQUIRE(X && params && norm && out && fallbacks, "mlp_infer: null buffer");
  FSI_REQUIRE(stride >= n, "mlp_infer: output stride below the row count");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_grid(n)), dim3(kBlock), 0, (hipStream_t)stream, n, X, params, norm, out,
                     stride, fallbacks);
  FSI_TRY_LAUNCH("mlp_infer");
  return 0;
}

}  // extern "C"
