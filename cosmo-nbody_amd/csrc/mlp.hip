// Learned opening angle: group features, a 10 -> 64 -> 64 -> 1 MLP with
// BF16 weights and FP32 accumulation on MFMA (v_mfma_f32_16x16x16_bf16),
// its backward pass and an Adam step, and the per-group theta selection.
// No reference counterpart.
//
// Layout: a wavefront handles 16 rows at a time and keeps activations
// transposed (H^T: hidden units x rows), so each layer is
//   H_out^T [16-unit tile x 16 rows] = W^T tile (A operand) x H_in^T (B operand)
// and the MFMA output layout of one layer (lane l: row l%16, units
// 4*(l/16)+0..3 of the tile) is exactly the B-operand layout of the next:
// no shuffles between layers.  The weight-gradient products (K = rows) need
// the other orientation; their operands go through 2 KB of LDS per wave.
//
// Parameters live in one FP32 vector (master copy, Adam state alongside):
//   W1[16][64] (rows 10..15 unused, kept zero), b1[64], W2[64][64], b2[64],
//   W3[64], b3[1]   -> kParamCount floats,
// and a BF16 copy of W1 | W2 | W3 (kWeightCount values) that the kernels read.
#include <math.h>

#include "common.h"
#include "cosmo_mlp.h"

namespace cosmo {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = COSMO_MLP_IN_PAD;  // 16 (11 used)
constexpr int kHid = COSMO_MLP_HIDDEN; // 64
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;
static_assert(kParams == COSMO_MLP_PARAMS, "parameter layout");
constexpr int kBfW1 = 0, kBfW2 = kIn * kHid, kBfW3 = kBfW2 + kHid * kHid;
constexpr int kBfCount = kBfW3 + kHid;
static_assert(kBfCount == COSMO_MLP_BF16_WEIGHTS, "bf16 layout");

// FP32 -> BF16 with round-to-nearest-even (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float x0, float x1, float x2, float x3) {
  bf16x4 v;
  v[0] = static_cast<short>(f2bf(x0));
  v[1] = static_cast<short>(f2bf(x1));
  v[2] = static_cast<short>(f2bf(x2));
  v[3] = static_cast<short>(f2bf(x3));
  return v;
}

// A operand of W^T for output tile `mt`, input chunk `kt`: lane l holds
// W[k = 16 kt + 4 (l/16) + j][m = 16 mt + l%16], j = 0..3 (W row-major [K][N]).
__device__ __forceinline__ bf16x4 load_wt(const uint16_t *__restrict__ w, int n_cols, int mt,
                                          int kt, int lane) {
  const int m = 16 * mt + (lane & 15);
  const int k0 = 16 * kt + 4 * (lane >> 4);
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(w[(k0 + j) * n_cols + m]);
  return v;
}

// A operand of W (not transposed) for the input-gradient product
// dH^T = W dZ^T: lane l holds W[k = 16 kt + l%16][m = 16 mt + 4 (l/16) + j].
__device__ __forceinline__ bf16x4 load_w(const uint16_t *__restrict__ w, int n_cols, int kt,
                                         int mt, int lane) {
  const int k = 16 * kt + (lane & 15);
  const int m0 = 16 * mt + 4 * (lane >> 4);
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(w[k * n_cols + m0 + j]);
  return v;
}

struct Fwd {
  f32x4 z1[4], z2[4];  // pre-activations, transposed tiles
  bf16x4 h1[4];        // relu(z1) in BF16 (layer-2 input)
  float y;             // output for row lane%16 (all four lane groups)
};

// Forward pass of one 16-row chunk. x: B operand of X^T (lane: row l%16,
// features 4*(l/16)+0..3).
__device__ __forceinline__ void mlp_forward_chunk(bf16x4 x, const bf16x4 (&a1)[4],
                                                  const bf16x4 (&a2)[4][4],
                                                  const float *__restrict__ p, int lane,
                                                  Fwd &f) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  const int g4 = 4 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f.z1[t] = mfma16(a1[t], x, zero);
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      f.z1[t][j] += p[kOffB1 + 16 * t + g4 + j];
      h[j] = f.z1[t][j] > 0.f ? f.z1[t][j] : 0.f;
    }
    f.h1[t] = pack4(h[0], h[1], h[2], h[3]);
  }
  float y = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc = zero;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = mfma16(a2[t][kt], f.h1[kt], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int u = 16 * t + g4 + j;
      acc[j] += p[kOffB2 + u];
      const float h2 = acc[j] > 0.f ? acc[j] : 0.f;
      y += bf2f(f2bf(p[kOffW3 + u])) * h2;
    }
    f.z2[t] = acc;
  }
  y += __shfl_xor(y, 16, kWave);
  y += __shfl_xor(y, 32, kWave);
  f.y = y + p[kOffB3];
}

__device__ __forceinline__ bf16x4 load_x(const uint16_t *__restrict__ X, int64_t row,
                                         int64_t rows, int lane) {
  bf16x4 x = {0, 0, 0, 0};
  if (row < rows) {
    const uint2 raw = *reinterpret_cast<const uint2 *>(X + row * kIn + 4 * (lane >> 4));
    x[0] = static_cast<short>(raw.x & 0xffffu);
    x[1] = static_cast<short>(raw.x >> 16);
    x[2] = static_cast<short>(raw.y & 0xffffu);
    x[3] = static_cast<short>(raw.y >> 16);
  }
  return x;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t *__restrict__ X,
                                                     int64_t rows,
                                                     const uint16_t *__restrict__ wbf,
                                                     const float *__restrict__ p,
                                                     float *__restrict__ Y) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * blockDim.x / kWave;
  bf16x4 a1[4], a2[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a1[t] = load_wt(wbf + kBfW1, kHid, t, 0, lane);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) a2[t][kt] = load_wt(wbf + kBfW2, kHid, t, kt, lane);
  }
  for (int64_t chunk = wave; chunk * 16 < rows; chunk += nwaves) {
    const int64_t row = chunk * 16 + (lane & 15);
    Fwd f;
    mlp_forward_chunk(load_x(X, row, rows, lane), a1, a2, p, lane, f);
    if (lane < 16 && row < rows) Y[row] = f.y;
  }
}

// Forward + backward of the mean squared error over `rows` samples.  Each
// block writes its gradient sum to partial[block][kParams] and its squared
// error sum to partial[block][kParams]; a second kernel reduces over blocks
// in a fixed order.
__global__ __launch_bounds__(256) void k_mlp_grad(const uint16_t *__restrict__ X,
                                                  const float *__restrict__ T, int64_t rows,
                                                  const uint16_t *__restrict__ wbf,
                                                  const float *__restrict__ p,
                                                  float *__restrict__ partial) {
  __shared__ uint16_t lds_h[4][16][kHid];  // per wave: 16 rows x 64 units
  __shared__ uint16_t lds_d[4][16][kHid];
  __shared__ float red[kParams + 1];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int g4 = 4 * (lane >> 4), c16 = lane & 15;
  const int64_t gwave = int64_t(blockIdx.x) * 4 + wv;
  const int64_t nwaves = int64_t(gridDim.x) * 4;
  const float inv_rows = 1.0f / static_cast<float>(rows);
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};

  bf16x4 a1[4], a2[4][4], w2n[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a1[t] = load_wt(wbf + kBfW1, kHid, t, 0, lane);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      a2[t][kt] = load_wt(wbf + kBfW2, kHid, t, kt, lane);
      w2n[t][kt] = load_w(wbf + kBfW2, kHid, t, kt, lane);  // W2[16t + l%16][16kt + g4 + j]
    }
  }
  f32x4 gw2[4][4], gw1[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gw1[i] = zero;
#pragma unroll
    for (int j = 0; j < 4; ++j) gw2[i][j] = zero;
  }
  float gw3[4][4], gb2[4][4], gb1[4][4], gb3 = 0.f, sq = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) gw3[t][j] = gb2[t][j] = gb1[t][j] = 0.f;

  for (int64_t chunk = gwave; chunk * 16 < rows; chunk += nwaves) {
    const int64_t row = chunk * 16 + c16;
    const bool valid = row < rows;
    const bf16x4 x = load_x(X, row, rows, lane);
    Fwd f;
    mlp_forward_chunk(x, a1, a2, p, lane, f);
    const float err = valid ? (f.y - T[row]) : 0.f;
    if (lane < 16) sq += err * err;
    const float dy = 2.0f * err * inv_rows;
    if (lane < 16) gb3 += dy;
    // layer 3 and the layer-2 pre-activation gradient (transposed layout)
    float dz2[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int u = 16 * t + g4 + j;
        const float z = f.z2[t][j];
        gw3[t][j] += dy * (z > 0.f ? z : 0.f);
        const float d = (z > 0.f) ? dy * bf2f(f2bf(p[kOffW3 + u])) : 0.f;
        dz2[t][j] = d;
        gb2[t][j] += d;
      }
    // stage h1 and dz2 (BF16) row-major in LDS for the weight-gradient products
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        lds_h[wv][c16][16 * t + g4 + j] = static_cast<uint16_t>(f.h1[t][j]);
        lds_d[wv][c16][16 * t + g4 + j] = f2bf(dz2[t][j]);
      }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS writes of this wave done
    __builtin_amdgcn_wave_barrier();
    // dW2[k][m] += sum_rows h1[row][k] dz2[row][m]:  A = h1^T tile (k x rows),
    // B = dz2 tile (rows x m); lane holds rows 4*(l/16)+j.
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = static_cast<short>(lds_h[wv][g4 + j][16 * kt + c16]);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        bf16x4 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = static_cast<short>(lds_d[wv][g4 + j][16 * mt + c16]);
        gw2[kt][mt] = mfma16(a, b, gw2[kt][mt]);
      }
    }
    // dH1^T = W2 dZ2^T: A = W2 tile (k x m), B = dz2^T in BF16 (transposed layout)
    bf16x4 dz2b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) dz2b[t] = pack4(dz2[t][0], dz2[t][1], dz2[t][2], dz2[t][3]);
    float dz1[4][4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      f32x4 acc = zero;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) acc = mfma16(w2n[kt][mt], dz2b[mt], acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float d = (f.z1[kt][j] > 0.f) ? acc[j] : 0.f;
        dz1[kt][j] = d;
        gb1[kt][j] += d;
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) lds_d[wv][c16][16 * t + g4 + j] = f2bf(dz1[t][j]);
    // X tile row-major for the A operand of dW1 = X^T dZ1
#pragma unroll
    for (int j = 0; j < 4; ++j) lds_h[wv][c16][g4 + j] = static_cast<uint16_t>(x[j]);
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    {
      bf16x4 a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = static_cast<short>(lds_h[wv][g4 + j][c16]);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        bf16x4 b;
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = static_cast<short>(lds_d[wv][g4 + j][16 * mt + c16]);
        gw1[mt] = mfma16(a, b, gw1[mt]);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  // Block reduction in a fixed order: wave 0 writes, waves 1..3 add in turn.
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
      // C layout of the dW products: lane holds (k = 16 kt + g4 + j, m = 16 mt + l%16)
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int idx = kOffW2 + (16 * kt + g4 + j) * kHid + 16 * mt + c16;
            red[idx] = (w == 0 ? 0.f : red[idx]) + gw2[kt][mt][j];
          }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int idx = kOffW1 + (g4 + j) * kHid + 16 * mt + c16;
          red[idx] = (w == 0 ? 0.f : red[idx]) + gw1[mt][j];
        }
      // per-lane sums over rows -> reduce over the 16 lanes of each group
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float a = gw3[t][j], b2 = gb2[t][j], b1 = gb1[t][j];
#pragma unroll
          for (int off = 1; off < 16; off <<= 1) {
            a += __shfl_xor(a, off, kWave);
            b2 += __shfl_xor(b2, off, kWave);
            b1 += __shfl_xor(b1, off, kWave);
          }
          if (c16 == 0) {
            const int u = 16 * t + g4 + j;
            red[kOffW3 + u] = (w == 0 ? 0.f : red[kOffW3 + u]) + a;
            red[kOffB2 + u] = (w == 0 ? 0.f : red[kOffB2 + u]) + b2;
            red[kOffB1 + u] = (w == 0 ? 0.f : red[kOffB1 + u]) + b1;
          }
        }
      float s3 = gb3, s = sq;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        s3 += __shfl_xor(s3, off, kWave);
        s += __shfl_xor(s, off, kWave);
      }
      if (lane == 0) {
        red[kOffB3] = (w == 0 ? 0.f : red[kOffB3]) + s3;
        red[kParams] = (w == 0 ? 0.f : red[kParams]) + s;
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i <= kParams; i += blockDim.x)
    partial[int64_t(blockIdx.x) * (kParams + 1) + i] = red[i];
}

// grad[i] = sum over blocks of partial[b][i] in block order; the last entry
// is the sum of squared errors.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ partial, int nblocks,
                                                    float *__restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > kParams) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += partial[int64_t(b) * (kParams + 1) + i];
  grad[i] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters, then the BF16 copy.
// W1 rows 10..15 stay zero.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ m,
                                                  float *__restrict__ v,
                                                  const float *__restrict__ grad, float lr,
                                                  float beta1, float beta2, float eps,
                                                  float bc1, float bc2,
                                                  uint16_t *__restrict__ wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const bool pad = (i < kOffB1) && ((i / kHid) >= COSMO_MLP_IN);
  float w = p[i];
  if (!pad) {
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  } else {
    w = 0.f;
  }
  p[i] = w;
  if (i < kOffB1) wbf[kBfW1 + i] = f2bf(w);
  else if (i >= kOffW2 && i < kOffB2) wbf[kBfW2 + (i - kOffW2)] = f2bf(w);
  else if (i >= kOffW3 && i < kOffB3) wbf[kBfW3 + (i - kOffW3)] = f2bf(w);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(const float *__restrict__ p,
                                                     uint16_t *__restrict__ wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  if (i < kOffB1) wbf[kBfW1 + i] = f2bf(p[i]);
  else if (i >= kOffW2 && i < kOffB2) wbf[kBfW2 + (i - kOffW2)] = f2bf(p[i]);
  else if (i >= kOffW3 && i < kOffB3) wbf[kBfW3 + (i - kOffW3)] = f2bf(p[i]);
}

// ---------------------------------------------------------------- features

__device__ __forceinline__ float min_image(float d) { return d - rintf(d); }

__device__ __forceinline__ float node_extent(const float *__restrict__ lo,
                                             const float *__restrict__ hi, int64_t k) {
  const float e0 = hi[3 * k] - lo[3 * k], e1 = hi[3 * k + 1] - lo[3 * k + 1],
              e2 = hi[3 * k + 2] - lo[3 * k + 2];
  float e = e0 > e1 ? e0 : e1;
  return e > e2 ? e : e2;
}

// Straddle profile: the nodes that the AABB test's straddle limit lets
// across a minimum-image boundary are (down to the size test) the
// straddle-level nodes, lighter than the limit with a parent at or above it.
// Along each axis, bin b of the profile holds the sums of (M/limit)^2 and
// M/limit over the straddle-level nodes whose box spans the bin, in 2^-32
// fixed point (order-independent integer sums).  prof layout: [2][3][nbins],
// first the squared masses, then the masses; diff has the same layout.
constexpr double kProfScale = 4294967296.0;

__global__ __launch_bounds__(256) void k_profile_clear(unsigned long long *__restrict__ d,
                                                       int len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) d[i] = 0ull;
}

__global__ __launch_bounds__(256) void k_profile_nodes(
    const int32_t *__restrict__ parents, const double *__restrict__ n_mass,
    const float *__restrict__ n_min, const float *__restrict__ n_max, int64_t n, double limit,
    int nbins, unsigned long long *__restrict__ diff) {
  const int64_t k = n + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= 2 * n - 1) return;
  const double m = n_mass[k];
  if (!(m < limit)) return;
  const int p = parents[k];
  if (p >= 0 && !(n_mass[p] >= limit)) return;
  const double r = m / limit;
  const unsigned long long w[2] = {static_cast<unsigned long long>(r * r * kProfScale),
                                   static_cast<unsigned long long>(r * kProfScale)};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    int b0 = static_cast<int>(n_min[3 * k + ax] * static_cast<float>(nbins));
    int b1 = static_cast<int>(n_max[3 * k + ax] * static_cast<float>(nbins));
    b0 = b0 < 0 ? 0 : (b0 > nbins - 1 ? nbins - 1 : b0);
    b1 = b1 < 0 ? 0 : (b1 > nbins - 1 ? nbins - 1 : b1);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      unsigned long long *row = diff + (q * 3 + ax) * nbins;
      atomicAdd(&row[b0], w[q]);
      if (b1 + 1 < nbins) atomicAdd(&row[b1 + 1], 0ull - w[q]);
    }
  }
}

// Prefix sums of the six difference rows: one workgroup per row, each
// thread a contiguous chunk of bins; the chunk totals are scanned in LDS
// (Hillis-Steele).  Integer sums, so the result does not depend on the
// association.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_profile_scan(
    const unsigned long long *__restrict__ diff, int nbins,
    unsigned long long *__restrict__ prof) {
  __shared__ unsigned long long part[kScanThreads];
  const int t = threadIdx.x;
  const unsigned long long *d = diff + int64_t(blockIdx.x) * nbins;
  unsigned long long *o = prof + int64_t(blockIdx.x) * nbins;
  const int per = (nbins + kScanThreads - 1) / kScanThreads;
  const int b0 = t * per;
  const int b1 = min(b0 + per, nbins);
  unsigned long long sum = 0ull;
  for (int b = b0; b < b1; ++b) sum += d[b];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const unsigned long long v = t >= off ? part[t - off] : 0ull;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned long long run = part[t] - sum;  // exclusive prefix of this chunk
  for (int b = b0; b < b1; ++b) {
    run += d[b];
    o[b] = run;
  }
}

// The straddle-level nodes of up to kMaxProfiles limits in one pass over the
// nodes (each node read once): profile j gets the node's weights when the
// node is lighter than limits[j] and its parent is not.  Same sums as
// k_profile_nodes per limit (integer atomics, order-independent).
constexpr int kMaxProfiles = 8;
struct ProfileLimits {
  double limit[kMaxProfiles];
  int count;
};

__global__ __launch_bounds__(256) void k_profile_nodes_multi(
    const int32_t *__restrict__ parents, const double *__restrict__ n_mass,
    const float *__restrict__ n_min, const float *__restrict__ n_max, int64_t n,
    ProfileLimits lim, int nbins, unsigned long long *__restrict__ diff) {
  const int64_t k = n + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= 2 * n - 1) return;
  const double m = n_mass[k];
  const int p = parents[k];
  const double mp = p >= 0 ? n_mass[p] : 0.0;
  int b0[3], b1[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    int lo = static_cast<int>(n_min[3 * k + ax] * static_cast<float>(nbins));
    int hi = static_cast<int>(n_max[3 * k + ax] * static_cast<float>(nbins));
    b0[ax] = lo < 0 ? 0 : (lo > nbins - 1 ? nbins - 1 : lo);
    b1[ax] = hi < 0 ? 0 : (hi > nbins - 1 ? nbins - 1 : hi);
  }
  for (int j = 0; j < lim.count; ++j) {
    const double limit = lim.limit[j];
    if (!(m < limit)) continue;
    if (p >= 0 && !(mp >= limit)) continue;
    const double r = m / limit;
    const unsigned long long w[2] = {static_cast<unsigned long long>(r * r * kProfScale),
                                     static_cast<unsigned long long>(r * kProfScale)};
    unsigned long long *dj = diff + int64_t(j) * 6 * nbins;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        unsigned long long *row = dj + (q * 3 + ax) * nbins;
        atomicAdd(&row[b0[ax]], w[q]);
        if (b1[ax] + 1 < nbins) atomicAdd(&row[b1[ax] + 1], 0ull - w[q]);
      }
    }
  }
}

// One wavefront per group: the group's box from coalesced loads (one
// particle per lane, exact min/max reductions), the ancestor features by
// lane 0, and the straddle-profile features of each of the J profiles by
// lane j.  feat: J blocks of [groups][9]; block j equals what
// k_group_features computes with profile j (the same operations).
__global__ __launch_bounds__(256) void k_group_features_multi(
    const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
    const int32_t *__restrict__ children, const int32_t *__restrict__ parents,
    const double *__restrict__ n_mass, const double *__restrict__ n_com,
    const float *__restrict__ n_min, const float *__restrict__ n_max, int64_t n, float a,
    const float *__restrict__ a_dev, float n_grid, const unsigned long long *__restrict__ prof,
    int nprof, int nbins, int64_t groups, float *__restrict__ feat) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t g = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (g >= groups) return;  // wave-uniform
  const int64_t b0 = g * 64;
  const int64_t b1 = (b0 + 64 < n) ? b0 + 64 : n;
  const int cnt = static_cast<int>(b1 - b0);
  // lanes past the group's end repeat its first particle (min/max unchanged)
  const int64_t i = (b0 + lane < b1) ? b0 + lane : b0;
  float lo[3] = {px[i], py[i], pz[i]}, hi[3] = {lo[0], lo[1], lo[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const float l = __shfl_xor(lo[k], off), h = __shfl_xor(hi[k], off);
      lo[k] = l < lo[k] ? l : lo[k];
      hi[k] = h > hi[k] ? h : hi[k];
    }
  }
  const float floor_len = 1.0f / (64.0f * n_grid);
  float ext = 0.f, vol = 1.f, cen[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float e = hi[k] - lo[k];
    ext = e > ext ? e : ext;
    vol *= (e > floor_len ? e : floor_len);
    cen[k] = 0.5f * (hi[k] + lo[k]);
  }
  float f[9] = {};
  f[0] = log2f(fmaxf(ext * n_grid, 1e-6f));
  f[1] = log2f(static_cast<float>(cnt) / (vol * static_cast<float>(n)));
  f[2] = a_dev ? a_dev[0] : a;
  if (lane == 0) {
    // smallest ancestor of the group's first leaf holding >= cnt particles' mass
    const double need = (static_cast<double>(cnt) - 0.5) / static_cast<double>(n);
    int64_t node = b0;
    while (n_mass[node] < need) {
      const int p = parents[node];
      if (p < 0) break;
      node = p;
    }
    for (int lvl = 0; lvl < 2; ++lvl) {
      const int p = parents[node];
      float r = 0.f, mlog = 0.f;
      if (p >= 0) {
        const int2 ch = reinterpret_cast<const int2 *>(children)[p];
        const int64_t s = (ch.x == node) ? ch.y : ch.x;
        const float dx = min_image(static_cast<float>(n_com[3 * s]) - cen[0]);
        const float dy = min_image(static_cast<float>(n_com[3 * s + 1]) - cen[1]);
        const float dz = min_image(static_cast<float>(n_com[3 * s + 2]) - cen[2]);
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        r = node_extent(n_min, n_max, s) / fmaxf(d, 1e-6f);
        mlog = log2f(fmaxf(static_cast<float>(n_mass[s]) * static_cast<float>(n) /
                               static_cast<float>(cnt), 1e-6f));
        node = p;
      }
      f[3 + 2 * lvl] = r;
      f[4 + 2 * lvl] = mlog;
    }
  }
#pragma unroll
  for (int k = 3; k < 7; ++k) f[k] = __shfl(f[k], 0);
  // lane j < nprof: the features of profile j (bins at the group centre +
  // 1/2 in each axis)
  if (lane < nprof) {
    const unsigned long long *pj = prof + int64_t(lane) * 6 * nbins;
    float s2 = 0.f, s1 = 0.f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      float x = cen[ax] + 0.5f;
      x -= floorf(x);
      int b = static_cast<int>(x * static_cast<float>(nbins));
      b = b < 0 ? 0 : (b > nbins - 1 ? nbins - 1 : b);
      s2 += static_cast<float>(pj[ax * nbins + b]);
      s1 += static_cast<float>(pj[(3 + ax) * nbins + b]);
    }
    f[7] = log2f(1.0f + s2 * 2.3283064365386963e-10f);
    f[8] = log2f(1.0f + s1 * 2.3283064365386963e-10f);
    float *o = feat + (int64_t(lane) * groups + g) * COSMO_MLP_GROUP_FEATURES;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = f[k];
  }
}

// One thread per group of 64 Morton-consecutive particles: the nine
// per-group features (the tenth input is the candidate theta).
__global__ __launch_bounds__(256) void k_group_features(
    const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
    const int32_t *__restrict__ children, const int32_t *__restrict__ parents,
    const double *__restrict__ n_mass, const double *__restrict__ n_com,
    const float *__restrict__ n_min, const float *__restrict__ n_max, int64_t n, float a,
    const float *__restrict__ a_dev, float n_grid, const unsigned long long *__restrict__ prof,
    int nbins, float *__restrict__ feat) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t b0 = g * 64;
  if (b0 >= n) return;
  const int64_t b1 = (b0 + 64 < n) ? b0 + 64 : n;
  const int cnt = static_cast<int>(b1 - b0);
  float lo[3] = {px[b0], py[b0], pz[b0]}, hi[3] = {lo[0], lo[1], lo[2]};
  for (int64_t i = b0 + 1; i < b1; ++i) {
    const float q[3] = {px[i], py[i], pz[i]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = q[k] < lo[k] ? q[k] : lo[k];
      hi[k] = q[k] > hi[k] ? q[k] : hi[k];
    }
  }
  const float floor_len = 1.0f / (64.0f * n_grid);
  float ext = 0.f, vol = 1.f, cen[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float e = hi[k] - lo[k];
    ext = e > ext ? e : ext;
    vol *= (e > floor_len ? e : floor_len);
    cen[k] = 0.5f * (hi[k] + lo[k]);
  }
  float *f = feat + g * COSMO_MLP_GROUP_FEATURES;
  f[0] = log2f(fmaxf(ext * n_grid, 1e-6f));
  f[1] = log2f(static_cast<float>(cnt) / (vol * static_cast<float>(n)));
  f[2] = a_dev ? a_dev[0] : a;
  // smallest ancestor of the group's first leaf holding >= cnt particles' mass
  const double need = (static_cast<double>(cnt) - 0.5) / static_cast<double>(n);
  int64_t node = b0;
  while (n_mass[node] < need) {
    const int p = parents[node];
    if (p < 0) break;
    node = p;
  }
  for (int lvl = 0; lvl < 2; ++lvl) {
    const int p = parents[node];
    float r = 0.f, mlog = 0.f;
    if (p >= 0) {
      const int2 ch = reinterpret_cast<const int2 *>(children)[p];
      const int64_t s = (ch.x == node) ? ch.y : ch.x;
      const float dx = min_image(static_cast<float>(n_com[3 * s]) - cen[0]);
      const float dy = min_image(static_cast<float>(n_com[3 * s + 1]) - cen[1]);
      const float dz = min_image(static_cast<float>(n_com[3 * s + 2]) - cen[2]);
      const float d = sqrtf(dx * dx + dy * dy + dz * dz);
      r = node_extent(n_min, n_max, s) / fmaxf(d, 1e-6f);
      mlog = log2f(fmaxf(static_cast<float>(n_mass[s]) * static_cast<float>(n) /
                             static_cast<float>(cnt), 1e-6f));
      node = p;
    }
    f[3 + 2 * lvl] = r;
    f[4 + 2 * lvl] = mlog;
  }
  // the straddle profile where the group's minimum-image boundaries lie,
  // half a box from its centre in each axis
  float s2 = 0.f, s1 = 0.f;
  if (prof) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      float x = cen[ax] + 0.5f;
      x -= floorf(x);
      int b = static_cast<int>(x * static_cast<float>(nbins));
      b = b < 0 ? 0 : (b > nbins - 1 ? nbins - 1 : b);
      s2 += static_cast<float>(prof[ax * nbins + b]);
      s1 += static_cast<float>(prof[(3 + ax) * nbins + b]);
    }
  }
  f[7] = log2f(1.0f + s2 * 2.3283064365386963e-10f);
  f[8] = log2f(1.0f + s1 * 2.3283064365386963e-10f);
}

// Rows g*K + k of the MLP input: [theta_k, features of g..., cand2_k, 0 pad]
// in BF16 (cand2: the second coordinate of candidate k, or 0).  With fidx
// the features of candidate k come from block fidx[k] of feat ([blocks][groups][9]).
__global__ __launch_bounds__(256) void k_assemble_rows(const float *__restrict__ feat,
                                                       int64_t groups,
                                                       const float *__restrict__ thetas,
                                                       const float *__restrict__ cand2,
                                                       const int32_t *__restrict__ fidx, int K,
                                                       uint16_t *__restrict__ X) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= groups * K) return;
  const int64_t g = r / K;
  const int k = static_cast<int>(r % K);
  const int64_t blk = fidx ? fidx[k] : 0;
  const float *f = feat + (blk * groups + g) * COSMO_MLP_GROUP_FEATURES;
  uint16_t *x = X + r * kIn;
  x[0] = f2bf(thetas[k]);
#pragma unroll
  for (int j = 0; j < COSMO_MLP_GROUP_FEATURES; ++j) x[1 + j] = f2bf(f[j]);
  x[1 + COSMO_MLP_GROUP_FEATURES] = cand2 ? f2bf(cand2[k]) : uint16_t(0);
#pragma unroll
  for (int j = 2 + COSMO_MLP_GROUP_FEATURES; j < kIn; ++j) x[j] = 0;
}

// k* = argmin_k cost[k] + lambda * 2^pred[g*K+k]: theta_group[g] =
// thetas[k*] and, when given, mu_group[g] = mus[k*] and k_group[g] = k*;
// any non-finite prediction of the group selects the fixed fallbacks
// (k_group -1).
__global__ __launch_bounds__(256) void k_select_theta(const float *__restrict__ pred,
                                                      int64_t groups,
                                                      const float *__restrict__ thetas,
                                                      const float *__restrict__ mus,
                                                      const float *__restrict__ cost, int K,
                                                      float lambda, float fallback,
                                                      float fallback_mu,
                                                      float *__restrict__ theta_group,
                                                      float *__restrict__ mu_group,
                                                      int32_t *__restrict__ k_group,
                                                      unsigned long long *__restrict__ n_fallback) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= groups) return;
  float best = INFINITY;
  int kb = -1;
  bool ok = true;
  for (int k = 0; k < K; ++k) {
    const float pk = pred[g * K + k];
    if (!isfinite(pk)) ok = false;
    const float score = cost[k] + lambda * exp2f(pk);
    if (score < best) {
      best = score;
      kb = k;
    }
  }
  float th = fallback, mu = fallback_mu;
  int kk = -1;
  if (!ok || !isfinite(best) || kb < 0) {
    if (n_fallback) atomicAdd(n_fallback, 1ull);
  } else {
    th = thetas[kb];
    if (mus) mu = mus[kb];
    kk = kb;
  }
  theta_group[g] = th;
  if (mu_group) mu_group[g] = mu;
  if (k_group) k_group[g] = kk;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_mlp_forward(const uint16_t *X, int64_t rows, const uint16_t *wbf, const float *params,
                      float *Y, void *stream) {
  if (rows < 0) return COSMO_E_INVALID;
  if (rows == 0) return COSMO_OK;
  if (!X || !wbf || !params || !Y) return COSMO_E_INVALID;
  const int64_t chunks = (rows + 15) / 16;
  const int64_t blocks = (chunks + 3) / 4;
  const unsigned grid = static_cast<unsigned>(blocks < 4096 ? blocks : 4096);
  k_mlp_forward<<<grid, 256, 0, as_stream(stream)>>>(X, rows, wbf, params, Y);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_mlp_grad(const uint16_t *X, const float *T, int64_t rows, const uint16_t *wbf,
                   const float *params, float *partial, int nblocks, float *grad,
                   void *stream) {
  if (rows < 1 || nblocks < 1 || !X || !T || !wbf || !params || !partial || !grad)
    return COSMO_E_INVALID;
  hipStream_t s = as_stream(stream);
  k_mlp_grad<<<nblocks, 256, 0, s>>>(X, T, rows, wbf, params, partial);
  k_mlp_reduce<<<blocks_for(kParams + 1, 256), 256, 0, s>>>(partial, nblocks, grad);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_mlp_adam(float *params, float *m, float *v, const float *grad, float lr, float beta1,
                   float beta2, float eps, int step, uint16_t *wbf, void *stream) {
  if (!params || !m || !v || !grad || !wbf || step < 1) return COSMO_E_INVALID;
  const float bc1 = 1.f - powf(beta1, static_cast<float>(step));
  const float bc2 = 1.f - powf(beta2, static_cast<float>(step));
  k_mlp_adam<<<blocks_for(kParams, 256), 256, 0, as_stream(stream)>>>(params, m, v, grad, lr,
                                                                      beta1, beta2, eps, bc1,
                                                                      bc2, wbf);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_mlp_to_bf16(const float *params, uint16_t *wbf, void *stream) {
  if (!params || !wbf) return COSMO_E_INVALID;
  k_mlp_to_bf16<<<blocks_for(kParams, 256), 256, 0, as_stream(stream)>>>(params, wbf);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_straddle_profile(const int32_t *parents, const double *n_mass, const float *n_min,
                           const float *n_max, int64_t n, double straddle_mass,
                           uint64_t *profile, uint64_t *scratch, int nbins, void *stream) {
  if (n < 1 || n > kMaxBodies || nbins < 1 || nbins > COSMO_MLP_PROFILE_MAX_BINS ||
      !(straddle_mass > 0.0))
    return COSMO_E_INVALID;
  if (!parents || !n_mass || !n_min || !n_max || !profile || !scratch) return COSMO_E_INVALID;
  hipStream_t s = as_stream(stream);
  unsigned long long *diff = reinterpret_cast<unsigned long long *>(scratch);
  const int len = 6 * nbins;
  k_profile_clear<<<blocks_for(len, 256), 256, 0, s>>>(diff, len);
  if (n > 1)
    k_profile_nodes<<<blocks_for(n - 1, 256), 256, 0, s>>>(parents, n_mass, n_min, n_max, n,
                                                           straddle_mass, nbins, diff);
  k_profile_scan<<<6, kScanThreads, 0, s>>>(diff, nbins,
                                             reinterpret_cast<unsigned long long *>(profile));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

static int group_features_impl(const float *px, const float *py, const float *pz,
                               const int32_t *children, const int32_t *parents,
                               const double *n_mass, const double *n_com, const float *n_min,
                               const float *n_max, int64_t n, double a, const float *a_dev,
                               double n_grid, const uint64_t *profile, int nbins, float *feat,
                               void *stream) {
  if (n < 1 || n > kMaxBodies) return COSMO_E_INVALID;
  if (!px || !py || !pz || !children || !parents || !n_mass || !n_com || !n_min || !n_max ||
      !feat)
    return COSMO_E_INVALID;
  if (profile && (nbins < 1 || nbins > COSMO_MLP_PROFILE_MAX_BINS)) return COSMO_E_INVALID;
  const int64_t groups = (n + 63) / 64;
  k_group_features<<<blocks_for(groups, 256), 256, 0, as_stream(stream)>>>(
      px, py, pz, children, parents, n_mass, n_com, n_min, n_max, n, static_cast<float>(a),
      a_dev, static_cast<float>(n_grid),
      reinterpret_cast<const unsigned long long *>(profile), nbins, feat);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_group_features(const float *px, const float *py, const float *pz,
                         const int32_t *children, const int32_t *parents, const double *n_mass,
                         const double *n_com, const float *n_min, const float *n_max, int64_t n,
                         double a, double n_grid, const uint64_t *profile, int nbins,
                         float *feat, void *stream) {
  return group_features_impl(px, py, pz, children, parents, n_mass, n_com, n_min, n_max, n, a,
                             nullptr, n_grid, profile, nbins, feat, stream);
}

int cosmo_group_features_dev(const float *px, const float *py, const float *pz,
                             const int32_t *children, const int32_t *parents,
                             const double *n_mass, const double *n_com, const float *n_min,
                             const float *n_max, int64_t n, const float *a, double n_grid,
                             const uint64_t *profile, int nbins, float *feat, void *stream) {
  if (!a) return COSMO_E_INVALID;
  return group_features_impl(px, py, pz, children, parents, n_mass, n_com, n_min, n_max, n,
                             0.0, a, n_grid, profile, nbins, feat, stream);
}

int cosmo_straddle_profiles(const int32_t *parents, const double *n_mass, const float *n_min,
                            const float *n_max, int64_t n, const double *straddle_masses,
                            int count, uint64_t *profiles, uint64_t *scratch, int nbins,
                            void *stream) {
  if (n < 1 || n > kMaxBodies || nbins < 1 || nbins > COSMO_MLP_PROFILE_MAX_BINS ||
      count < 1 || count > COSMO_MLP_MAX_PROFILES || !straddle_masses)
    return COSMO_E_INVALID;
  if (!parents || !n_mass || !n_min || !n_max || !profiles || !scratch) return COSMO_E_INVALID;
  ProfileLimits lim;
  lim.count = count;
  for (int j = 0; j < kMaxProfiles; ++j) {
    lim.limit[j] = j < count ? straddle_masses[j] : 0.0;
    if (j < count && !(straddle_masses[j] > 0.0)) return COSMO_E_INVALID;
  }
  hipStream_t s = as_stream(stream);
  unsigned long long *diff = reinterpret_cast<unsigned long long *>(scratch);
  const int len = 6 * nbins * count;
  k_profile_clear<<<blocks_for(len, 256), 256, 0, s>>>(diff, len);
  if (n > 1)
    k_profile_nodes_multi<<<blocks_for(n - 1, 256), 256, 0, s>>>(parents, n_mass, n_min, n_max,
                                                                 n, lim, nbins, diff);
  // the prefix sums of every row of every profile: one workgroup per row
  k_profile_scan<<<6 * count, kScanThreads, 0, s>>>(
      diff, nbins, reinterpret_cast<unsigned long long *>(profiles));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_group_features_multi(const float *px, const float *py, const float *pz,
                               const int32_t *children, const int32_t *parents,
                               const double *n_mass, const double *n_com, const float *n_min,
                               const float *n_max, int64_t n, double a, const float *a_dev,
                               double n_grid, const uint64_t *profiles, int count, int nbins,
                               float *feat, void *stream) {
  if (n < 1 || n > kMaxBodies || count < 1 || count > COSMO_MLP_MAX_PROFILES ||
      nbins < 1 || nbins > COSMO_MLP_PROFILE_MAX_BINS)
    return COSMO_E_INVALID;
  if (!px || !py || !pz || !children || !parents || !n_mass || !n_com || !n_min || !n_max ||
      !profiles || !feat)
    return COSMO_E_INVALID;
  const int64_t groups = (n + 63) / 64;
  k_group_features_multi<<<blocks_for(groups * kWave, 256), 256, 0, as_stream(stream)>>>(
      px, py, pz, children, parents, n_mass, n_com, n_min, n_max, n, static_cast<float>(a),
      a_dev, static_cast<float>(n_grid), reinterpret_cast<const unsigned long long *>(profiles),
      count,
      nbins, groups, feat);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_assemble_rows(const float *feat, int64_t groups, const float *thetas,
                        const float *cand2, const int32_t *fidx, int K, uint16_t *X,
                        void *stream) {
  if (groups < 0 || K < 1 || (groups && (!feat || !thetas || !X))) return COSMO_E_INVALID;
  if (groups == 0) return COSMO_OK;
  k_assemble_rows<<<blocks_for(groups * K, 256), 256, 0, as_stream(stream)>>>(
      feat, groups, thetas, cand2, fidx, K, X);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_select_theta(const float *pred, int64_t groups, const float *thetas,
                       const float *mus, const float *cost, int K, double lambda,
                       double fallback, double fallback_mu, float *theta_group,
                       float *mu_group, int32_t *k_group, uint64_t *n_fallback, void *stream) {
  if (groups < 0 || K < 1 || (groups && (!pred || !thetas || !cost || !theta_group)) ||
      (mu_group && !mus))
    return COSMO_E_INVALID;
  if (groups == 0) return COSMO_OK;
  k_select_theta<<<blocks_for(groups, 256), 256, 0, as_stream(stream)>>>(
      pred, groups, thetas, mus, cost, K, static_cast<float>(lambda),
      static_cast<float>(fallback), static_cast<float>(fallback_mu), theta_group, mu_group,
      k_group, reinterpret_cast<unsigned long long *>(n_fallback));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
