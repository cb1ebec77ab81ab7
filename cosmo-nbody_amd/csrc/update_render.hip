// Kick-drift update and fixed-point density splat.
//
// Reference: src/kernels/physics_kernels.py:155-203 (integrate) and
// src/kernels/render_kernels.py:12-47 (render_density).  The library is
// compiled with -ffp-contract=off, so the FP32 expressions below round
// after every operation exactly as written; the CPU oracle is built the
// same way and the results are bit-identical.
#include <math.h>

#include <algorithm>

#include "common.h"

namespace cosmo {

// Coefficients come either as arguments (coefs == nullptr) or from device
// memory (graph replays).
__global__ __launch_bounds__(256) void k_kick_drift(float *__restrict__ px, float *__restrict__ py,
                                                   float *__restrict__ pz, float *__restrict__ vx,
                                                   float *__restrict__ vy, float *__restrict__ vz,
                                                   const float *__restrict__ force, float c_drag,
                                                   float c_kick, float c_drift,
                                                   const float *__restrict__ coefs, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (coefs) {
    c_drag = coefs[0];
    c_kick = coefs[1];
    c_drift = coefs[2];
  }
  const float ux = vx[i] * c_drag + force[3 * i + 0] * c_kick;
  const float uy = vy[i] * c_drag + force[3 * i + 1] * c_kick;
  const float uz = vz[i] * c_drag + force[3 * i + 2] * c_kick;
  vx[i] = ux;
  vy[i] = uy;
  vz[i] = uz;
  const float x = px[i] + ux * c_drift;
  const float y = py[i] + uy * c_drift;
  const float z = pz[i] + uz * c_drift;
  px[i] = x - floorf(x);
  py[i] = y - floorf(y);
  pz[i] = z - floorf(z);
}

struct Mvp {
  float m[16];
};

__device__ __forceinline__ void splat_one(float x, float y, float z, const float *m, int res,
                                          float scale, unsigned long long *grid) {
  const float xc = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
  const float yc = ((m[4] * x + m[5] * y) + m[6] * z) + m[7];
  const float wc = ((m[12] * x + m[13] * y) + m[14] * z) + m[15];
  if (!(wc > 0.01f)) return;
  const float resf = static_cast<float>(res);
  const float inv = 1.0f / wc;
  const float sx = ((xc * inv + 1.0f) * 0.5f) * resf;
  const float sy = ((1.0f - yc * inv) * 0.5f) * resf;
  if (sx >= 0.0f && sx < resf && sy >= 0.0f && sy < resf) {
    const float val = 2.5f * inv;
    const unsigned long long q = static_cast<unsigned long long>(rintf(val * scale));
    atomicAdd(&grid[int64_t(static_cast<int>(sy)) * res + static_cast<int>(sx)], q);
  }
}

// render_kernels.py:35-47 with the float atomic replaced by a u64 atomic on
// a fixed-point value, so the grid does not depend on arrival order.
__global__ __launch_bounds__(256) void k_render_density(const float *__restrict__ px,
                                                        const float *__restrict__ py,
                                                        const float *__restrict__ pz,
                                                        unsigned long long *__restrict__ grid,
                                                        Mvp mvp, int res, float scale, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  splat_one(px[i], py[i], pz[i], mvp.m, res, scale, grid);
}

__global__ __launch_bounds__(256) void k_render_density_dev(const float *__restrict__ px,
                                                            const float *__restrict__ py,
                                                            const float *__restrict__ pz,
                                                            unsigned long long *__restrict__ grid,
                                                            const float *__restrict__ mvp, int res,
                                                            float scale, int64_t n) {
  __shared__ float m[16];
  if (threadIdx.x < 16) m[threadIdx.x] = mvp[threadIdx.x];
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  splat_one(px[i], py[i], pz[i], m, res, scale, grid);
}

// count[0] += number of non-finite values among v[0..n) (NaN or +-inf).
__global__ __launch_bounds__(256) void k_count_nonfinite(const float *__restrict__ v, int64_t n,
                                                         unsigned long long *__restrict__ count) {
  unsigned long long c = 0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    c += isfinite(v[i]) ? 0ull : 1ull;
  // wave sum, one atomic per wavefront that saw any
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd(count, c);
}

// grid[p] += bins[p] * 2^-frac_bits: the fixed-point splat added to a float32
// density grid (camera.py:26).  The bin is exact in FP64 below 2^53 and the
// product is rounded once to FP32 before the add.
__global__ __launch_bounds__(256) void k_grid_add_fixed(const unsigned long long *__restrict__ bins,
                                                        float *__restrict__ grid, int64_t cells,
                                                        double scale) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const unsigned long long b = bins[i];
  if (b != 0ull) grid[i] += static_cast<float>(static_cast<double>(b) * scale);
}

// Colour mapping of a density frame (src/utils/visualizer.py:115-171), fused
// after the splat: img = log1p(120 x) / (max(img) + 1e-9), val = img^2.2,
// the four-band palette, x1.3, clipped to [0, 1]; x is the fixed-point bin
// in FP32.  log1p is increasing, so max(img) = log1p(120 max x) and the
// reduction is an exact integer maximum over the bins.  The FP32 constants
// are those the reference's float32 arrays meet: T1 = 0.6, T2 = 0.72,
// T3 = 0.82 and the band widths as rounded from their FP64 differences.
__global__ void k_grid_max_clear(unsigned long long *mx) {
  if (threadIdx.x == 0) mx[0] = 0ull;
}

__global__ __launch_bounds__(256) void k_grid_max(const unsigned long long *__restrict__ grid,
                                                  int64_t len, unsigned long long *mx) {
  __shared__ unsigned long long part[256];
  unsigned long long m = 0ull;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += int64_t(gridDim.x) * blockDim.x)
    m = grid[i] > m ? grid[i] : m;
  part[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const unsigned long long o = part[threadIdx.x + w];
      part[threadIdx.x] = o > part[threadIdx.x] ? o : part[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(mx, part[0]);
}

__global__ __launch_bounds__(256) void k_colorize(const unsigned long long *__restrict__ grid,
                                                  int64_t len, float scale,
                                                  const unsigned long long *__restrict__ mx,
                                                  float *__restrict__ frame,
                                                  uint8_t *__restrict__ rgb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const float top = log1pf((static_cast<float>(mx[0]) * scale) * 120.0f);
  const float img = log1pf((static_cast<float>(grid[i]) * scale) * 120.0f) / (top + 1e-9f);
  const float val = powf(img, 2.2f);
  const float T1 = 0.6f, T2 = 0.72f, T3 = 0.82f;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if (val <= T1) {
    b = (val / T1) * 0.8f;
  } else if (val <= T2) {
    const float t = (val - T1) / 0.12f;
    b = (1.0f - t) * 0.8f;
    g = t * 0.9f;
  } else if (val <= T3) {
    const float t = (val - T2) / 0.09999999999999998f;
    g = 0.9f + t * 0.1f;
    r = t;
  } else {
    const float t = (val - T3) / 0.18000000000000005f;
    r = 1.0f;
    g = 1.0f - t;
  }
  const float c[3] = {fminf(fmaxf(r * 1.3f, 0.0f), 1.0f), fminf(fmaxf(g * 1.3f, 0.0f), 1.0f),
                      fminf(fmaxf(b * 1.3f, 0.0f), 1.0f)};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (frame) frame[3 * i + k] = c[k];
    // engine.py:139: (frame * 255).astype(uint8), a truncation
    if (rgb) rgb[3 * i + k] = static_cast<uint8_t>(c[k] * 255.0f);
  }
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_kick_drift(float *px, float *py, float *pz, float *vx, float *vy, float *vz,
                     const float *force, float c_drag, float c_kick, float c_drift, int64_t n,
                     void *stream) {
  if (n < 0) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !vx || !vy || !vz || !force) return COSMO_E_INVALID;
  k_kick_drift<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(px, py, pz, vx, vy, vz, force,
                                                                   c_drag, c_kick, c_drift,
                                                                   nullptr, n);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_kick_drift_dev(float *px, float *py, float *pz, float *vx, float *vy, float *vz,
                         const float *force, const float *coefs, int64_t n, void *stream) {
  if (n < 0) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !vx || !vy || !vz || !force || !coefs) return COSMO_E_INVALID;
  k_kick_drift<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(px, py, pz, vx, vy, vz, force,
                                                                   0.f, 0.f, 0.f, coefs, n);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_integrate(float *px, float *py, float *pz, float *vx, float *vy, float *vz,
                    const float *force, double a, double H, double dt, int64_t n,
                    void *stream) {
  if (!(a > 0.0)) return COSMO_E_INVALID;
  // physics_kernels.py:180-195: drag = 1 - H dt, force and drift scaled by
  // dt/a; formed in FP64, rounded once to FP32.
  const float c_drag = static_cast<float>(1.0 - H * dt);
  const float c_kick = static_cast<float>(dt / a);
  return cosmo_kick_drift(px, py, pz, vx, vy, vz, force, c_drag, c_kick, c_kick, n, stream);
}

int cosmo_render_density(const float *px, const float *py, const float *pz, uint64_t *grid,
                         const float *mvp, int32_t res, int64_t n, int frac_bits,
                         void *stream) {
  if (n < 0 || res < 1 || res > 65536 || frac_bits < 0 || frac_bits > 30 || !mvp)
    return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !grid) return COSMO_E_INVALID;
  Mvp m;
  for (int k = 0; k < 16; ++k) m.m[k] = mvp[k];
  const float scale = ldexpf(1.0f, frac_bits);
  k_render_density<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
      px, py, pz, reinterpret_cast<unsigned long long *>(grid), m, res, scale, n);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_grid_add_fixed(const uint64_t *bins, float *grid, int64_t cells, int frac_bits,
                         void *stream) {
  if (cells < 0 || frac_bits < 0 || frac_bits > 30) return COSMO_E_INVALID;
  if (cells == 0) return COSMO_OK;
  if (!bins || !grid) return COSMO_E_INVALID;
  k_grid_add_fixed<<<blocks_for(cells, 256), 256, 0, as_stream(stream)>>>(
      reinterpret_cast<const unsigned long long *>(bins), grid, cells, ldexp(1.0, -frac_bits));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_count_nonfinite(const float *v, int64_t n, uint64_t *count, void *stream) {
  if (n < 0 || !count) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!v) return COSMO_E_INVALID;
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(blocks_for(n, 256), 4096));
  k_count_nonfinite<<<g, 256, 0, as_stream(stream)>>>(
      v, n, reinterpret_cast<unsigned long long *>(count));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_render_density_dev(const float *px, const float *py, const float *pz, uint64_t *grid,
                             const float *mvp_dev, int32_t res, int64_t n, int frac_bits,
                             void *stream) {
  if (n < 0 || res < 1 || res > 65536 || frac_bits < 0 || frac_bits > 30 || !mvp_dev)
    return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !grid) return COSMO_E_INVALID;
  k_render_density_dev<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(
      px, py, pz, reinterpret_cast<unsigned long long *>(grid), mvp_dev, res,
      ldexpf(1.0f, frac_bits), n);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_colorize_frame(const uint64_t *grid, int32_t res, int frac_bits, uint64_t *max_scratch,
                         float *frame, uint8_t *rgb, void *stream) {
  if (res < 1 || res > 65536 || frac_bits < 0 || frac_bits > 30) return COSMO_E_INVALID;
  if (!grid || !max_scratch || (!frame && !rgb)) return COSMO_E_INVALID;
  hipStream_t s = as_stream(stream);
  const int64_t len = int64_t(res) * res;
  unsigned long long *mx = reinterpret_cast<unsigned long long *>(max_scratch);
  const unsigned long long *g = reinterpret_cast<const unsigned long long *>(grid);
  k_grid_max_clear<<<1, 64, 0, s>>>(mx);
  const unsigned blocks = static_cast<unsigned>(len / 256 < 1024 ? (len + 255) / 256 : 1024);
  k_grid_max<<<blocks, 256, 0, s>>>(g, len, mx);
  k_colorize<<<blocks_for(len, 256), 256, 0, s>>>(g, len, ldexpf(1.0f, -frac_bits), mx, frame,
                                                   rgb);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
