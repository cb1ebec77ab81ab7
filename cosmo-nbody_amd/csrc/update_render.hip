// Kick-drift update and fixed-point density splat.
//
// Reference: src/kernels/physics_kernels.py:155-203 (integrate) and
// src/kernels/render_kernels.py:12-47 (render_density).  The library is
// compiled with -ffp-contract=off, so the FP32 expressions below round
// after every operation exactly as written; the CPU oracle is built the
// same way and the results are bit-identical.
#include <math.h>

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

}  // extern "C"
