// Zel'dovich initial conditions on the GPU.
//
// Reference: src/utils/cosmology.py:12-73 (generate_zeldovich_ics).  The
// phases are an input (the caller draws them, so that one phase array gives
// the same realisation on the host and here); everything after the phases
// runs on the device: the spectrum of each displacement component, three
// inverse FFTs (hipFFT, double complex, in place), the real parts, the
// standard deviation of the x component, and the lattice + displacement
// with the periodic wrap and the velocity coupling, rounded to FP32.  Every
// step restates the numpy expression order of the host version
// (src/utils/cosmology.py in this package), so the FP32 outputs agree with
// it except where the FFT libraries' different rounding crosses an FP32
// rounding boundary.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "common.h"

namespace cosmo {

// numpy.fft.fftfreq(n)[m] * n: integers (m or m - n) times 1/n, times n.
__device__ __forceinline__ double fft_wavenumber(int64_t m, int64_t n) {
  const int64_t half = (n - 1) / 2 + 1;
  const int64_t k = m < half ? m : m - n;
  return (static_cast<double>(k) * (1.0 / static_cast<double>(n))) * static_cast<double>(n);
}

// cosmology.py:27-45 for component `comp`: the spectrum
// (i k_c / (|k|^2 + 1e-10)) * sqrt(P(k)) e^{i phase} with P = |k|^-2.5,
// P(0) = 0, |k|(0) = 1e-10 and the k = 0 term zero.  ij ordering: the
// flat index is (i n + j) n + l with k_x from i.
__global__ __launch_bounds__(256) void k_za_spectrum(const double *__restrict__ phase,
                                                     double2 *__restrict__ out, int64_t n,
                                                     int comp) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n3 = n * n * n;
  if (idx >= n3) return;
  const int64_t i = idx / (n * n), j = (idx / n) % n, l = idx % n;
  const double kx = fft_wavenumber(i, n), ky = fft_wavenumber(j, n), kz = fft_wavenumber(l, n);
  const double k2 = kx * kx + ky * ky + kz * kz;
  double kmag = sqrt(k2);
  double P;
  if (idx == 0) {
    kmag = 1e-10;
    P = 0.0;
  } else {
    P = pow(kmag, -2.5);
  }
  const double s = sqrt(P);
  double sn, cs;
  sincos(phase[idx], &sn, &cs);
  const double dr = s * cs, di = s * sn;
  const double kc = comp == 0 ? kx : (comp == 1 ? ky : kz);
  const double t = idx == 0 ? 0.0 : kc / (kmag * kmag + 1e-10);
  // (0 + i t) (dr + i di) = (-t di, t dr)
  out[idx] = make_double2(-(t * di), t * dr);
}

// numpy's ifftn scales by 1/n per axis; the real part is the displacement.
__global__ __launch_bounds__(256) void k_real_part(const double2 *__restrict__ in,
                                                   double *__restrict__ out, int64_t n3,
                                                   double inv) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx < n3) out[idx] = in[idx].x * inv;
}

// Sum of x (SQUARES: of (x - mean[0])) over a grid-stride range into one
// partial per block, in a fixed order (deterministic for a given launch).
template <bool SQUARES>
__global__ __launch_bounds__(256) void k_partial_sum(const double *__restrict__ x, int64_t n,
                                                     const double *__restrict__ mean,
                                                     double *__restrict__ partial) {
  __shared__ double red[256];
  const double mu = SQUARES ? mean[0] : 0.0;
  double acc = 0.0;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x) {
    const double v = x[k] - mu;
    acc += SQUARES ? v * v : x[k];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// One block: total of the partials divided by n into out[0].
__global__ __launch_bounds__(256) void k_final_mean(const double *__restrict__ partial,
                                                    int nparts, int64_t n,
                                                    double *__restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) acc += partial[k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / static_cast<double>(n);
}

// numpy.remainder(x, 1.0): fmod, shifted into [0, 1) for negative results.
__device__ __forceinline__ double wrap_unit(double x) {
  double m = fmod(x, 1.0);
  if (m != 0.0 && m < 0.0) m += 1.0;
  return m == 0.0 ? 0.0 : m;
}

// cosmology.py:48-73: d *= scale with scale = (1/n) amplitude / std(d_x);
// pos = (lattice + d) % 1, vel = d * coupling, both rounded to FP32.
__global__ __launch_bounds__(256) void k_za_finish(const double *__restrict__ disp,
                                                   const double *__restrict__ var, int64_t n,
                                                   double amplitude, double coupling,
                                                   float *__restrict__ px, float *__restrict__ py,
                                                   float *__restrict__ pz, float *__restrict__ vx,
                                                   float *__restrict__ vy,
                                                   float *__restrict__ vz) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n3 = n * n * n;
  if (idx >= n3) return;
  const double scale = (1.0 / static_cast<double>(n)) * amplitude / sqrt(var[0]);
  const double step = 1.0 / static_cast<double>(n);
  const int64_t i = idx / (n * n), j = (idx / n) % n, l = idx % n;
  const double g[3] = {static_cast<double>(i) * step, static_cast<double>(j) * step,
                       static_cast<double>(l) * step};
  float *pos[3] = {px, py, pz};
  float *vel[3] = {vx, vy, vz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d = disp[c * n3 + idx] * scale;
    pos[c][idx] = static_cast<float>(wrap_unit(g[c] + d));
    vel[c][idx] = static_cast<float>(d * coupling);
  }
}

constexpr int kSumBlocks = 1024;

// One double-complex 3-D plan per (device, grid size), created on first use
// and kept until the process ends: later calls allocate nothing.  The plan
// is bound to the caller's stream on every call.
static bool cached_plan(int64_t n, hipfftHandle *out) {
  static std::mutex mu;
  static std::map<std::pair<int, int64_t>, hipfftHandle> plans;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(dev, n);
  auto it = plans.find(key);
  if (it == plans.end()) {
    hipfftHandle plan;
    if (hipfftPlan3d(&plan, static_cast<int>(n), static_cast<int>(n), static_cast<int>(n),
                     HIPFFT_Z2Z) != HIPFFT_SUCCESS)
      return false;
    it = plans.emplace(key, plan).first;
  }
  *out = it->second;
  return true;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int64_t cosmo_zeldovich_work_bytes(int64_t n_grid) {
  if (n_grid < 1 || n_grid > 1024) return 0;
  const int64_t n3 = n_grid * n_grid * n_grid;
  // complex FFT buffer, three displacement components, partial sums, mean, variance
  return n3 * int64_t(sizeof(double2)) + 3 * n3 * int64_t(sizeof(double)) +
         (kSumBlocks + 2) * int64_t(sizeof(double));
}

int cosmo_zeldovich_ics(int64_t n_grid, const double *phase, double amplitude,
                        double coupling, void *work, float *pos_x, float *pos_y, float *pos_z,
                        float *vel_x, float *vel_y, float *vel_z, void *stream) {
  if (n_grid < 1 || n_grid > 1024) return COSMO_E_INVALID;
  if (!phase || !work || !pos_x || !pos_y || !pos_z || !vel_x || !vel_y || !vel_z)
    return COSMO_E_INVALID;
  const int64_t n = n_grid, n3 = n * n * n;
  hipStream_t s = as_stream(stream);
  double2 *buf = reinterpret_cast<double2 *>(work);
  double *disp = reinterpret_cast<double *>(buf + n3);
  double *partial = disp + 3 * n3;
  double *mean = partial + kSumBlocks;
  double *var = mean + 1;
  hipfftHandle plan;
  if (!cached_plan(n, &plan)) return COSMO_E_FFT;
  int rc = COSMO_OK;
  if (hipfftSetStream(plan, s) != HIPFFT_SUCCESS) rc = COSMO_E_FFT;
  const unsigned g = blocks_for(n3, 256);
  // numpy's ifftn scales each axis by 1/n; here 1/n^3 scales the real part
  // once (the same number for a power-of-two n)
  const double inv = 1.0 / static_cast<double>(n3);
  for (int c = 0; c < 3 && rc == COSMO_OK; ++c) {
    k_za_spectrum<<<g, 256, 0, s>>>(phase, buf, n, c);
    if (hipGetLastError() != hipSuccess) { rc = COSMO_E_HIP; break; }
    hipfftDoubleComplex *z = reinterpret_cast<hipfftDoubleComplex *>(buf);
    if (hipfftExecZ2Z(plan, z, z, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) { rc = COSMO_E_FFT; break; }
    k_real_part<<<g, 256, 0, s>>>(buf, disp + c * n3, n3, inv);
    if (hipGetLastError() != hipSuccess) rc = COSMO_E_HIP;
  }
  if (rc != COSMO_OK) return rc;
  // numpy.std(d_x) (ddof 0): mean, then the mean square of the deviations
  const int parts = static_cast<int>(std::min<int64_t>(kSumBlocks, blocks_for(n3, 256)));
  k_partial_sum<false><<<parts, 256, 0, s>>>(disp, n3, nullptr, partial);
  k_final_mean<<<1, 256, 0, s>>>(partial, parts, n3, mean);
  k_partial_sum<true><<<parts, 256, 0, s>>>(disp, n3, mean, partial);
  k_final_mean<<<1, 256, 0, s>>>(partial, parts, n3, var);
  k_za_finish<<<g, 256, 0, s>>>(disp, var, n, amplitude, coupling, pos_x, pos_y, pos_z, vel_x,
                                vel_y, vel_z);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
