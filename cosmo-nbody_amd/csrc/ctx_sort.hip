// Context, status strings, Morton encoding and the (code, index) radix sort.
//
// Reference: src/kernels/morton.py (compute_codes) and src/core/state.py:20-40
// (cp.argsort of the codes followed by eight gathers).  Here the sort is a
// stable LSD radix sort with 8-bit digits whose passes each read a tile of
// 4096 keys per workgroup: a digit count, a per-digit scan over tiles, and a
// scatter whose within-tile ranks come from wave64 match masks (8 ballots),
// so equal digits keep their input order and the result is the permutation
// of a sort on (code, index).  The scatter reorders each tile by digit in
// LDS before it writes, so its global writes are contiguous runs.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "common.h"

namespace cosmo {

static thread_local hipError_t g_last_hip = hipSuccess;
void set_hip_error(hipError_t e) { g_last_hip = e; }

// ------------------------------------------------------------------ Morton

// morton.py:33-36 as written: keeps the low 10 bits of v.
__device__ __forceinline__ uint64_t spread_ref10(uint64_t v) {
  v = (v | (v << 16)) & 0x030000FFull;
  v = (v | (v << 8)) & 0x0300F00Full;
  v = (v | (v << 4)) & 0x030C30C3ull;
  v = (v | (v << 2)) & 0x09249249ull;
  return v;
}

// 21 bits per axis: bit k of v moves to bit 3k.
__device__ __forceinline__ uint64_t spread_full21(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// morton.py:57-60 uint64(x * 2^21); x*2^21 is exact in FP32.  Not-positive
// values (and NaN) give 0, values >= 2^40 saturate.
__device__ __forceinline__ uint64_t fixed21(float x) {
  const float s = x * 2097152.0f;
  if (!(s > 0.0f)) return 0;
  if (s >= 1099511627776.0f) return 1099511627776ull;
  return static_cast<uint64_t>(s);
}

template <int MASKS>
__global__ __launch_bounds__(256) void k_compute_codes(
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pz, uint64_t *__restrict__ codes,
    uint32_t *__restrict__ indices, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t ix = fixed21(px[i]), iy = fixed21(py[i]), iz = fixed21(pz[i]);
  uint64_t c;
  if (MASKS == COSMO_MASK_REF10)
    c = spread_ref10(ix) | (spread_ref10(iy) << 1) | (spread_ref10(iz) << 2);
  else
    c = spread_full21(ix) | (spread_full21(iy) << 1) | (spread_full21(iz) << 2);
  codes[i] = c;
  if (indices) indices[i] = static_cast<uint32_t>(i);
}

// ------------------------------------------------------------- radix sort

// Lanes of this wave holding the same 8-bit digit (among `valid` lanes).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Exclusive scan of one value per thread over an NT-thread block.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wave_sums[NT / kWave];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(x, off, kWave);
    if (lane >= off) x += t;
  }
  if (lane == kWave - 1) wave_sums[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) {
    const uint32_t s = wave_sums[w];
    before += (w < wave) ? s : 0u;
    all += s;
  }
  __syncthreads();
  if (total) *total = all;
  return before + x - v;
}

// Digit counts of one tile: plain LDS atomics (the order of equal digits
// does not matter for a count), one histogram per wave to spread conflicts.
// A launch has one block per 4096-key tile (512 at 128^3, two per CU), so
// the tile is spread over kCountThreads threads (kCountItems keys each) to
// keep enough wavefronts in flight; the totals are formed by the first
// kRadix threads.
constexpr int kCountThreads = 1024;
constexpr int kCountItems = kSortTile / kCountThreads;
static_assert(kCountItems * kCountThreads == kSortTile, "tile split");

__global__ __launch_bounds__(kCountThreads) void k_radix_count(
    const uint64_t *__restrict__ keys, int64_t n, int shift,
    uint32_t *__restrict__ hist, int ntiles) {
  constexpr int kWaves = kCountThreads / kWave;
  __shared__ uint32_t cnt[kWaves][kRadix];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  for (int i = tid; i < kWaves * kRadix; i += kCountThreads) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kSortTile + wave * (kCountItems * kWave);
  uint64_t k[kCountItems];
#pragma unroll
  for (int r = 0; r < kCountItems; ++r) {
    const int64_t idx = base + r * kWave + lane;
    k[r] = idx < n ? keys[idx] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < kCountItems; ++r) {
    const int64_t idx = base + r * kWave + lane;
    if (idx < n) atomicAdd(&cnt[wave][uint32_t(k[r] >> shift) & (kRadix - 1)], 1u);
  }
  __syncthreads();
  if (tid < kRadix) {
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) total += cnt[w][tid];
    hist[int64_t(tid) * ntiles + blockIdx.x] = total;
  }
}

// One block per digit: exclusive scan of that digit's per-tile counts, in
// runs of kRowscanThreads tiles.
constexpr int kRowscanThreads = 1024;

__global__ __launch_bounds__(kRowscanThreads) void k_radix_rowscan(
    uint32_t *__restrict__ hist, int ntiles, uint32_t *__restrict__ digit_tot) {
  uint32_t *row = hist + int64_t(blockIdx.x) * ntiles;
  uint32_t running = 0;
  for (int start = 0; start < ntiles; start += kRowscanThreads) {
    const int j = start + threadIdx.x;
    const uint32_t v = (j < ntiles) ? row[j] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<kRowscanThreads>(v, &total);
    if (j < ntiles) row[j] = running + ex;
    running += total;
  }
  if (threadIdx.x == 0) digit_tot[blockIdx.x] = running;
}

// Stable scatter of one tile, staged through LDS: every key gets its rank
// among the tile's keys of the same digit (wave match masks, then waves in
// order), the tile is reordered by digit in LDS, and the write-out walks
// the reordered tile so that consecutive threads write consecutive
// addresses of one digit's output run.  The tile's LDS (48 KiB of staging
// and 16 KiB of per-wavefront digit counts) allows two blocks per CU, so the
// tile is spread over kScatterThreads threads (kScatterItems keys each) to
// keep the CU's 32 wavefront slots busy; the per-digit bookkeeping runs on
// the first kRadix threads.  Measured (same permutation): with 256 threads
// of 16 keys the sort took 0.264 ms at 128^3 and 2.23 ms at 256^3, with 512
// threads 0.243 and 1.94 ms, with 1024 threads 0.247 and 1.78 ms.
constexpr int kScatterThreads = 1024;
constexpr int kScatterItems = kSortTile / kScatterThreads;
static_assert(kScatterItems * kScatterThreads == kSortTile, "tile split");
static_assert(kScatterThreads >= kRadix, "one thread per digit");

__global__ __launch_bounds__(kScatterThreads) void k_radix_scatter(
    const uint64_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    uint64_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, int64_t n,
    int shift, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ digit_tot, int ntiles) {
  constexpr int kWaves = kScatterThreads / kWave;
  __shared__ uint32_t wcnt[kWaves][kRadix];
  __shared__ uint32_t lstart[kRadix];  // first tile-local slot of each digit
  __shared__ uint32_t gofs[kRadix];    // output position - tile-local slot
  __shared__ uint64_t sk[kSortTile];
  __shared__ uint32_t sv[kSortTile];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const bool digit_thread = tid < kRadix;
  for (int i = tid; i < kWaves * kRadix; i += kScatterThreads) (&wcnt[0][0])[i] = 0;
  const uint32_t dstart =
      block_exclusive_scan<kScatterThreads>(digit_thread ? digit_tot[tid] : 0u, nullptr);
  const uint32_t gbase = digit_thread ? dstart + hist[int64_t(tid) * ntiles + blockIdx.x] : 0u;
  __syncthreads();

  const int64_t tile0 = int64_t(blockIdx.x) * kSortTile;
  const int64_t base = tile0 + wave * (kScatterItems * kWave);
  const uint64_t lt = lanemask_lt();
  uint64_t k[kScatterItems];
  uint32_t v[kScatterItems], rank[kScatterItems];
#pragma unroll
  for (int r = 0; r < kScatterItems; ++r) {
    const int64_t idx = base + r * kWave + lane;
    const bool valid = idx < n;
    k[r] = valid ? keys_in[idx] : ~0ull;
    v[r] = valid ? (vals_in ? vals_in[idx] : static_cast<uint32_t>(idx)) : 0u;
    const uint32_t d = uint32_t(k[r] >> shift) & (kRadix - 1);
    const uint64_t m = match_digit(d, valid);
    const uint32_t prior = wcnt[wave][d];
    rank[r] = prior + (uint32_t)__popcll(m & lt);
    if (valid && lane == __ffsll((unsigned long long)m) - 1)
      wcnt[wave][d] = prior + (uint32_t)__popcll(m);
  }
  __syncthreads();
  uint32_t run = 0;
  if (digit_thread) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t t = wcnt[w][tid];
      wcnt[w][tid] = run;
      run += t;
    }
  }
  const uint32_t ls = block_exclusive_scan<kScatterThreads>(run, nullptr);
  if (digit_thread) {
    lstart[tid] = ls;
    gofs[tid] = gbase - ls;  // modular: gofs + slot is the output position
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScatterItems; ++r) {
    const int64_t idx = base + r * kWave + lane;
    if (idx < n) {
      const uint32_t d = uint32_t(k[r] >> shift) & (kRadix - 1);
      const uint32_t slot = lstart[d] + wcnt[wave][d] + rank[r];
      sk[slot] = k[r];
      sv[slot] = v[r];
    }
  }
  __syncthreads();
  const int64_t valid_in_tile = n - tile0 < kSortTile ? n - tile0 : kSortTile;
#pragma unroll
  for (int r = 0; r < kScatterItems; ++r) {
    const int j = r * kScatterThreads + tid;
    if (j < valid_in_tile) {
      const uint64_t key = sk[j];
      const uint32_t pos = gofs[uint32_t(key >> shift) & (kRadix - 1)] + uint32_t(j);
      keys_out[pos] = key;
      vals_out[pos] = sv[j];
    }
  }
}

struct PermuteArgs {
  const float *src[8];
  float *dst[8];
};

template <int NA>
__global__ __launch_bounds__(256) void k_permute(const uint32_t *__restrict__ perm,
                                                 PermuteArgs a, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
#pragma unroll
  for (int k = 0; k < NA; ++k) a.dst[k][i] = a.src[k][j];
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

const char *cosmo_status_string(int status) {
  switch (status) {
    case COSMO_OK:
      return "ok";
    case COSMO_E_INVALID:
      return "invalid argument";
    case COSMO_E_HIP: {
      static thread_local char buf[160];
      snprintf(buf, sizeof buf, "HIP error: %s", hipGetErrorString(g_last_hip));
      return buf;
    }
    case COSMO_E_TOO_LARGE:
      return "size exceeds the capacity (the context's n_max, or COSMO_SORT_NOALLOC_MAX "
             "for the no-allocation sort)";
    case COSMO_E_MALFORMED_TREE:
      return "malformed tree";
    case COSMO_E_FFT:
      return "hipFFT error";
    default:
      return "unknown status";
  }
}

int cosmo_ctx_create(int64_t n_max, cosmo_ctx **out) {
  if (!out || n_max < 1 || n_max > kMaxBodies) return COSMO_E_INVALID;
  *out = nullptr;
  cosmo_ctx *c = new cosmo_ctx();
  memset(c, 0, sizeof *c);
  COSMO_CHECK_HIP(hipGetDevice(&c->device));
  c->n_max = n_max;
  c->max_tiles = (n_max + kSortTile - 1) / kSortTile;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&c->keys_alt, sizeof(uint64_t) * n_max);
  if (e == hipSuccess) e = hipMalloc(&c->vals_a, sizeof(uint32_t) * n_max);
  if (e == hipSuccess) e = hipMalloc(&c->vals_b, sizeof(uint32_t) * n_max);
  if (e == hipSuccess) e = hipMalloc(&c->hist, sizeof(uint32_t) * kRadix * c->max_tiles);
  if (e == hipSuccess) e = hipMalloc(&c->digit_tot, sizeof(uint32_t) * kRadix);
  if (e == hipSuccess) e = hipMalloc(&c->nodes, cosmo_tree_bytes(n_max, COSMO_MAC_AABB));
  if (e == hipSuccess) e = hipMalloc(&c->scratch_i32, sizeof(int32_t) * (4 * n_max + 4));
  if (e == hipSuccess) e = hipMalloc(&c->split_frontier, sizeof(SplitFrontier));
  c->plan_cap = std::min<int64_t>(kSplitMaxPlans, (n_max + 63) / 64);
  if (e == hipSuccess) e = hipMalloc(&c->split_plans, sizeof(SplitFrontier) * c->plan_cap);
  if (e == hipSuccess)
    e = hipMalloc(&c->split_partial,
                  sizeof(float) * 3 *
                      std::min<int64_t>(kSplitWaveBudget * 64,
                                        kSplitMaxFrontier * std::min<int64_t>(n_max, kSplitMaxCount)));
  if (e != hipSuccess) {
    set_hip_error(e);
    cosmo_ctx_destroy(c);
    return COSMO_E_HIP;
  }
  *out = c;
  return COSMO_OK;
}

int cosmo_ctx_destroy(cosmo_ctx *c) {
  if (!c) return COSMO_OK;
  (void)hipFree(c->split_frontier);
  (void)hipFree(c->split_partial);
  (void)hipFree(c->split_plans);
  (void)hipFree(c->keys_alt);
  (void)hipFree(c->vals_a);
  (void)hipFree(c->vals_b);
  (void)hipFree(c->hist);
  (void)hipFree(c->digit_tot);
  (void)hipFree(c->nodes);
  (void)hipFree(c->scratch_i32);
  delete c;
  return COSMO_OK;
}

void *cosmo_ctx_nodes(cosmo_ctx *c) { return c ? c->nodes : nullptr; }

int64_t cosmo_tree_bytes(int64_t n, int mac) {
  if (n < 1) return 0;
  // REF: one record per node id and the sentinel record 2n - 1
  if (mac != COSMO_MAC_AABB) return 2 * n * int64_t(sizeof(cosmo_inode_ref));
  // AABB: the traversal reads a whole internal-record size at every node,
  // leaves included, so the buffer carries one internal record of tail room
  return n * int64_t(sizeof(cosmo_leaf)) + n * int64_t(sizeof(cosmo_inode_aabb));
}

int cosmo_compute_codes(const float *px, const float *py, const float *pz,
                        uint64_t *codes, uint32_t *indices, int64_t n,
                        int mask_set, void *stream) {
  if (n < 0 || n > kMaxBodies) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !codes) return COSMO_E_INVALID;
  const unsigned g = blocks_for(n, 256);
  if (mask_set == COSMO_MASK_REF10)
    k_compute_codes<COSMO_MASK_REF10><<<g, 256, 0, as_stream(stream)>>>(px, py, pz, codes, indices, n);
  else if (mask_set == COSMO_MASK_FULL21)
    k_compute_codes<COSMO_MASK_FULL21><<<g, 256, 0, as_stream(stream)>>>(px, py, pz, codes, indices, n);
  else
    return COSMO_E_INVALID;
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

// The LSD radix sort on a given workspace: keys_alt and vals_alt hold n
// entries, hist kRadix * ceil(n / kSortTile), digit_tot kRadix.  Every pass
// overwrites what it reads later, so the workspace needs no clearing.
static int sort_with(uint64_t *keys_alt, uint32_t *vals_alt, uint32_t *hist, uint32_t *digit_tot,
                     const uint64_t *codes_in, uint64_t *codes_out, uint32_t *perm, int64_t n,
                     int key_bits, void *stream) {
  hipStream_t s = as_stream(stream);
  const int ntiles = static_cast<int>((n + kSortTile - 1) / kSortTile);
  int passes = (key_bits + kRadixBits - 1) / kRadixBits;
  // Pass p writes to (codes_out, perm) when (passes-1-p) is even, so the
  // last pass lands there.  With codes_in == codes_out the first pass must
  // not write codes_out: keep the pass count even (an extra pass over a
  // digit that is zero everywhere is a stable copy).
  if (codes_in == codes_out && (passes & 1) && passes < 8) passes++;
  const uint64_t *kin = codes_in;
  const uint32_t *vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    const bool to_out = ((passes - 1 - p) & 1) == 0;
    uint64_t *kout = to_out ? codes_out : keys_alt;
    uint32_t *vout = to_out ? perm : vals_alt;
    const int shift = p * kRadixBits;
    k_radix_count<<<ntiles, kCountThreads, 0, s>>>(kin, n, shift, hist, ntiles);
    k_radix_rowscan<<<kRadix, kRowscanThreads, 0, s>>>(hist, ntiles, digit_tot);
    k_radix_scatter<<<ntiles, kScatterThreads, 0, s>>>(kin, vin, kout, vout, n, shift, hist,
                                                       digit_tot, ntiles);
    COSMO_CHECK_LAUNCH();
    kin = kout;
    vin = vout;
  }
  return COSMO_OK;
}

// ------------------------------------------- sort without allocation
//
// Workspace of cosmo_sort_by_code_noalloc in one file-scope device array:
// alternate keys and values for COSMO_SORT_NOALLOC_MAX entries, the tile
// histograms and the digit totals (24.5 MiB per device, held until the
// process ends).  Host code never passes `scratch` itself: its device
// address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kNoAllocMax = COSMO_SORT_NOALLOC_MAX;
constexpr int64_t kNoAllocTiles = (kNoAllocMax + kSortTile - 1) / kSortTile;
constexpr size_t kNoAllocKeysOfs = 0;
constexpr size_t kNoAllocValsOfs = kNoAllocKeysOfs + sizeof(uint64_t) * kNoAllocMax;
constexpr size_t kNoAllocHistOfs = kNoAllocValsOfs + sizeof(uint32_t) * kNoAllocMax;
constexpr size_t kNoAllocTotOfs = kNoAllocHistOfs + sizeof(uint32_t) * kRadix * kNoAllocTiles;
constexpr size_t kNoAllocBytes = kNoAllocTotOfs + sizeof(uint32_t) * kRadix;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device and kept).
static int scratch_on_device(unsigned char **out) {
  int dev = 0;
  COSMO_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return COSMO_E_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    COSMO_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return COSMO_E_HIP;
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return COSMO_OK;
}

int cosmo_sort_by_code_noalloc(const uint64_t *codes_in, uint64_t *codes_out, uint32_t *perm,
                               int64_t n, int key_bits, void *stream) {
  if (n < 0 || key_bits < 1 || key_bits > 64) return COSMO_E_INVALID;
  if (n > kNoAllocMax) return COSMO_E_TOO_LARGE;
  if (n == 0) return COSMO_OK;
  if (!codes_in || !codes_out || !perm) return COSMO_E_INVALID;
  unsigned char *base = nullptr;
  const int rc = scratch_on_device(&base);
  if (rc != COSMO_OK) return rc;
  return sort_with(reinterpret_cast<uint64_t *>(base + kNoAllocKeysOfs),
                   reinterpret_cast<uint32_t *>(base + kNoAllocValsOfs),
                   reinterpret_cast<uint32_t *>(base + kNoAllocHistOfs),
                   reinterpret_cast<uint32_t *>(base + kNoAllocTotOfs), codes_in, codes_out, perm,
                   n, key_bits, stream);
}

int cosmo_sort_by_code(cosmo_ctx *ctx, const uint64_t *codes_in, uint64_t *codes_out,
                       uint32_t *perm, int64_t n, int key_bits, void *stream) {
  if (!ctx || n < 0 || key_bits < 1 || key_bits > 64) return COSMO_E_INVALID;
  if (n > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (n == 0) return COSMO_OK;
  if (!codes_in || !codes_out || !perm) return COSMO_E_INVALID;
  return sort_with(ctx->keys_alt, ctx->vals_a, ctx->hist, ctx->digit_tot, codes_in, codes_out,
                   perm, n, key_bits, stream);
}

int cosmo_permute(const uint32_t *perm, int n_arrays, const float *const *src,
                  float *const *dst, int64_t n, void *stream) {
  if (n < 0 || n_arrays < 0 || n_arrays > 8 || (n_arrays && (!src || !dst)))
    return COSMO_E_INVALID;
  if (n == 0 || n_arrays == 0) return COSMO_OK;
  if (!perm) return COSMO_E_INVALID;
  PermuteArgs a;
  memset(&a, 0, sizeof a);
  for (int k = 0; k < n_arrays; ++k) {
    if (!src[k] || !dst[k]) return COSMO_E_INVALID;
    a.src[k] = src[k];
    a.dst[k] = dst[k];
  }
  const unsigned g = blocks_for(n, 256);
  hipStream_t s = as_stream(stream);
  switch (n_arrays) {
    case 1: k_permute<1><<<g, 256, 0, s>>>(perm, a, n); break;
    case 2: k_permute<2><<<g, 256, 0, s>>>(perm, a, n); break;
    case 3: k_permute<3><<<g, 256, 0, s>>>(perm, a, n); break;
    case 4: k_permute<4><<<g, 256, 0, s>>>(perm, a, n); break;
    case 5: k_permute<5><<<g, 256, 0, s>>>(perm, a, n); break;
    case 6: k_permute<6><<<g, 256, 0, s>>>(perm, a, n); break;
    case 7: k_permute<7><<<g, 256, 0, s>>>(perm, a, n); break;
    default: k_permute<8><<<g, 256, 0, s>>>(perm, a, n); break;
  }
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
