// Shared helpers of the gfx950 kernels: status handling, launch geometry,
// wave64 primitives and the context layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cosmo_abi.h"

namespace cosmo {

constexpr int kWave = 64;

// Records the HIP error of the most recent failing call (thread-local) so
// that cosmo_status_string() can name it.
void set_hip_error(hipError_t e);

#define COSMO_CHECK_LAUNCH()                        \
  do {                                              \
    hipError_t e_ = hipGetLastError();              \
    if (e_ != hipSuccess) {                         \
      ::cosmo::set_hip_error(e_);                   \
      return COSMO_E_HIP;                           \
    }                                               \
  } while (0)

#define COSMO_CHECK_HIP(call)                       \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) {                         \
      ::cosmo::set_hip_error(e_);                   \
      return COSMO_E_HIP;                           \
    }                                               \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(int64_t n, int threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

// Largest particle count any entry point accepts: node ids (2n-1) and the
// tree-build search range (up to 4n) stay inside int32.
constexpr int64_t kMaxBodies = int64_t(1) << 29;
// Largest leaf bucket of the AABB records (a bucket's leaves are summed in
// one pass of the traversal's leaf loop).
constexpr int kMaxBucket = 64;

// Radix sort geometry: 4096-key tiles, 8-bit digits.
constexpr int kSortTile = 4096;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

// Split traversal (small launches): at most this many top-tree subtrees per
// wavefront group (one per thread of the frontier builder), top-tree paths
// of at most kSplitMaxPath nodes, and at most kSplitMaxCount particles per
// launch.  A launch of G groups gets F = min(kSplitMaxFrontier,
// kSplitWaveBudget / G) subtrees per group, so F * G <= kSplitWaveBudget
// wavefronts (G <= kSplitMaxCount / 64 = 1024 keeps F >= 16) and the
// per-subtree partial sums fit kSplitWaveBudget * 64 * 3 floats.
constexpr int kSplitMaxFrontier = 256;
constexpr int kSplitMaxPath = 24;
constexpr int64_t kSplitMaxCount = 65536;
constexpr int64_t kSplitWaveBudget = 16384;
static_assert(kSplitWaveBudget / (kSplitMaxCount / 64) >= 1, "one subtree per group at least");
static_assert(kSplitWaveBudget >= kSplitMaxFrontier * 64, "partial buffer bound");

// Frontiers planned once per tree for a batch loop (see forces.hip): at
// most this many windows.
constexpr int64_t kSplitMaxPlans = 1024;

__host__ __device__ inline int split_fanout(int64_t count) {
  const int64_t groups = (count + 63) / 64;
  const int64_t f = kSplitWaveBudget / groups;
  return static_cast<int>(f < 1 ? 1 : (f > kSplitMaxFrontier ? kSplitMaxFrontier : f));
}

struct SplitFrontier {
  int count;
  int pad[3];
  int node[kSplitMaxFrontier];
  int slot[kSplitMaxFrontier];   // stack slot of the subtree root
  int plen[kSplitMaxFrontier];   // ancestors on the path from the root
  uint32_t lmask[kSplitMaxFrontier];  // bit k: leftmost subtree below path[k]
  int path[kSplitMaxFrontier][kSplitMaxPath];
  int pslot[kSplitMaxFrontier][kSplitMaxPath];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Mask of lanes below this one.
__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

}  // namespace cosmo

struct cosmo_ctx {
  int device;
  int64_t n_max;
  int64_t max_tiles;
  uint64_t *keys_alt;   // [n_max]
  uint32_t *vals_a;     // [n_max]
  uint32_t *vals_b;     // [n_max]
  uint32_t *hist;       // [kRadix * max_tiles]
  uint32_t *digit_tot;  // [kRadix]
  void *nodes;          // packed records, cosmo_tree_bytes(n_max, AABB)
  int32_t *scratch_i32; // [4 n_max + 4] validator counters
  void *split_frontier; // cosmo::SplitFrontier of the split traversal
  float *split_partial; // [F][count][3] <= [kSplitWaveBudget * 64][3] (see split_fanout)
  void *split_plans;    // cosmo::SplitFrontier[plan_cap]: one per window of a batch loop
  int64_t plan_cap;     // min(kSplitMaxPlans, ceil(n_max / 64))
  // what the planned frontiers were built for (plan_epoch 0: none)
  uint64_t plan_epoch;
  const int32_t *plan_root;
  int64_t plan_n, plan_win, plan_span;
  int plan_fanout;
};
