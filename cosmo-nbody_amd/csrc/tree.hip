// Radix-tree build (Karras 2012), FP64 upward pass, root/top-cache helpers,
// node packing and the tree validator.
//
// Reference: src/kernels/tree_kernels.py.  The build keeps the reference's
// delta function including its equal-code tie-break (:55-56), so `children`
// is bit-identical to the reference algorithm's for any input.  The upward
// pass climbs from each leaf with an arrival counter per node (:164-191):
// the second arrival forms the parent as left + right in FP64.  Nodes whose
// leaves all lie in one block's chunk are climbed in LDS; only nodes that
// span chunks use device memory.  The eight XCDs have separate L2 caches, so
// a child's moments must be visible to the sibling's thread on another XCD.
// Acquire/release at agent scope would do it with a write-back and an
// invalidation of the whole L2 per climb step (8 ms at 128^3); instead every
// moment store and load of the device-wide climb is a relaxed agent-scope
// atomic (coherent at the device: written through and read past the local
// L2), and each thread waits for its stores to complete (s_waitcnt vmcnt(0))
// before it bumps the arrival counter.  The second arrival reads the sibling
// only after its counter update returned, i.e. after the first arrival's
// stores completed.
#include "common.h"

namespace cosmo {

__device__ __forceinline__ int clz64(uint64_t v) {
  return v ? __builtin_clzll(v) : 64;
}

// tree_kernels.py:36-58.
__device__ __forceinline__ int delta_fn(const uint64_t *__restrict__ codes,
                                        uint64_t ci, int i, int j, int n) {
  if (j < 0 || j >= n) return -1;
  const uint64_t cj = codes[j];
  if (ci == cj) return 64 + (63 - clz64(static_cast<uint64_t>(i ^ j)));
  return clz64(ci ^ cj);
}

__global__ __launch_bounds__(256) void k_reset_tree(int32_t *__restrict__ child,
                                                    int32_t *__restrict__ parent,
                                                    int32_t *__restrict__ counters,
                                                    int32_t *__restrict__ root_idx,
                                                    int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) root_idx[0] = 0;
  if (i < m) {
    reinterpret_cast<int2 *>(child)[i] = make_int2(-1, -1);
    parent[i] = -1;
    counters[i] = 0;
  }
}

// tree_kernels.py:60-129, one thread per internal node i in [0, n-2].  far
// (nullable) receives j, the other end of node i's leaf range [min(i,j),
// max(i,j)] (no reference counterpart; the leaf buckets of the packed
// records are read from it).
__global__ __launch_bounds__(256) void k_build_tree(const uint64_t *__restrict__ codes,
                                                    int32_t *__restrict__ children,
                                                    int32_t *__restrict__ parents,
                                                    int32_t *__restrict__ far, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const uint64_t ci = codes[i];
  const int idx = i + n;
  const int d_prev = delta_fn(codes, ci, i, i - 1, n);
  const int d_next = delta_fn(codes, ci, i, i + 1, n);
  const int dir = (d_next > d_prev) ? 1 : -1;
  const int min_delta = (d_next > d_prev) ? d_prev : d_next;
  int l_max = 2;
  while (delta_fn(codes, ci, i, i + l_max * dir, n) > min_delta) l_max *= 2;
  int l = 0;
  for (int t = l_max / 2; t >= 1; t /= 2)
    if (delta_fn(codes, ci, i, i + (l + t) * dir, n) > min_delta) l += t;
  const int j = i + l * dir;
  const int delta_node = delta_fn(codes, ci, i, j, n);
  int s = 0;
  for (int t = l; t > 0;) {
    const int step = (t + 1) / 2;
    if (delta_fn(codes, ci, i, i + (s + step) * dir, n) > delta_node) s += step;
    t -= step;
  }
  int gamma = i + s * dir;
  if (dir == -1) gamma -= 1;
  const int left = gamma, right = gamma + 1;
  const int lo = min(i, j), hi = max(i, j);
  const int c0 = (lo == left) ? left : left + n;
  const int c1 = (hi == right) ? right : right + n;
  reinterpret_cast<int2 *>(children)[idx] = make_int2(c0, c1);
  parents[c0] = idx;
  parents[c1] = idx;
  if (far) far[i] = j;
}

// Device-coherent (agent-scope) relaxed accesses for data exchanged between
// threads of one kernel on different XCDs.
template <typename T>
__device__ __forceinline__ void st_dev(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_dev(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this thread's earlier memory operations have completed
// (vmcnt(0), expcnt/lgkmcnt unconstrained); compiler fences keep the
// surrounding accesses on their side.
__device__ __forceinline__ void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void store_moments(double *n_mass, double *n_com, float *n_min,
                                              float *n_max, int64_t k, double m, double c0,
                                              double c1, double c2, float lo0, float lo1,
                                              float lo2, float hi0, float hi1, float hi2) {
  st_dev(&n_mass[k], m);
  st_dev(&n_com[3 * k + 0], c0);
  st_dev(&n_com[3 * k + 1], c1);
  st_dev(&n_com[3 * k + 2], c2);
  st_dev(&n_min[3 * k + 0], lo0);
  st_dev(&n_min[3 * k + 1], lo1);
  st_dev(&n_min[3 * k + 2], lo2);
  st_dev(&n_max[3 * k + 0], hi0);
  st_dev(&n_max[3 * k + 1], hi1);
  st_dev(&n_max[3 * k + 2], hi2);
}

__device__ __forceinline__ float min_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_ref(float a, float b) { return (b > a) ? b : a; }

// Moments of one node as its climbing thread holds them.
struct Moments {
  double m, c0, c1, c2;
  float lo0, lo1, lo2, hi0, hi1, hi2;
};

// tree_kernels.py:156-162: leaf mass, mass-weighted position (FP64) and AABB.
__device__ __forceinline__ Moments leaf_moments(const float *__restrict__ px,
                                                const float *__restrict__ py,
                                                const float *__restrict__ pz,
                                                const float *__restrict__ mass, int k) {
  const float x = px[k], y = py[k], z = pz[k];
  const double m = static_cast<double>(mass[k]);
  return {m,  static_cast<double>(x) * m, static_cast<double>(y) * m, static_cast<double>(z) * m,
          x,  y, z, x, y, z};
}

// tree_kernels.py:172-187: parent = left + right.  IEEE addition is
// commutative, so "mine + sibling" equals "left + right" bit for bit; min/max
// keep the left/right order explicitly.
__device__ __forceinline__ void combine(Moments &v, const Moments &s, bool mine_left) {
  v.m = v.m + s.m;
  v.c0 = v.c0 + s.c0;
  v.c1 = v.c1 + s.c1;
  v.c2 = v.c2 + s.c2;
  if (mine_left) {
    v.lo0 = min_ref(v.lo0, s.lo0); v.lo1 = min_ref(v.lo1, s.lo1); v.lo2 = min_ref(v.lo2, s.lo2);
    v.hi0 = max_ref(v.hi0, s.hi0); v.hi1 = max_ref(v.hi1, s.hi1); v.hi2 = max_ref(v.hi2, s.hi2);
  } else {
    v.lo0 = min_ref(s.lo0, v.lo0); v.lo1 = min_ref(s.lo1, v.lo1); v.lo2 = min_ref(s.lo2, v.lo2);
    v.hi0 = max_ref(s.hi0, v.hi0); v.hi1 = max_ref(s.hi1, v.hi1); v.hi2 = max_ref(s.hi2, v.hi2);
  }
}

__device__ __forceinline__ void store_plain(double *n_mass, double *n_com, float *n_min,
                                            float *n_max, int64_t k, const Moments &v) {
  n_mass[k] = v.m;
  n_com[3 * k + 0] = v.c0;
  n_com[3 * k + 1] = v.c1;
  n_com[3 * k + 2] = v.c2;
  n_min[3 * k + 0] = v.lo0;
  n_min[3 * k + 1] = v.lo1;
  n_min[3 * k + 2] = v.lo2;
  n_max[3 * k + 0] = v.hi0;
  n_max[3 * k + 1] = v.hi1;
  n_max[3 * k + 2] = v.hi2;
}

// tree_kernels.py:164-191 across the device: the climb of a node whose
// moments are complete in registers, through nodes that no single block
// owns.  Every moment store and load here is device-coherent (see the file
// header) and each thread's stores complete before it bumps the counter.
__device__ void climb_device(const int32_t *__restrict__ children,
                             const int32_t *__restrict__ parents, double *n_mass,
                             double *n_com, float *n_min, float *n_max, int32_t *counts,
                             Moments v, int curr, int n) {
  store_moments(n_mass, n_com, n_min, n_max, curr, v.m, v.c0, v.c1, v.c2, v.lo0, v.lo1, v.lo2,
                v.hi0, v.hi1, v.hi2);
  // A climb is at most n - 1 levels; the bound only matters for a malformed
  // tree (the validator reports those), where it keeps every thread finite.
  for (int level = 0; level < n; ++level) {
    const int p = parents[curr];
    if (p == -1) break;
    wait_stores();
    const int old = __hip_atomic_fetch_add(&counts[p], 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    // the sibling's moments may only be read after the counter update has
    // returned: no load below may be moved above it by the compiler (the
    // hardware issues them after the branch on `old`, which waits for the
    // returned value)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (old == 0) break;
    const int2 ch = reinterpret_cast<const int2 *>(children)[p];
    const bool mine_left = (ch.x == curr);
    const int64_t sib = mine_left ? ch.y : ch.x;
    Moments s;
    s.m = ld_dev(&n_mass[sib]);
    s.c0 = ld_dev(&n_com[3 * sib + 0]);
    s.c1 = ld_dev(&n_com[3 * sib + 1]);
    s.c2 = ld_dev(&n_com[3 * sib + 2]);
    s.lo0 = ld_dev(&n_min[3 * sib + 0]);
    s.lo1 = ld_dev(&n_min[3 * sib + 1]);
    s.lo2 = ld_dev(&n_min[3 * sib + 2]);
    s.hi0 = ld_dev(&n_max[3 * sib + 0]);
    s.hi1 = ld_dev(&n_max[3 * sib + 1]);
    s.hi2 = ld_dev(&n_max[3 * sib + 2]);
    combine(v, s, mine_left);
    store_moments(n_mass, n_com, n_min, n_max, p, v.m, v.c0, v.c1, v.c2, v.lo0, v.lo1, v.lo2,
                  v.hi0, v.hi1, v.hi2);
    curr = p;
  }
}

// tree_kernels.py:131-191, blocked.  Block b owns the leaf chunk
// [b0, b0 + K) with b0 = b K.  An internal node lies inside the chunk
// exactly when both of its children do, and its index (one end of its leaf
// range) is then in [b0, b0 + K), so those nodes get an LDS arrival counter
// and an LDS moment slot each.  Phase 1 climbs from every leaf through the
// chunk's own nodes with LDS atomics.  A climb stops where the parent's
// index or the sibling lies outside the chunk; a first arrival whose
// sibling never came by the barrier (the sibling's range leaves the chunk)
// stops there too.  Those threads hold a complete node whose parent spans
// several chunks, and phase 2 continues them with the device-wide climb on
// `counts`, which sees only such nodes.  The chunk's own internal nodes are
// written out after the barrier with coalesced stores, and their counters
// are set to 2, the value the per-node climb leaves on every internal node.
// Sums are formed in the same order as the per-node climb (bit-identical).
constexpr int kUpChunk = 1024;

__global__ __launch_bounds__(kUpChunk) void k_multipoles(
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pz, const float *__restrict__ mass,
    const int32_t *__restrict__ children, const int32_t *__restrict__ parents,
    double *n_mass, double *n_com, float *n_min, float *n_max, int32_t *counts,
    int n) {
  __shared__ double s_m[kUpChunk], s_c0[kUpChunk], s_c1[kUpChunk], s_c2[kUpChunk];
  __shared__ float s_lo0[kUpChunk], s_lo1[kUpChunk], s_lo2[kUpChunk];
  __shared__ float s_hi0[kUpChunk], s_hi1[kUpChunk], s_hi2[kUpChunk];
  __shared__ int s_cnt[kUpChunk];
  // the chunk's topology, loaded once with coalesced reads: parents of its
  // leaves and of its internal nodes, children of its internal nodes
  __shared__ int s_par_leaf[kUpChunk], s_par_int[kUpChunk];
  __shared__ int2 s_ch[kUpChunk];
  const int t = threadIdx.x;
  const int b0 = blockIdx.x * kUpChunk;
  const int leaf = b0 + t;
  s_cnt[t] = 0;
  s_par_leaf[t] = leaf < n ? parents[leaf] : -1;
  s_par_int[t] = leaf < n - 1 ? parents[n + leaf] : -1;
  s_ch[t] = leaf < n - 1 ? reinterpret_cast<const int2 *>(children)[n + leaf]
                         : make_int2(-1, -1);
  __syncthreads();
  // 0: climb ended (root, or completed by the sibling's thread);
  // 1: first arrival at chunk node `wait`; 2: continue across chunks.
  int state = 0, wait = 0, curr = leaf;
  Moments v{};
  if (leaf < n) {
    v = leaf_moments(px, py, pz, mass, leaf);
    store_plain(n_mass, n_com, n_min, n_max, leaf, v);
    for (int level = 0; level < kUpChunk; ++level) {
      // curr is a leaf or an internal node of the chunk
      const int p = curr < n ? s_par_leaf[curr - b0] : s_par_int[curr - n - b0];
      if (p == -1) break;
      const int slot = p - n - b0;
      if (slot < 0 || slot >= kUpChunk) { state = 2; break; }
      const int2 ch = s_ch[slot];
      const bool mine_left = (ch.x == curr);
      const int sib = mine_left ? ch.y : ch.x;
      const int sslot = (sib >= n ? sib - n : sib) - b0;
      if (sslot < 0 || sslot >= kUpChunk) { state = 2; break; }
      const int old = __hip_atomic_fetch_add(&s_cnt[slot], 1, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) { state = 1; wait = slot; break; }
      Moments s;
      if (sib >= n) {
        s = {s_m[sslot],   s_c0[sslot],  s_c1[sslot],  s_c2[sslot],  s_lo0[sslot],
             s_lo1[sslot], s_lo2[sslot], s_hi0[sslot], s_hi1[sslot], s_hi2[sslot]};
      } else {
        s = leaf_moments(px, py, pz, mass, sib);
      }
      combine(v, s, mine_left);
      s_m[slot] = v.m; s_c0[slot] = v.c0; s_c1[slot] = v.c1; s_c2[slot] = v.c2;
      s_lo0[slot] = v.lo0; s_lo1[slot] = v.lo1; s_lo2[slot] = v.lo2;
      s_hi0[slot] = v.hi0; s_hi1[slot] = v.hi1; s_hi2[slot] = v.hi2;
      curr = p;
    }
  }
  __syncthreads();
  if (state == 1 && s_cnt[wait] == 1) state = 2;
  const int i = b0 + t;
  if (i < n - 1 && s_cnt[t] == 2) {
    const Moments w{s_m[t],   s_c0[t],  s_c1[t],  s_c2[t],  s_lo0[t],
                    s_lo1[t], s_lo2[t], s_hi0[t], s_hi1[t], s_hi2[t]};
    store_plain(n_mass, n_com, n_min, n_max, int64_t(n) + i, w);
    counts[n + i] = 2;
  }
  if (state == 2) climb_device(children, parents, n_mass, n_com, n_min, n_max, counts, v, curr, n);
}

// tree_kernels.py:193-214.
__global__ __launch_bounds__(256) void k_normalize_com(const double *__restrict__ n_mass,
                                                       double *__restrict__ n_com,
                                                       int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double nm = n_mass[i];
  if (nm > 0) {
    const double inv = 1.0 / nm;
    n_com[3 * i + 0] *= inv;
    n_com[3 * i + 1] *= inv;
    n_com[3 * i + 2] *= inv;
  }
}

// tree_kernels.py:193-214 on the reference's float32 node arrays
// (octree.py:37-38), in the reference's arithmetic: inv = 1.0 / m is FP64
// (a Python float literal over a float32 element) and com *= inv rounds the
// FP64 product back to the float32 array.
__global__ __launch_bounds__(256) void k_normalize_com_f32(const float *__restrict__ n_mass,
                                                           float *__restrict__ n_com,
                                                           int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double nm = static_cast<double>(n_mass[i]);
  if (nm > 0) {
    const double inv = 1.0 / nm;
    n_com[3 * i + 0] = static_cast<float>(static_cast<double>(n_com[3 * i + 0]) * inv);
    n_com[3 * i + 1] = static_cast<float>(static_cast<double>(n_com[3 * i + 1]) * inv);
    n_com[3 * i + 2] = static_cast<float>(static_cast<double>(n_com[3 * i + 2]) * inv);
  }
}

// FP64 node mass and CoM sums rounded to the reference's float32 arrays.
__global__ __launch_bounds__(256) void k_round_moments(const double *__restrict__ m64,
                                                       const double *__restrict__ c64,
                                                       float *__restrict__ n_mass,
                                                       float *__restrict__ n_com, int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  n_mass[i] = static_cast<float>(m64[i]);
  n_com[3 * i + 0] = static_cast<float>(c64[3 * i + 0]);
  n_com[3 * i + 1] = static_cast<float>(c64[3 * i + 1]);
  n_com[3 * i + 2] = static_cast<float>(c64[3 * i + 2]);
}

// tree_kernels.py:216-234: the largest parentless index wins (deterministic).
__global__ __launch_bounds__(256) void k_find_root(const int32_t *__restrict__ parents,
                                                   int32_t *__restrict__ root_idx, int64_t m) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < m && parents[i] == -1) atomicMax(root_idx, static_cast<int32_t>(i));
}

// tree_kernels.py:264-314, restated serially on one lane (256 slots).
__global__ void k_top_tree_cache(const int32_t *__restrict__ child,
                                 const int32_t *__restrict__ root_idx,
                                 int32_t *__restrict__ top_nodes,
                                 int32_t *__restrict__ node_to_cache, int cache_size) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int root = root_idx[0];
  if (root == -1) return;
  top_nodes[0] = root;
  node_to_cache[root] = 0;
  int read_idx = 0, write_idx = 1;
  while (read_idx < write_idx && write_idx < cache_size) {
    const int curr = top_nodes[read_idx];
    const int c1 = child[2 * curr], c2 = child[2 * curr + 1];
    if (c1 != -1 && write_idx < cache_size) {
      top_nodes[write_idx] = c1;
      node_to_cache[c1] = write_idx;
      write_idx++;
    }
    if (c2 != -1 && write_idx < cache_size) {
      top_nodes[write_idx] = c2;
      node_to_cache[c2] = write_idx;
      write_idx++;
    }
    read_idx++;
  }
  while (write_idx < cache_size) top_nodes[write_idx++] = -1;
}

// Writes node k's FP32 traversal record (leaf or internal, see cosmo_abi.h)
// from its FP64 mass, normalised FP64 CoM and FP32 box.
template <int MAC>
__device__ __forceinline__ void write_record(void *__restrict__ out, int64_t n, int64_t k,
                                             double nm, double c0, double c1, double c2,
                                             float lo0, float lo1, float lo2, float hi0,
                                             float hi1, float hi2, int2 ch, float boxf,
                                             int first, int count) {
  const float fx = static_cast<float>(c0), fy = static_cast<float>(c1), fz = static_cast<float>(c2);
  const float fm = static_cast<float>(nm);
  if (MAC == COSMO_MAC_REF) {
    // One 32-byte record per node id.  physics_kernels.py:129-130 for an
    // internal node; a leaf has sx2 = -1 (always accepted).  A leaf's
    // children are the massless sentinel record 2n - 1 that follows the
    // nodes: a leaf whose test fails (only on a non-finite distance) and is
    // opened pushes two sentinels, which the walk pops and skips, so the
    // loop needs no leaf check before a push.  A node lacking a child
    // (c1 == -1 is the reference's leaf, physics_kernels.py:127; a missing
    // c2 only occurs in a malformed tree) is written as a leaf.
    const int sentinel = static_cast<int>(2 * n - 1);
    float4 *dst = reinterpret_cast<float4 *>(out) + 2 * k;
    dst[0] = make_float4(fx, fy, fz, fm);
    if (k < n || ch.x < 0 || ch.y < 0) {
      dst[1] = make_float4(-1.0f, 0.0f, __int_as_float(sentinel), __int_as_float(sentinel));
    } else {
      float sx = hi0 - lo0;
      sx = (sx > boxf * 0.5f) ? boxf : sx;
      // |sx^2|: the sign bit marks leaves, also for a NaN extent
      dst[1] = make_float4(fabsf(sx * sx), sx, __int_as_float(ch.x), __int_as_float(ch.y));
    }
    if (k == 2 * n - 2) {  // the sentinel: massless, itself as children
      dst[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      dst[3] = make_float4(-1.0f, 0.0f, __int_as_float(sentinel), __int_as_float(sentinel));
    }
    return;
  }
  if (k < n) {
    reinterpret_cast<float4 *>(out)[k] = make_float4(fx, fy, fz, fm);
    return;
  }
  char *ibase = reinterpret_cast<char *>(out) + n * sizeof(cosmo_leaf);
  {
    const float e0 = hi0 - lo0, e1 = hi1 - lo1, e2 = hi2 - lo2;
    float sz = 0.0f;
    sz = e0 > sz ? e0 : sz;
    sz = e1 > sz ? e1 : sz;
    sz = e2 > sz ? e2 : sz;
    const float o0 = fx - 0.5f * (hi0 + lo0);
    const float o1 = fy - 0.5f * (hi1 + lo1);
    const float o2 = fz - 0.5f * (hi2 + lo2);
    const float dl = sqrtf(__builtin_fmaf(o2, o2, __builtin_fmaf(o1, o1, o0 * o0)));
    float4 *dst = reinterpret_cast<float4 *>(ibase + (k - n) * sizeof(cosmo_inode_aabb));
    dst[0] = make_float4(fx, fy, fz, fm);
    dst[1] = make_float4(sz, dl, __int_as_float(ch.x), __int_as_float(ch.y));
    dst[2] = make_float4(fx - lo0, fy - lo1, fz - lo2, hi0 - fx);
    dst[3] = make_float4(hi1 - fy, hi2 - fz, __int_as_float(first), __int_as_float(count));
  }
}

// tree_kernels.py:193-214 on one node: c *= 1/m for m > 0.
__device__ __forceinline__ void normalize_one(double nm, double &c0, double &c1, double &c2) {
  if (nm > 0) {
    const double inv = 1.0 / nm;
    c0 *= inv;
    c1 *= inv;
    c2 *= inv;
  }
}

// Packs node k into its FP32 record; NORMALIZE also divides the FP64 CoM
// sums by the mass first (normalize_com fused) and writes them back.  With
// far (the other end of each internal node's leaf range, from the build)
// an AABB record of a node with at most `bucket` leaves also holds its leaf
// range (first, count); count 0 marks every other node.
// T is double (this build's FP64 node arrays) or float (the reference's
// float32 arrays, widened exactly); NORMALIZE needs double.
template <int MAC, bool NORMALIZE, typename T = double>
__global__ __launch_bounds__(256) void k_pack_nodes(
    const int32_t *__restrict__ children, const T *__restrict__ n_mass,
    T *__restrict__ n_com, const float *__restrict__ n_min,
    const float *__restrict__ n_max, int64_t m, float boxf, void *__restrict__ out,
    const int32_t *__restrict__ far, int bucket) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int64_t n = (m + 1) / 2;
  const double nm = n_mass[k];
  double c0 = n_com[3 * k + 0], c1 = n_com[3 * k + 1], c2 = n_com[3 * k + 2];
  if (NORMALIZE && nm > 0) {
    normalize_one(nm, c0, c1, c2);
    n_com[3 * k + 0] = c0;
    n_com[3 * k + 1] = c1;
    n_com[3 * k + 2] = c2;
  }
  const int2 ch = k < n ? make_int2(-1, -1) : reinterpret_cast<const int2 *>(children)[k];
  int first = 0, count = 0;
  if (far && k >= n) {
    const int i = static_cast<int>(k - n), j = far[k - n];
    const int c = (i < j ? j - i : i - j) + 1;
    if (c <= bucket) {
      first = i < j ? i : j;
      count = c;
    }
  }
  write_record<MAC>(out, n, k, nm, c0, c1, c2, n_min[3 * k + 0], n_min[3 * k + 1],
                    n_min[3 * k + 2], n_max[3 * k + 0], n_max[3 * k + 1], n_max[3 * k + 2], ch,
                    boxf, first, count);
}

// ------------------------------------------------------------ validator

// scratch layout: [0] flag bits, [1] parentless count, [2] largest parentless
// id, [3..3+M) arrival counters, [3+M..3+2M) leaves below each node.
__global__ __launch_bounds__(256) void k_validate_clear(int32_t *scratch, int64_t len) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) scratch[i] = (i == 2) ? -1 : 0;
}

__global__ __launch_bounds__(256) void k_validate_links(const int32_t *__restrict__ children,
                                                        const int32_t *__restrict__ parents,
                                                        int n, int32_t *scratch) {
  const int64_t M = 2 * int64_t(n) - 1;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= M) return;
  int bad = 0;
  const int2 ch = reinterpret_cast<const int2 *>(children)[k];
  if (k < n) {
    if (ch.x != -1 || ch.y != -1) bad = 4;
  } else {
    if (ch.x < 0 || ch.x >= M || ch.y < 0 || ch.y >= M || ch.x == ch.y) bad = 4;
    else if (parents[ch.x] != k || parents[ch.y] != k) bad = 4;
  }
  if (bad) atomicOr(&scratch[0], bad);
  if (parents[k] == -1) {
    atomicAdd(&scratch[1], 1);
    atomicMax(&scratch[2], static_cast<int32_t>(k));
  }
}

// Counts leaves below every node with the same arrival-counter climb as the
// upward pass; each node completes at most once, so the climb terminates
// on any parents array.
__global__ __launch_bounds__(256) void k_validate_climb(const int32_t *__restrict__ children,
                                                        const int32_t *__restrict__ parents,
                                                        int n, int32_t *scratch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t M = 2 * int64_t(n) - 1;
  int32_t *arrive = scratch + 3;
  int32_t *leaves = scratch + 3 + M;
  int cnt = 1;
  leaves[i] = 1;
  int curr = i;
  for (;;) {
    const int p = parents[curr];
    if (p < 0 || p >= M) break;
    const int old = __hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) break;
    const int2 ch = reinterpret_cast<const int2 *>(children)[p];
    const int sib = (ch.x == curr) ? ch.y : ch.x;
    if (sib < 0 || sib >= M) break;
    cnt += leaves[sib];
    leaves[p] = cnt;
    curr = p;
  }
}

__global__ void k_validate_finish(int n, const int32_t *scratch, int32_t *flags_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t M = 2 * int64_t(n) - 1;
  int f = scratch[0];
  const int n_root = scratch[1], root = scratch[2];
  if (n_root != 1) f |= 1;
  if (n >= 2 && !(n_root == 1 && root == n)) f |= 2;
  if (n_root < 1 || root < 0 || root >= M || scratch[3 + M + root] != n) f |= 8;
  flags_out[0] = f;
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_reset_tree_buffers(int32_t *child, int32_t *parent, int32_t *counters,
                             int32_t *root_idx, int64_t n_nodes, void *stream) {
  if (n_nodes < 0 || n_nodes > 2 * kMaxBodies) return COSMO_E_INVALID;
  if (!child || !parent || !counters || !root_idx) return COSMO_E_INVALID;
  const unsigned g = blocks_for(n_nodes > 0 ? n_nodes : 1, 256);
  k_reset_tree<<<g, 256, 0, as_stream(stream)>>>(child, parent, counters, root_idx, n_nodes);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

static int build_tree_impl(const uint64_t *codes, int32_t *children, int32_t *parents,
                           int32_t *far, int64_t n, void *stream) {
  if (n < 0 || n > kMaxBodies) return COSMO_E_INVALID;
  if (n < 2) return COSMO_OK;
  if (!codes || !children || !parents) return COSMO_E_INVALID;
  k_build_tree<<<blocks_for(n - 1, 256), 256, 0, as_stream(stream)>>>(codes, children, parents,
                                                                       far, static_cast<int>(n));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_build_tree(const uint64_t *codes, int32_t *children, int32_t *parents, int64_t n,
                     void *stream) {
  return build_tree_impl(codes, children, parents, nullptr, n, stream);
}

int cosmo_build_tree_spans(const uint64_t *codes, int32_t *children, int32_t *parents,
                           int32_t *far, int64_t n, void *stream) {
  if (n >= 2 && !far) return COSMO_E_INVALID;
  return build_tree_impl(codes, children, parents, far, n, stream);
}

int cosmo_compute_multipoles(const float *px, const float *py, const float *pz,
                             const float *mass, const int32_t *children,
                             const int32_t *parents, double *n_mass, double *n_com,
                             float *n_min, float *n_max, int32_t *counts, int64_t n,
                             void *stream) {
  if (n < 0 || n > kMaxBodies) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (!px || !py || !pz || !mass || !children || !parents || !n_mass || !n_com || !n_min ||
      !n_max || !counts)
    return COSMO_E_INVALID;
  k_multipoles<<<blocks_for(n, kUpChunk), kUpChunk, 0, as_stream(stream)>>>(
      px, py, pz, mass, children, parents, n_mass, n_com, n_min, n_max, counts,
      static_cast<int>(n));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_compute_multipoles_f32(cosmo_ctx *ctx, const float *px, const float *py,
                                 const float *pz, const float *mass, const int32_t *children,
                                 const int32_t *parents, float *n_mass, float *n_com,
                                 float *n_min, float *n_max, int32_t *counts, int64_t n,
                                 void *stream) {
  if (!ctx || n < 0 || n > kMaxBodies) return COSMO_E_INVALID;
  if (n == 0) return COSMO_OK;
  if (n > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (!n_mass || !n_com) return COSMO_E_INVALID;
  // FP64 sums in the context's record buffer (80 n_max >= 32 (2n - 1) bytes)
  const int64_t m = 2 * n - 1;
  double *m64 = reinterpret_cast<double *>(ctx->nodes);
  double *c64 = m64 + m;
  const int rc = cosmo_compute_multipoles(px, py, pz, mass, children, parents, m64, c64, n_min,
                                          n_max, counts, n, stream);
  if (rc != COSMO_OK) return rc;
  k_round_moments<<<blocks_for(m, 256), 256, 0, as_stream(stream)>>>(m64, c64, n_mass, n_com, m);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_normalize_com_f32(const float *n_mass, float *n_com, int64_t n_nodes, void *stream) {
  if (n_nodes < 0 || n_nodes > 2 * kMaxBodies) return COSMO_E_INVALID;
  if (n_nodes == 0) return COSMO_OK;
  if (!n_mass || !n_com) return COSMO_E_INVALID;
  k_normalize_com_f32<<<blocks_for(n_nodes, 256), 256, 0, as_stream(stream)>>>(n_mass, n_com,
                                                                               n_nodes);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_normalize_com(const double *n_mass, double *n_com, int64_t n_nodes, void *stream) {
  if (n_nodes < 0 || n_nodes > 2 * kMaxBodies) return COSMO_E_INVALID;
  if (n_nodes == 0) return COSMO_OK;
  if (!n_mass || !n_com) return COSMO_E_INVALID;
  k_normalize_com<<<blocks_for(n_nodes, 256), 256, 0, as_stream(stream)>>>(n_mass, n_com,
                                                                           n_nodes);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_find_root(const int32_t *parents, int32_t *root_idx, int64_t n_nodes, void *stream) {
  if (n_nodes < 0 || n_nodes > 2 * kMaxBodies) return COSMO_E_INVALID;
  if (n_nodes == 0) return COSMO_OK;
  if (!parents || !root_idx) return COSMO_E_INVALID;
  k_find_root<<<blocks_for(n_nodes, 256), 256, 0, as_stream(stream)>>>(parents, root_idx,
                                                                       n_nodes);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_build_top_tree_cache(const int32_t *child, const int32_t *root_idx,
                               int32_t *top_nodes, int32_t *node_to_cache,
                               int64_t cache_size, void *stream) {
  if (cache_size < 0 || cache_size > (1 << 20)) return COSMO_E_INVALID;
  if (cache_size == 0) return COSMO_OK;
  if (!child || !root_idx || !top_nodes || !node_to_cache) return COSMO_E_INVALID;
  k_top_tree_cache<<<1, 64, 0, as_stream(stream)>>>(child, root_idx, top_nodes, node_to_cache,
                                                    static_cast<int>(cache_size));
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

static int pack_impl(const int32_t *children, const double *n_mass, double *n_com,
                     const float *n_min, const float *n_max, int64_t m, double box, int mac,
                     bool normalize, const int32_t *far, int bucket, void *nodes, void *stream) {
  if (m < 0 || m > 2 * kMaxBodies || (m > 0 && (m & 1) == 0)) return COSMO_E_INVALID;
  if (bucket < 0 || bucket > kMaxBucket || (bucket > 0 && (mac != COSMO_MAC_AABB || !far)))
    return COSMO_E_INVALID;
  if (m == 0) return COSMO_OK;
  if (!children || !n_mass || !n_com || !n_min || !n_max || !nodes) return COSMO_E_INVALID;
  const unsigned g = blocks_for(m, 256);
  hipStream_t s = as_stream(stream);
  const float boxf = static_cast<float>(box);
  if (bucket == 0) far = nullptr;
  if (mac == COSMO_MAC_REF && normalize)
    k_pack_nodes<COSMO_MAC_REF, true><<<g, 256, 0, s>>>(children, n_mass, n_com, n_min, n_max, m, boxf, nodes, far, bucket);
  else if (mac == COSMO_MAC_REF)
    k_pack_nodes<COSMO_MAC_REF, false><<<g, 256, 0, s>>>(children, n_mass, n_com, n_min, n_max, m, boxf, nodes, far, bucket);
  else if (mac == COSMO_MAC_AABB && normalize)
    k_pack_nodes<COSMO_MAC_AABB, true><<<g, 256, 0, s>>>(children, n_mass, n_com, n_min, n_max, m, boxf, nodes, far, bucket);
  else if (mac == COSMO_MAC_AABB)
    k_pack_nodes<COSMO_MAC_AABB, false><<<g, 256, 0, s>>>(children, n_mass, n_com, n_min, n_max, m, boxf, nodes, far, bucket);
  else
    return COSMO_E_INVALID;
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_pack_nodes_f32(const int32_t *children, const float *n_mass, const float *n_com,
                         const float *n_min, const float *n_max, int64_t n_nodes, double box,
                         int mac, void *nodes, void *stream) {
  const int64_t m = n_nodes;
  if (m < 0 || m > 2 * kMaxBodies || (m > 0 && (m & 1) == 0)) return COSMO_E_INVALID;
  if (mac != COSMO_MAC_REF && mac != COSMO_MAC_AABB) return COSMO_E_INVALID;
  if (m == 0) return COSMO_OK;
  if (!children || !n_mass || !n_com || !n_min || !n_max || !nodes) return COSMO_E_INVALID;
  const unsigned g = blocks_for(m, 256);
  hipStream_t s = as_stream(stream);
  const float boxf = static_cast<float>(box);
  float *com = const_cast<float *>(n_com);  // not written without NORMALIZE
  if (mac == COSMO_MAC_REF)
    k_pack_nodes<COSMO_MAC_REF, false, float><<<g, 256, 0, s>>>(children, n_mass, com, n_min, n_max, m, boxf, nodes, nullptr, 0);
  else
    k_pack_nodes<COSMO_MAC_AABB, false, float><<<g, 256, 0, s>>>(children, n_mass, com, n_min, n_max, m, boxf, nodes, nullptr, 0);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_pack_nodes(const int32_t *children, const double *n_mass, const double *n_com,
                     const float *n_min, const float *n_max, int64_t n_nodes, double box,
                     int mac, void *nodes, void *stream) {
  return pack_impl(children, n_mass, const_cast<double *>(n_com), n_min, n_max, n_nodes, box,
                   mac, false, nullptr, 0, nodes, stream);
}

int cosmo_normalize_pack(const int32_t *children, const double *n_mass, double *n_com,
                         const float *n_min, const float *n_max, int64_t n_nodes, double box,
                         int mac, void *nodes, void *stream) {
  return pack_impl(children, n_mass, n_com, n_min, n_max, n_nodes, box, mac, true, nullptr, 0,
                   nodes, stream);
}

int cosmo_normalize_pack_buckets(const int32_t *children, const int32_t *far,
                                 const double *n_mass, double *n_com, const float *n_min,
                                 const float *n_max, int64_t n_nodes, double box, int mac,
                                 int bucket, void *nodes, void *stream) {
  return pack_impl(children, n_mass, n_com, n_min, n_max, n_nodes, box, mac, true, far, bucket,
                   nodes, stream);
}

int cosmo_validate_tree(cosmo_ctx *ctx, const int32_t *children, const int32_t *parents,
                        int64_t n, int32_t *flags_out, void *stream) {
  if (!ctx || n < 1 || !children || !parents || !flags_out) return COSMO_E_INVALID;
  if (n > ctx->n_max) return COSMO_E_TOO_LARGE;
  const int64_t M = 2 * n - 1;
  const int64_t len = 3 + 2 * M;
  hipStream_t s = as_stream(stream);
  k_validate_clear<<<blocks_for(len, 256), 256, 0, s>>>(ctx->scratch_i32, len);
  k_validate_links<<<blocks_for(M, 256), 256, 0, s>>>(children, parents, static_cast<int>(n),
                                                      ctx->scratch_i32);
  k_validate_climb<<<blocks_for(n, 256), 256, 0, s>>>(children, parents, static_cast<int>(n),
                                                      ctx->scratch_i32);
  k_validate_finish<<<1, 64, 0, s>>>(static_cast<int>(n), ctx->scratch_i32, flags_out);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
