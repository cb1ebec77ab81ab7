// Barnes-Hut traversal and the FP64 direct sum.
//
// Reference: src/kernels/physics_kernels.py:14-153 (compute_forces), one
// thread per particle with a private 64-entry stack.  Here one wavefront
// walks the tree for 64 Morton-consecutive particles with ONE shared stack:
// each entry carries the 64-bit mask of the lanes that must visit it.  On a
// pop every masked lane runs its own acceptance test, accepted lanes
// interact, and the lanes that open the node become the mask of the pushed
// children (c2 then c1, so c1 is visited first, physics_kernels.py:144-149).
// Each lane therefore visits exactly its own nodes in its own depth-first
// order, so its interaction set and FP32 summation order are those of the
// per-particle walk.  A per-lane depth counter equals the depth of the
// reference's private stack and reproduces its overflow drop
// (`stack_top + 1 < 64`, :143).  Masks of stacked entries are nested (an
// entry's lanes opened every node below it), so the shared stack never
// holds more than 64 entries and lives in three VGPRs indexed by lane
// (v_readlane with a wave-uniform slot; a push selects one lane).  Node records are
// wave-uniform loads (scalar cache).
#include "common.h"

namespace cosmo {

struct TravParams {
  double th2d;      // REF: theta^2 in FP64 (the reference's arithmetic)
  float th2;        // REF: theta^2
  float inv_theta;  // AABB: 1/theta
  float G;
  float eps2;
  float half_box;
};

// Unit roundoff of FP32.
constexpr float kU = 5.9604645e-8f;

// Minimum image of r = fl(c - p) when |r| == 0.5.  Only then can the FP32
// difference and the exact one fall on different sides of 1/2 (rounding is
// monotonic and 1/2 is representable), so only then is the image decided
// from the exact difference: c - p > 1/2  <=>  c - 1/2 > p, where c - 1/2 is
// exact for c in [1/4, 1] (Sterbenz), which holds whenever c - p ~ 1/2 with
// p >= 0.  An exact tie keeps r, as rint(+-0.5) = 0 does.
__device__ __forceinline__ float wrap_tie(float c, float p, float r) {
  const float s = (c - 0.5f > p) ? 1.0f : ((p - 0.5f > c) ? -1.0f : 0.0f);
  return r - s;
}

// Keeps the compiler from hoisting work on v out of a rarely taken branch
// (an empty asm that claims to modify v; emits no instruction).
__device__ __forceinline__ float pin(float v) {
  asm volatile("" : "+v"(v));
  return v;
}

__device__ __forceinline__ void neumaier_add(float &s, float &c, float x) {
  const float t = s + x;
  if (fabsf(s) >= fabsf(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

template <int MAC>
struct InodeOf;
template <>
struct InodeOf<COSMO_MAC_REF> {
  typedef cosmo_inode_ref type;
};
template <>
struct InodeOf<COSMO_MAC_AABB> {
  typedef cosmo_inode_aabb type;
};

// Minimum-image displacement from a particle to a record's CoM (rint image,
// which gives the right d^2 in every case; see fix_ties for the direction).
struct Disp {
  float rx, ry, rz, dx, dy, dz, d2;
};

__device__ __forceinline__ Disp displace(float cx, float cy, float cz, float x, float y,
                                         float z) {
  Disp d;
  d.rx = cx - x;
  d.ry = cy - y;
  d.rz = cz - z;
  d.dx = d.rx - rintf(d.rx);
  d.dy = d.ry - rintf(d.ry);
  d.dz = d.rz - rintf(d.rz);
  d.d2 = __builtin_fmaf(d.dz, d.dz, __builtin_fmaf(d.dy, d.dy, d.dx * d.dx));
  return d;
}

// FP64 restatement of the reference's displacement (physics_kernels.py:118-126).
__device__ __forceinline__ double exact_d2(float cx, float cy, float cz, float x, float y,
                                          float z) {
  double ex = static_cast<double>(pin(cx)) - static_cast<double>(x);
  double ey = static_cast<double>(pin(cy)) - static_cast<double>(y);
  double ez = static_cast<double>(pin(cz)) - static_cast<double>(z);
  ex -= rint(ex);
  ey -= rint(ey);
  ez -= rint(ez);
  return ex * ex + ey * ey + ez * ez;
}

// A component with |c - p| rounding to exactly 1/2 may belong to the other
// image: d^2 is the same either way, the direction is not.  Such components
// only occur for d^2 >= 1/4.
__device__ __forceinline__ void fix_ties(Disp &d, float cx, float cy, float cz, float x, float y,
                                         float z) {
  if (__builtin_expect(d.d2 >= 0.25f, 0)) {
    d.dx = (fabsf(d.rx) == 0.5f) ? wrap_tie(pin(cx), x, d.rx) : d.dx;
    d.dy = (fabsf(d.ry) == 0.5f) ? wrap_tie(pin(cy), y, d.ry) : d.dy;
    d.dz = (fabsf(d.rz) == 0.5f) ? wrap_tie(pin(cz), z, d.rz) : d.dz;
  }
}

template <bool COMP>
__device__ __forceinline__ void accumulate(const Disp &d, float gm, float eps2, float &ax,
                                           float &ay, float &az, float &cx, float &cy,
                                           float &cz) {
  const float inv = __builtin_amdgcn_rsqf(d.d2 + eps2);
  const float f = gm * ((inv * inv) * inv);
  if (COMP) {
    neumaier_add(ax, cx, f * d.dx);
    neumaier_add(ay, cy, f * d.dy);
    neumaier_add(az, cz, f * d.dz);
  } else {
    ax = __builtin_fmaf(f, d.dx, ax);
    ay = __builtin_fmaf(f, d.dy, ay);
    az = __builtin_fmaf(f, d.dz, az);
  }
}

// PPL particles per lane: a wavefront walks for 64*PPL Morton-consecutive
// particles (lane l holds particles base + 64h + l, h < PPL).  Each stack
// entry carries one 64-bit lane mask per half; the masks of each half are
// nested along the stack, so the depth bound of 64 still holds.  With PPL = 2
// one scalar node load serves twice the particles, at the price of a larger
// union of visited nodes.  Leaves (node ids < n_tree) take a separate path:
// they are always accepted and never push.
template <int MAC, bool COMP, bool STATS, bool GROUP_THETA, int PPL>
__global__ __launch_bounds__(256) void k_traverse(
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pz, const void *__restrict__ tree, int n_tree,
    const int32_t *__restrict__ root_idx, float *__restrict__ force, TravParams p,
    const float *__restrict__ theta_group, int64_t lo, int64_t hi,
    unsigned long long *__restrict__ stats) {
  typedef typename InodeOf<MAC>::type Inode;
  const float4 *__restrict__ leaves = reinterpret_cast<const float4 *>(tree);
  const Inode *__restrict__ inodes = reinterpret_cast<const Inode *>(
      reinterpret_cast<const char *>(tree) + int64_t(n_tree) * sizeof(cosmo_leaf));
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  const int64_t base = lo + wave * (kWave * PPL);
  if (base >= hi) return;  // wave-uniform
  float x[PPL], y[PPL], z[PPL], th2[PPL], inv_theta[PPL];
  double th2d[PPL];
  float ax[PPL], ay[PPL], az[PPL], cx[PPL], cy[PPL], cz[PPL];
  int depth[PPL];
  bool active[PPL];
  uint32_t n_int = 0, n_vis = 0, n_pop = 0;
  int stk_node = 0;
  uint32_t stk_lo[PPL], stk_hi[PPL];
  const int root = root_idx[0];
#pragma unroll
  for (int h = 0; h < PPL; ++h) {
    const int64_t i = base + h * kWave + lane;
    active[h] = i < hi;
    x[h] = y[h] = z[h] = 0.0f;
    if (active[h]) {
      x[h] = px[i];
      y[h] = py[i];
      z[h] = pz[i];
    }
    th2[h] = p.th2;
    inv_theta[h] = p.inv_theta;
    th2d[h] = p.th2d;
    if (GROUP_THETA) {
      const int64_t g = wave * PPL + h;
      const float t = (base + h * kWave < hi) ? theta_group[g] : 0.5f;
      th2[h] = t * t;
      inv_theta[h] = 1.0f / t;
      th2d[h] = static_cast<double>(t) * static_cast<double>(t);
    }
    ax[h] = ay[h] = az[h] = cx[h] = cy[h] = cz[h] = 0.0f;
    depth[h] = active[h] ? 1 : 0;
    const uint64_t am = __ballot(active[h]);
    stk_lo[h] = (lane == 0) ? static_cast<uint32_t>(am) : 0u;
    stk_hi[h] = (lane == 0) ? static_cast<uint32_t>(am >> 32) : 0u;
  }
  if (lane == 0) stk_node = root;
  int top = 1;

  while (top > 0) {
    --top;
    if (STATS) n_pop++;
    const int node = __builtin_amdgcn_readlane(stk_node, top);
    bool in[PPL];
#pragma unroll
    for (int h = 0; h < PPL; ++h) {
      const uint32_t mlo = __builtin_amdgcn_readlane(stk_lo[h], top);
      const uint32_t mhi = __builtin_amdgcn_readlane(stk_hi[h], top);
      in[h] = ((lane < 32 ? mlo >> lane : mhi >> (lane - 32)) & 1u) != 0;
      depth[h] -= in[h] ? 1 : 0;
    }
    if (node < n_tree) {
      // leaf: accepted unconditionally (physics_kernels.py:128,135)
      const float4 L = leaves[node];
      if (L.w <= 0.0f) continue;
      const float gm = p.G * L.w;
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        if (!in[h]) continue;
        Disp d = displace(L.x, L.y, L.z, x[h], y[h], z[h]);
        bool near_ok = d.d2 > 1e-7f;
        // the reference's d^2 > 1e-7 cut in FP64 where FP32 rounding (at most
        // 2e-10 near the cut) could decide otherwise
        if (MAC == COSMO_MAC_REF && __builtin_expect(fabsf(d.d2 - 1e-7f) <= 2e-10f, 0))
          near_ok = exact_d2(L.x, L.y, L.z, x[h], y[h], z[h]) > 1e-7;
        if (near_ok) {
          fix_ties(d, L.x, L.y, L.z, x[h], y[h], z[h]);
          accumulate<COMP>(d, gm, p.eps2, ax[h], ay[h], az[h], cx[h], cy[h], cz[h]);
          if (STATS) n_int++;
        }
        if (STATS) n_vis++;
      }
      continue;
    }
    const Inode rec = inodes[node - n_tree];
    if (rec.m <= 0.0f) continue;
    const float gm = p.G * rec.m;
    bool open[PPL];
#pragma unroll
    for (int h = 0; h < PPL; ++h) {
      open[h] = false;
      if (!in[h]) continue;
      Disp d = displace(rec.cx, rec.cy, rec.cz, x[h], y[h], z[h]);
      bool accept, near_ok = d.d2 > 1e-7f;
      if constexpr (MAC == COSMO_MAC_REF) {
        // The reference decides in FP64 (physics_kernels.py:122-136).  The
        // FP32 decision is final unless rounding could flip it: with
        // |d_k - d_exact,k| <= u |c_k - p_k| <= u and a few u on the products,
        // |error of sx^2 - theta^2 d^2| <= theta^2 u (10.1 d^2 + 3 |d|_1)
        // <= 11.8 u theta^2 |d|_1 (|d| <= sqrt(3)/2), and |d^2 - d_exact^2|
        // <= 2e-10 near the 1e-7 cut.  Lanes inside those bands redo both
        // tests in FP64 exactly as the reference does.
        const float sx = rec.sx;
        const float t = th2[h] * d.d2;
        const float sx2 = sx * sx;
        accept = sx2 < t;
        const float l1 = fabsf(d.dx) + fabsf(d.dy) + fabsf(d.dz);
        const bool unsure = (fabsf(d.d2 - 1e-7f) <= 2e-10f) ||
                            (fabsf(sx2 - t) <= (16.0f * kU) * th2[h] * l1);
        if (__builtin_expect(unsure, 0)) {
          const double e2 = exact_d2(rec.cx, rec.cy, rec.cz, x[h], y[h], z[h]);
          const double sxd = static_cast<double>(pin(sx));
          accept = (sxd * sxd) < (th2d[h] * e2);
          near_ok = e2 > 1e-7;
        }
      } else {
        // (s/theta + dl)^2 < d^2, and every particle of the node sits on the
        // same side of the minimum-image boundary as its CoM in each axis
        const cosmo_inode_aabb &ra = rec;
        const float r = __builtin_fmaf(ra.s, inv_theta[h], ra.dl);
        const float hb = p.half_box;
        const bool inside = (d.dx + ra.hi[0] < hb) && (d.dx - ra.lo[0] > -hb) &&
                            (d.dy + ra.hi[1] < hb) && (d.dy - ra.lo[1] > -hb) &&
                            (d.dz + ra.hi[2] < hb) && (d.dz - ra.lo[2] > -hb);
        accept = (r * r < d.d2) && inside;
      }
      if (accept) {
        if (near_ok) {
          fix_ties(d, rec.cx, rec.cy, rec.cz, x[h], y[h], z[h]);
          accumulate<COMP>(d, gm, p.eps2, ax[h], ay[h], az[h], cx[h], cy[h], cz[h]);
          if (STATS) n_int++;
        }
      } else {
        open[h] = depth[h] + 1 < 64;
      }
      if (STATS) n_vis++;
    }
    uint32_t olo[PPL], ohi[PPL];
    bool any = false;
#pragma unroll
    for (int h = 0; h < PPL; ++h) {
      const uint64_t om = __ballot(open[h]);
      olo[h] = static_cast<uint32_t>(om);
      ohi[h] = static_cast<uint32_t>(om >> 32);
      any = any || (om != 0);
    }
    if (any) {
      // internal nodes have two children (build_tree): push c2, then c1
#pragma unroll
      for (int h = 0; h < PPL; ++h)
        if (open[h]) depth[h] += 2;
      bool here = (lane == top);
      stk_node = here ? rec.c2 : stk_node;
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        stk_lo[h] = here ? olo[h] : stk_lo[h];
        stk_hi[h] = here ? ohi[h] : stk_hi[h];
      }
      ++top;
      here = (lane == top);
      stk_node = here ? rec.c1 : stk_node;
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        stk_lo[h] = here ? olo[h] : stk_lo[h];
        stk_hi[h] = here ? ohi[h] : stk_hi[h];
      }
      ++top;
    }
  }
#pragma unroll
  for (int h = 0; h < PPL; ++h) {
    if (active[h]) {
      const int64_t i = base + h * kWave + lane;
      force[3 * i + 0] = ax[h] + cx[h];
      force[3 * i + 1] = ay[h] + cy[h];
      force[3 * i + 2] = az[h] + cz[h];
    }
  }
  if (STATS) {
    uint32_t a = n_int, b = n_vis;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off, kWave);
      b += __shfl_xor(b, off, kWave);
    }
    if (lane == 0) {
      atomicAdd(&stats[0], static_cast<unsigned long long>(a));
      atomicAdd(&stats[1], static_cast<unsigned long long>(b));
      atomicAdd(&stats[2], static_cast<unsigned long long>(n_pop));
    }
  }
}

// FP64 direct sum: a block of 1024 threads takes 64 targets (one per lane);
// its 16 waves split the sources and their partial sums are combined in a
// fixed order, so results do not depend on scheduling.
constexpr int kDirectWaves = 16;

__global__ __launch_bounds__(1024) void k_direct_sum(
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pz, const float *__restrict__ mass, int64_t n,
    const int64_t *__restrict__ targets, int64_t nt, double G, double eps2,
    double *__restrict__ acc) {
  __shared__ double part[kDirectWaves][kWave][3];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int64_t t = int64_t(blockIdx.x) * kWave + lane;
  const bool valid = t < nt;
  const int64_t ti = valid ? targets[t] : 0;
  const double x = px[ti], y = py[ti], z = pz[ti];
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int64_t chunk = (n + kDirectWaves - 1) / kDirectWaves;
  const int64_t j0 = wave * chunk;
  const int64_t j1 = (j0 + chunk < n) ? j0 + chunk : n;
  for (int64_t j = j0; j < j1; ++j) {
    double dx = static_cast<double>(px[j]) - x;
    double dy = static_cast<double>(py[j]) - y;
    double dz = static_cast<double>(pz[j]) - z;
    dx -= rint(dx);
    dy -= rint(dy);
    dz -= rint(dz);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > 1e-7) {
      const double d_inv = 1.0 / sqrt(d2 + eps2);
      const double f = G * static_cast<double>(mass[j]) * (d_inv * d_inv * d_inv);
      ax += f * dx;
      ay += f * dy;
      az += f * dz;
    }
  }
  part[wave][lane][0] = ax;
  part[wave][lane][1] = ay;
  part[wave][lane][2] = az;
  __syncthreads();
  if (wave == 0 && valid) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int w = 0; w < kDirectWaves; ++w) {
      sx += part[w][lane][0];
      sy += part[w][lane][1];
      sz += part[w][lane][2];
    }
    acc[3 * t + 0] = sx;
    acc[3 * t + 1] = sy;
    acc[3 * t + 2] = sz;
  }
}

template <int MAC, bool COMP, bool STATS, bool GT, int PPL>
static void launch_traverse(hipStream_t s, const float *px, const float *py, const float *pz,
                            const void *nodes, int n_tree, const int32_t *root_idx, float *force,
                            TravParams p, const float *tg, int64_t lo, int64_t hi,
                            uint64_t *stats) {
  const int64_t waves = (hi - lo + kWave * PPL - 1) / (kWave * PPL);
  const unsigned grid = blocks_for(waves * kWave, 256);
  k_traverse<MAC, COMP, STATS, GT, PPL><<<grid, 256, 0, s>>>(
      px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi,
      reinterpret_cast<unsigned long long *>(stats));
}

template <int MAC, bool COMP, bool STATS, bool GT>
static void dispatch_ppl(bool pairs, hipStream_t s, const float *px, const float *py,
                         const float *pz, const void *nodes, int n_tree, const int32_t *root_idx,
                         float *force, TravParams p, const float *tg, int64_t lo, int64_t hi,
                         uint64_t *stats) {
  if (pairs)
    launch_traverse<MAC, COMP, STATS, GT, 2>(s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
  else
    launch_traverse<MAC, COMP, STATS, GT, 1>(s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
}

template <int MAC, bool COMP>
static void dispatch_traverse2(bool stats_on, bool gt, bool pairs, hipStream_t s,
                               const float *px, const float *py, const float *pz,
                               const void *nodes, int n_tree, const int32_t *root_idx,
                               float *force, TravParams p, const float *tg, int64_t lo, int64_t hi,
                               uint64_t *stats) {
  if (stats_on) {
    if (gt)
      dispatch_ppl<MAC, COMP, true, true>(pairs, s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
    else
      dispatch_ppl<MAC, COMP, true, false>(pairs, s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
  } else {
    if (gt)
      dispatch_ppl<MAC, COMP, false, true>(pairs, s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
    else
      dispatch_ppl<MAC, COMP, false, false>(pairs, s, px, py, pz, nodes, n_tree, root_idx, force, p, tg, lo, hi, stats);
  }
}

}  // namespace cosmo

using namespace cosmo;

extern "C" {

int cosmo_traverse(const float *px, const float *py, const float *pz, const void *nodes,
                   int64_t n_tree, const int32_t *root_idx, float *force, double theta,
                   const float *theta_group, double G, double soft, double box, int64_t off,
                   int64_t count, int64_t n_bodies, int mac, int flags, uint64_t *stats,
                   void *stream) {
  if (n_bodies < 0 || n_bodies > kMaxBodies || off < 0 || count < 0 || n_tree < 1 ||
      n_tree > kMaxBodies)
    return COSMO_E_INVALID;
  if (mac != COSMO_MAC_REF && mac != COSMO_MAC_AABB) return COSMO_E_INVALID;
  const bool stats_on = (flags & COSMO_TRAV_STATS) != 0;
  if (stats_on && !stats) return COSMO_E_INVALID;
  const int64_t lo = off;
  const int64_t hi = (off + count < n_bodies) ? off + count : n_bodies;
  if (hi <= lo) return COSMO_OK;
  if (!px || !py || !pz || !nodes || !root_idx || !force) return COSMO_E_INVALID;
  TravParams p;
  const float thf = static_cast<float>(theta);
  p.th2d = theta * theta;
  p.th2 = thf * thf;
  p.inv_theta = 1.0f / thf;
  p.G = static_cast<float>(G);
  const float sf = static_cast<float>(soft);
  p.eps2 = sf * sf;
  p.half_box = static_cast<float>(box) * 0.5f;
  hipStream_t s = as_stream(stream);
  const bool comp = (flags & COSMO_TRAV_COMPENSATED) != 0;
  const bool pairs = (flags & COSMO_TRAV_PAIRS) != 0;
  const bool gt = theta_group != nullptr;
  if (mac == COSMO_MAC_REF) {
    if (comp)
      dispatch_traverse2<COSMO_MAC_REF, true>(stats_on, gt, pairs, s, px, py, pz, nodes, static_cast<int>(n_tree), root_idx, force, p, theta_group, lo, hi, stats);
    else
      dispatch_traverse2<COSMO_MAC_REF, false>(stats_on, gt, pairs, s, px, py, pz, nodes, static_cast<int>(n_tree), root_idx, force, p, theta_group, lo, hi, stats);
  } else {
    if (comp)
      dispatch_traverse2<COSMO_MAC_AABB, true>(stats_on, gt, pairs, s, px, py, pz, nodes, static_cast<int>(n_tree), root_idx, force, p, theta_group, lo, hi, stats);
    else
      dispatch_traverse2<COSMO_MAC_AABB, false>(stats_on, gt, pairs, s, px, py, pz, nodes, static_cast<int>(n_tree), root_idx, force, p, theta_group, lo, hi, stats);
  }
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_compute_forces(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                         const int32_t *children, const double *n_mass, const double *n_com,
                         const float *n_min, const float *n_max, float *force, double theta,
                         double G, double soft, double box, const int32_t *root_idx,
                         const int32_t *top_nodes, const int32_t *node_to_cache, int64_t off,
                         int64_t count, int64_t n_bodies, void *stream) {
  (void)top_nodes;
  (void)node_to_cache;
  if (!ctx) return COSMO_E_INVALID;
  if (n_bodies > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (n_bodies < 1) return COSMO_OK;
  const int rc = cosmo_pack_nodes(children, n_mass, n_com, n_min, n_max, 2 * n_bodies - 1, box,
                                  COSMO_MAC_REF, ctx->nodes, stream);
  if (rc != COSMO_OK) return rc;
  return cosmo_traverse(px, py, pz, ctx->nodes, n_bodies, root_idx, force, theta, nullptr, G,
                        soft, box, off, count, n_bodies, COSMO_MAC_REF, 0, nullptr, stream);
}

int cosmo_direct_sum(const float *px, const float *py, const float *pz, const float *mass,
                     int64_t n, const int64_t *targets, int64_t n_targets, double G,
                     double soft, double *acc, void *stream) {
  if (n < 0 || n_targets < 0) return COSMO_E_INVALID;
  if (n_targets == 0) return COSMO_OK;
  if (!px || !py || !pz || !mass || !targets || !acc || n == 0) return COSMO_E_INVALID;
  const unsigned grid = static_cast<unsigned>((n_targets + kWave - 1) / kWave);
  k_direct_sum<<<grid, kDirectWaves * kWave, 0, as_stream(stream)>>>(
      px, py, pz, mass, n, targets, n_targets, G, soft * soft, acc);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

}  // extern "C"
