// Barnes-Hut traversal and the FP64 direct sum.
//
// Reference: src/kernels/physics_kernels.py:14-153 (compute_forces), one
// thread per particle with a private 64-entry stack.  Here one wavefront
// walks the tree for 64 Morton-consecutive particles with ONE shared stack
// of node ids, held in one VGPR indexed by lane (v_readlane pops,
// v_writelane pushes).  On a pop every lane of the entry runs its own
// acceptance test, accepted lanes interact, and the lanes that open the
// node push c2 then c1 (so c1 is visited first, physics_kernels.py:144-149).
// Which lanes belong to an entry is not stored: the entries a lane belongs
// to are a prefix of the stack whose length is the lane's own depth (the
// depth of the reference's private stack), so a lane is in the popped entry
// exactly when its depth exceeds the new top.  Each lane therefore visits
// exactly its own nodes in its own depth-first order, so its interaction set
// and FP32 summation order are those of the per-particle walk, and the
// depth counter reproduces the reference's overflow drop
// (`stack_top + 1 < 64`, :143).  Node records are wave-uniform scalar loads.
#include <algorithm>
#include <cstring>

#include "common.h"

namespace cosmo {

struct TravParams {
  double th2d;      // REF: theta^2 in FP64 (the reference's arithmetic)
  float th2;        // REF: theta^2
  float inv_theta;  // AABB: 1/theta
  float theta;
  float G;
  float eps2;
  float half_box;
  float straddle_mu;  // AABB: nodes lighter than this pass the minimum-image guard
  const float *straddle_group;  // AABB: each group's own straddle mass (nullable)
  int wave_stats;   // STATS: also store each wavefront's counts
  int big_bits;     // REF: bits of 2e-7 theta^2 (fixed theta; see k_traverse)
  const int32_t *order;  // launch order of the groups (nullable)
  int64_t n_waves;       // number of groups when order is given
  uint32_t *group_pops;  // nullable: each group's wavefront stack pops
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

// Writes `value` into lane `slot` of v (v_writelane ignores EXEC).  The slot
// goes through M0: a VALU instruction reads at most one SGPR, and value and
// slot are both wave-uniform.
__device__ __forceinline__ int write_lane(int v, int value, int slot) {
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tv_writelane_b32 %0, %1, m0"
               : "+v"(v)
               : "s"(value), "s"(slot)
               : "m0");
  return v;
}

// Neumaier's compensated step: s + x rounded, and its exact rounding error
// (big - t) + small (Fast2Sum on the larger-magnitude operand) added to c.
// Written with the operands picked before the sum is overwritten, so the new
// sum can take the old one's register.
__device__ __forceinline__ void neumaier_add(float &s, float &c, float x) {
  const bool sb = fabsf(s) >= fabsf(x);
  const float big = sb ? s : x;
  const float small = sb ? x : s;
  s = s + x;
  c += (big - s) + small;
}

// The hardware hands workgroup b to XCD b mod 8, each XCD with its own L2.
// xcd_block(b, nb) maps the launch's blocks so that XCD x runs the logical
// blocks of chunks x, x + 8, x + 16, ... of kXcdChunk consecutive blocks
// each: Morton-adjacent groups then meet in one L2, and the fine interleave
// of chunks keeps the XCDs' work balanced.  A bijection on [0, nb): the
// blocks past the last whole round of 8 chunks keep their own index.
// Measured on the walk (forces bit-identical): 32 blocks (128 groups, 8192
// particles) per chunk took the reference-test walk from 4.95 to 4.72 ms at
// 128^3 and 40.5 to 39.4 ms at 256^3, and the AABB walk on the halo state by
// 2%; chunks of 16, 64, 128 and 256 blocks gained less.
constexpr int kXcds = 8;
constexpr int kXcdChunk = 32;

__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int full = nb / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
  if (b >= full) return b;
  const int x = b % kXcds, j = b / kXcds;
  return ((j / kXcdChunk) * kXcds + x) * kXcdChunk + j % kXcdChunk;
}

// float in the constant address space: uniform loads through it are scalar
// loads (the packed records are read-only for the whole traversal).
typedef const __attribute__((address_space(4))) float ConstF;

__device__ __forceinline__ float4 ld4(const ConstF *p) {
  return make_float4(p[0], p[1], p[2], p[3]);
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
// TIGHT: guard on |d_k| == 1/2 itself (|r| < 1, so |r| == 1/2 exactly when
// the minimum-image d has |d| == 1/2): one instruction more than the d^2
// test, and the body almost never runs.  Measured: the AABB walk (many leaf
// terms beyond a quarter box) is 2% faster with it, the reference-test walk
// 1% slower, so each uses its own guard.
template <bool TIGHT = false>
__device__ __forceinline__ void fix_ties(Disp &d, float cx, float cy, float cz, float x, float y,
                                         float z) {
  const bool maybe = TIGHT ? fmaxf(fmaxf(fabsf(d.dx), fabsf(d.dy)), fabsf(d.dz)) == 0.5f
                           : d.d2 >= 0.25f;
  if (__builtin_expect(maybe, 0)) {
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

// The reference test on one node for one lane (physics_kernels.py:122-136):
// accept = sx^2 < theta^2 d^2 and near_ok = d^2 > 1e-7, both as the
// reference decides them in FP64.  The FP32 decision is final unless
// rounding could flip it.  With |d_k - d_exact,k| <= u |c_k - p_k| <= u and a
// few u on the products, E = |error of sx^2 - theta^2 d^2| <=
// theta^2 u (10.1 d^2 + 3 |d|_1) <= 11.8 u theta^2 |d|_1 (|d| <= sqrt(3)/2,
// |d|_1 <= 3/2).  A flip needs |sx^2 - theta^2 d^2| <= 2E, so
// theta^2 d^2 <= sx^2 + 35.4 u, |d|_1 <= sqrt(3) d <= 1.7321 sx / theta +
// 2.52e-3 and 2E <= u (40.9 theta sx + 0.0595 theta^2), which the band
// k1 sx + k2 covers.  Near the 1e-7 cut |d^2 - d_exact^2| <= 2e-10.  Lanes
// of the entry (in) inside either band redo both tests in FP64 exactly as
// the reference does.  A leaf has sx^2 = -1: always accepted, never inside
// the band.  A node with sx^2 bits >= big_bits (sx^2 >= 2e-7 theta^2) is
// "big": a lane within 1e-7 + 2e-10 of its CoM has theta^2 d^2 below half of
// sx^2, far outside the band, so it rejects the node and the cut cannot
// change a decision; for big nodes the cut and its band are skipped.
__device__ __forceinline__ void ref_decide(const float4 &A, const float4 &B, bool leaf,
                                           float x, float y, float z, const Disp &d,
                                           float th2, double th2d, float k1, float k2,
                                           int big_bits, bool in, bool &accept,
                                           bool &near_ok) {
  const float sx2 = B.x;
  const float t = th2 * d.d2;
  accept = sx2 < t;
  const float bnd = __builtin_fmaf(k1, B.y, k2);
  bool unsure;
  if (__float_as_int(B.x) >= big_bits) {
    near_ok = true;
    unsure = in & (fabsf(sx2 - t) <= bnd);
  } else {
    // (the empty asm keeps this path a branch: the compiler would
    // otherwise evaluate it for every node and select)
    asm volatile("");
    near_ok = d.d2 > 1e-7f;
    unsure = in & (int(fabsf(d.d2 - 1e-7f) <= 2e-10f) | int(fabsf(sx2 - t) <= bnd));
  }
  if (__builtin_expect(unsure, 0)) {
    const double e2 = exact_d2(A.x, A.y, A.z, x, y, z);
    if (!leaf) {
      const double sxd = static_cast<double>(pin(B.y));
      accept = (sxd * sxd) < (th2d * e2);
    }
    near_ok = e2 > 1e-7;
  }
}

// PPL particles per lane: a wavefront walks for 64*PPL Morton-consecutive
// particles (lane l holds particles base + 64h + l, h < PPL).  The stack holds
// node ids only.  The lanes of an entry are those that opened its parent, so
// along the stack (bottom to top) the lane sets are nested: the entries that
// hold a lane are a prefix of the stack, and the lane's own depth (the depth
// of the reference's private stack) says how long that prefix is.  A lane
// therefore visits the popped entry exactly when its depth exceeds the new
// top.  With PPL = 2 one scalar node load serves twice the particles, at the
// price of a larger union of visited nodes.  Leaves (node ids < n_tree) run
// the same code with the acceptance test forced true.
template <int MAC, bool COMP, bool STATS, bool GROUP_THETA, int PPL>
__global__ __launch_bounds__(256) void k_traverse(
    const float *__restrict__ px, const float *__restrict__ py,
    const float *__restrict__ pz, const void *__restrict__ tree, int n_tree,
    const int32_t *__restrict__ root_idx, float *__restrict__ force, TravParams p,
    const float *__restrict__ theta_group, int64_t lo, int64_t hi,
    unsigned long long *__restrict__ stats) {
  typedef typename InodeOf<MAC>::type Inode;
  const float4 *__restrict__ leaves = reinterpret_cast<const float4 *>(tree);
  const int lane = threadIdx.x & (kWave - 1);
  // slot = launch position of this wavefront; wave = the group it takes.
  // Without a launch order, runs of kXcdChunk consecutive blocks (Morton-
  // adjacent groups, whose walks share their near-field records) go to one
  // XCD and its L2 (see xcd_block).  A launch order puts the groups of one
  // cost bin together, and those lie scattered in space: it keeps the plain
  // block order, which spreads each bin over the XCDs.
  const int blk = p.order ? static_cast<int>(blockIdx.x)
                          : xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
  const int slot = __builtin_amdgcn_readfirstlane(
      static_cast<int>((int64_t(blk) * blockDim.x + threadIdx.x) / kWave));
  int64_t wave = slot;
  if (p.order) {
    if (slot >= p.n_waves) return;
    wave = p.order[slot];
    if (wave < 0 || wave >= p.n_waves) return;
  }
  const int64_t base = lo + wave * (kWave * PPL);
  if (base >= hi) return;  // wave-uniform
  float x[PPL], y[PPL], z[PPL], th2[PPL], k1[PPL], k2[PPL], inv_theta[PPL];
  double th2d[PPL];
  float ax[PPL], ay[PPL], az[PPL], cx[PPL], cy[PPL], cz[PPL];
  int depth[PPL];
  bool active[PPL];
  uint32_t w_int = 0, w_vis = 0, w_pop = 0, w_drop = 0;  // wave totals (uniform)
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
    float th = p.theta;
    if (GROUP_THETA) {
      const int64_t g = wave * PPL + h;
      const float t = (base + h * kWave < hi) ? theta_group[g] : 0.5f;
      th2[h] = t * t;
      inv_theta[h] = 1.0f / t;
      th2d[h] = static_cast<double>(t) * static_cast<double>(t);
      th = t;
    }
    // rounding band of the reference test as a function of the node's sx
    k1[h] = (48.0f * kU) * th;
    k2[h] = (0.08f * kU) * th2[h];
    ax[h] = ay[h] = az[h] = cx[h] = cy[h] = cz[h] = 0.0f;
    depth[h] = active[h] ? 1 : 0;
  }
  // threshold of the "big" nodes of ref_decide (largest theta of the
  // wavefront), compared as integer bits (sx^2 >= 0; a leaf's -1 is
  // negative), so the test is a scalar compare
  int big_bits = p.big_bits;
  if (GROUP_THETA) {
    float th2max = th2[0];
#pragma unroll
    for (int h = 1; h < PPL; ++h) th2max = fmaxf(th2max, th2[h]);
    big_bits = __builtin_amdgcn_readfirstlane(__float_as_int(th2max * 2e-7f));
  }
  // the straddle mass of this wavefront's group (one group per wavefront
  // when it is given per group)
  float straddle_mu = p.straddle_mu;
  if (MAC == COSMO_MAC_AABB && p.straddle_group)
    straddle_mu =
        __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(p.straddle_group[wave])));
  // m < straddle_mu as an integer compare of the bits on the scalar unit: a
  // node reaching the test has m > 0 (massless nodes are skipped first), and
  // for a non-negative straddle mass the bit patterns order as the values
  // (+inf included; a NaN m compares above every finite limit and keeps the
  // guard, as the float compare does).  A negative or NaN limit keeps the
  // guard for every node, as the float compare does, through bits 0.
  const int straddle_bits = __builtin_amdgcn_readfirstlane(
      straddle_mu >= 0.0f ? __float_as_int(straddle_mu) : 0);
  // internal record k (k >= n_tree) starts at 64 k - 48 n_tree bytes
  const uint32_t inode_bias =
      uint32_t(n_tree) * uint32_t(sizeof(Inode) - sizeof(cosmo_leaf));
  int stk = (lane == 0) ? root : 0;  // lane k holds stack slot k
  int top = 1;

  while (top > 0) {
    --top;
    // (the AABB walk always counts its pops, for cosmo_order_groups_by_pops:
    // one scalar add; the reference-test walk only in its counting variant,
    // where the add measurably slows the loop)
    if (STATS || MAC == COSMO_MAC_AABB) w_pop++;
    const int node = __builtin_amdgcn_readlane(stk, top);
    bool in[PPL];
    uint64_t m_in[PPL];  // the entry's lanes as a scalar mask
#pragma unroll
    for (int h = 0; h < PPL; ++h) {
      in[h] = depth[h] > top;
      m_in[h] = __builtin_amdgcn_ballot_w64(in[h]);
      // min(depth, top): a lane of the popped entry has depth exactly top + 1
      // (no lane's depth exceeds the stack's size), so the update subtracts
      // the compare's bit (one borrow-subtract, no copy of the old depth)
      if constexpr (MAC == COSMO_MAC_REF)
        depth[h] -= in[h] ? 1 : 0;
      else  // (here the compiler made the subtraction a select and a subtract)
        depth[h] = min(depth[h], top);
    }
    // Records are wave-uniform loads read through the constant address space
    // (the records are not written while a traversal runs), so every field
    // is a scalar load.  REF: one 32-byte record per node id, a leaf's with
    // sx^2 = -1 so that the test accepts it; no per-kind offset or field
    // selection in the loop.  AABB: leaf and internal records start with
    // (x, y, z, m); one wave-uniform byte offset is selected for either kind
    // and the whole internal-record size is loaded at once (the buffer has
    // tail room for the last leaf); a leaf uses only its first 16 bytes.
    bool leaf;
    float4 A, B, C, D;
    if constexpr (MAC == COSMO_MAC_REF) {
      // 32-bit byte offset: one scalar shift, and the load adds it to the
      // base (traverse_impl refuses trees whose records pass 4 GiB)
      const ConstF *rec = (const ConstF *)(reinterpret_cast<const char *>(tree) +
                                           uint32_t(node) * uint32_t(sizeof(Inode)));
      A = ld4(rec);
      B = ld4(rec + 4);
      C = D = A;
      leaf = __float_as_int(B.x) < 0;  // only read on the rare FP64 recheck
    } else {
      // 32-bit byte offsets (traverse_impl refuses trees whose records
      // pass 4 GiB): a leaf's record at 16 node, an internal node's at
      // 16 n + 64 (node - n) = 64 node - 48 n, one select between two
      // shifts on the scalar unit.  A leaf's B, C and D hold whatever
      // follows its 16 bytes (the buffer has tail room); the walk reads
      // them only for internal nodes: a leaf is accepted by every lane, so
      // nothing is pushed and its fields go nowhere.
      leaf = node < n_tree;
      const uint32_t ofs = leaf ? uint32_t(node) * uint32_t(sizeof(cosmo_leaf))
                                : uint32_t(node) * uint32_t(sizeof(Inode)) - inode_bias;
      const ConstF *rec = (const ConstF *)(reinterpret_cast<const char *>(tree) + ofs);
      A = ld4(rec);
      B = ld4(rec + 4);
      C = ld4(rec + 8);
      D = ld4(rec + 12);
      // keep the 64 bytes one load, issued before the massless-node exit
      // (the compiler would otherwise sink the B, C, D loads below it and
      // pay a second load latency on every pop); emits no instruction
      asm volatile("" ::"s"(B.x), "s"(B.y), "s"(B.z), "s"(B.w), "s"(C.x), "s"(C.y), "s"(C.z),
                   "s"(C.w), "s"(D.x), "s"(D.y), "s"(D.z), "s"(D.w));
    }
    // massless nodes are skipped (integer test on the uniform bits: m <= 0)
    if (__float_as_int(A.w) <= 0) continue;
    // every lane of a popped entry has depth == top, so the reference's
    // `stack_top + 1 < 64` (physics_kernels.py:143) is uniform here
    const bool room = top + 1 < 64;
    // lanes that open the node, and (AABB) lanes that reject it, as masks
    uint64_t openm[PPL], rejm[PPL];
    uint64_t om = 0;
    if constexpr (MAC == COSMO_MAC_REF) {
      // lanes may open the node only while the stack has room (:143).  A
      // leaf is always used (is_leaf = c1 == -1, :127); its test fails only
      // on a non-finite distance, and then the lanes that open it push the
      // two massless sentinel records that are a leaf's children, which
      // are popped and skipped: no interaction, as in the reference, and
      // no leaf check in the loop.
      const uint64_t m_room = room ? ~0ull : 0ull;
      // The reference test with the lane sets kept as scalar masks: each
      // compare is one ballot, the set logic runs on the scalar unit, and
      // the divergent interaction runs under inverse_ballot(mask) (no lane
      // bools materialised in vector registers).
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        Disp d = displace(A.x, A.y, A.z, x[h], y[h], z[h]);
        // ref_decide() on masks: accept = sx^2 < theta^2 d^2 and near_ok =
        // d^2 > 1e-7 as the reference decides them in FP64; lanes of the
        // entry inside the rounding band redo them in FP64
        const float sx2 = B.x;
        const float t = th2[h] * d.d2;
        uint64_t m_acc = __builtin_amdgcn_ballot_w64(sx2 < t);
        const float bnd = __builtin_fmaf(k1[h], B.y, k2[h]);
        uint64_t m_band = __builtin_amdgcn_ballot_w64(fabsf(sx2 - t) <= bnd);
        uint64_t m_near = ~0ull;
        if (__float_as_int(B.x) < big_bits) {
          // (the empty asm keeps this path a branch)
          asm volatile("");
          m_near = __builtin_amdgcn_ballot_w64(d.d2 > 1e-7f);
          m_band |= __builtin_amdgcn_ballot_w64(fabsf(d.d2 - 1e-7f) <= 2e-10f);
        }
        m_band &= m_in[h];
        if (__builtin_expect(m_band != 0, 0)) {
          bool accept = __builtin_amdgcn_inverse_ballot_w64(m_acc);
          bool near_ok = __builtin_amdgcn_inverse_ballot_w64(m_near);
          if (__builtin_amdgcn_inverse_ballot_w64(m_band)) {
            const double e2 = exact_d2(A.x, A.y, A.z, x[h], y[h], z[h]);
            if (!leaf) {
              const double sxd = static_cast<double>(pin(B.y));
              accept = (sxd * sxd) < (th2d[h] * e2);
            }
            near_ok = e2 > 1e-7;
          }
          m_acc = __builtin_amdgcn_ballot_w64(accept);
          m_near = __builtin_amdgcn_ballot_w64(near_ok);
        }
        const uint64_t m_hit = m_in[h] & m_acc & m_near;
        const uint64_t m_open = m_in[h] & ~m_acc & m_room;
        om |= m_open;
        openm[h] = m_open;
        rejm[h] = 0;
        if (__builtin_amdgcn_inverse_ballot_w64(m_hit)) {
          fix_ties(d, A.x, A.y, A.z, x[h], y[h], z[h]);
          accumulate<COMP>(d, A.w, p.eps2, ax[h], ay[h], az[h], cx[h], cy[h], cz[h]);
        }
        if (STATS) {
          w_vis += __builtin_popcountll(m_in[h]);
          w_int += __builtin_popcountll(m_hit);
          // lanes whose children the full stack drops (:143; leaves have none)
          if (__float_as_int(B.x) >= 0) w_drop += __builtin_popcountll(m_in[h] & ~m_acc & ~m_room);
        }
      }
    } else {
      const uint64_t m_room = room ? ~0ull : 0ull;
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        // every lane evaluates the node; only the entry's lanes (m_in) use it
        Disp d = displace(A.x, A.y, A.z, x[h], y[h], z[h]);
        // (s/theta + dl)^2 < d^2, and every particle of the node sits on the
        // same side of the minimum-image boundary as its CoM in each axis,
        // unless the node is lighter than straddle_mu (a node across the
        // boundary puts at most its mass on the wrong image).  A leaf is
        // always accepted.  The lane sets are scalar masks (one ballot per
        // compare, set logic on the scalar unit).
        uint64_t m_acc = ~0ull;
        if (!leaf) {
          const float r = __builtin_fmaf(B.x, inv_theta[h], B.y);
          m_acc = __builtin_amdgcn_ballot_w64(r * r < d.d2);
          // only internal nodes at or above the straddle mass must pass the
          // guard (a wave-uniform condition); the others skip it
          if (!(__float_as_int(A.w) < straddle_bits)) {
            asm volatile("");
            // the six one-sided tests as one max and one min: the same
            // decisions (a NaN operand, which fmaxf/fminf would drop, only
            // arises with a NaN or infinite CoM, whose d^2 is NaN and fails
            // the distance test anyway)
            const float hb = p.half_box;
            const float up = fmaxf(fmaxf(d.dx + C.w, d.dy + D.x), d.dz + D.y);
            const float dn = fminf(fminf(d.dx - C.x, d.dy - C.y), d.dz - C.z);
            m_acc &= __builtin_amdgcn_ballot_w64(up < hb) & __builtin_amdgcn_ballot_w64(dn > -hb);
          }
        }
        const uint64_t m_near = __builtin_amdgcn_ballot_w64(d.d2 > 1e-7f);
        const uint64_t m_hit = m_in[h] & m_acc & m_near;
        rejm[h] = m_in[h] & ~m_acc;
        openm[h] = rejm[h] & m_room;
        om |= openm[h];
        if (__builtin_amdgcn_inverse_ballot_w64(m_hit)) {
          fix_ties<true>(d, A.x, A.y, A.z, x[h], y[h], z[h]);
          accumulate<COMP>(d, A.w, p.eps2, ax[h], ay[h], az[h], cx[h], cy[h], cz[h]);
        }
        if (STATS) {
          w_vis += __builtin_popcountll(m_in[h]);
          w_int += __builtin_popcountll(m_hit);
        }
      }
    }
    // AABB records of small nodes carry their leaf range (count > 0): the
    // lanes that reject such a node sum its leaves directly, in leaf order,
    // which is the order in which the walk would reach them.  The leaf
    // records are wave-uniform scalar loads.
    int bcount = 0;
    if constexpr (MAC == COSMO_MAC_AABB) bcount = leaf ? 0 : __float_as_int(D.w);
    if (STATS && MAC == COSMO_MAC_AABB && bcount == 0 && !room) {
#pragma unroll
      for (int h = 0; h < PPL; ++h) w_drop += __builtin_popcountll(rejm[h]);
    }
    if constexpr (MAC == COSMO_MAC_AABB) {
      // AABB: the push with its writes made unconditional.  When no lane
      // opens the node (or it is a bucket, whose rejecting lanes sum its
      // leaves below instead) the two writes land in slots at or above the
      // top, which hold nothing live, the depth select changes no lane and
      // the top stays.  With one definition of the stack on every path the
      // compiler keeps the stack and the depths in place (no register
      // copies per pop).
      // The second slot is capped at 63: with a full stack (top = 63, no
      // room, so nothing is pushed) slot 64 would name lane 0 of the
      // register, the bottom entry.
      const bool push = bcount == 0 && om != 0;
      stk = write_lane(stk, __float_as_int(B.w), top);
      stk = write_lane(stk, __float_as_int(B.z), min(top + 1, kWave - 1));
#pragma unroll
      for (int h = 0; h < PPL; ++h)
        depth[h] = __builtin_amdgcn_inverse_ballot_w64(bcount == 0 ? openm[h] : 0ull) ? top + 2
                                                                                  : depth[h];
      top += push ? 2 : 0;
    }
    if (bcount > 0) {
      uint64_t bm = 0;  // lanes that reject the node (stack room is not needed)
      bool rej[PPL];
#pragma unroll
      for (int h = 0; h < PPL; ++h) {
        bm |= rejm[h];
        rej[h] = __builtin_amdgcn_inverse_ballot_w64(rejm[h]);
      }
      if (bm != 0) {
        const int first = __float_as_int(D.z);
        // the bucket's leaf records, read as wave-uniform scalar loads
        const ConstF *lrec = (const ConstF *)(leaves + first);
        // compensated mode: the bucket's terms form a short FP32 partial sum
        // that enters the running sum as one compensated term
        float bx[PPL], by[PPL], bz[PPL];
#pragma unroll
        for (int h = 0; h < PPL; ++h) bx[h] = by[h] = bz[h] = 0.0f;
        // Near buckets.  Let r = fl(c - p) per axis (c the node's CoM) and
        // e = max(|c - lo|, |hi - c|) the largest offset of the node's box
        // from its CoM along that axis.  Every leaf q of the node lies in
        // [lo, hi], so |q - p| <= |c - p| + e.  When |r| + e < 1/4 for
        // every axis of every rejecting lane, |q - p| < 1/2 with room to
        // spare, so fl(q - p) lies strictly inside (-1/2, 1/2): rint gives
        // 0 and no component ties at 1/2.  The minimum-image displacement is
        // then the plain difference, bit for bit what the general loop
        // computes (r - rint(r) = r - 0 = r), and the tie fix never fires.
        uint64_t far = 0;
        {
          const float ex = fmaxf(fabsf(C.x), fabsf(C.w));
          const float ey = fmaxf(fabsf(C.y), fabsf(D.x));
          const float ez = fmaxf(fabsf(C.z), fabsf(D.y));
#pragma unroll
          for (int h = 0; h < PPL; ++h) {
            const float mx = fabsf(A.x - x[h]) + ex;
            const float my = fabsf(A.y - y[h]) + ey;
            const float mz = fabsf(A.z - z[h]) + ez;
            // NaN or inf anywhere fails the compare and takes the general loop
            far |= __ballot(rej[h] & !(fmaxf(fmaxf(mx, my), mz) < 0.25f));
          }
        }
        auto near_leaf = [&](const float4 &L) {
          // massless leaves are skipped, as in the walk (a wave-uniform
          // condition folded into the lane mask rather than a branch)
          const bool live = __float_as_int(L.w) > 0;
#pragma unroll
          for (int h = 0; h < PPL; ++h) {
            Disp d;
            d.dx = d.rx = L.x - x[h];
            d.dy = d.ry = L.y - y[h];
            d.dz = d.rz = L.z - z[h];
            d.d2 = __builtin_fmaf(d.dz, d.dz, __builtin_fmaf(d.dy, d.dy, d.dx * d.dx));
            const bool hb = rej[h] & live & (d.d2 > 1e-7f);
            if (hb) {
              if (COMP)
                accumulate<false>(d, L.w, p.eps2, bx[h], by[h], bz[h], cx[h], cy[h], cz[h]);
              else
                accumulate<false>(d, L.w, p.eps2, ax[h], ay[h], az[h], cx[h], cy[h], cz[h]);
            }
            if (STATS && live) {
              w_vis += __builtin_popcountll(__ballot(rej[h]));
              w_int += __builtin_popcountll(__ballot(hb));
            }
          }
        };
        if (far == 0) {
          // four leaf records per scalar load, then two, then one
          int k = 0;
          for (; k + 4 <= bcount; k += 4) {
            const float4 L0 = ld4(lrec + 4 * k);
            const float4 L1 = ld4(lrec + 4 * k + 4);
            const float4 L2 = ld4(lrec + 4 * k + 8);
            const float4 L3 = ld4(lrec + 4 * k + 12);
            near_leaf(L0);
            near_leaf(L1);
            near_leaf(L2);
            near_leaf(L3);
          }
          if (k + 2 <= bcount) {
            const float4 L0 = ld4(lrec + 4 * k);
            const float4 L1 = ld4(lrec + 4 * k + 4);
            near_leaf(L0);
            near_leaf(L1);
            k += 2;
          }
          if (k < bcount) near_leaf(ld4(lrec + 4 * k));
        } else {
        for (int k = 0; k < bcount; ++k) {
          const float4 L = ld4(lrec + 4 * k);
          const float qx = L.x, qy = L.y, qz = L.z;
          const int qm = __float_as_int(L.w);
          if (qm <= 0) continue;  // massless leaves are skipped, as in the walk
#pragma unroll
          for (int h = 0; h < PPL; ++h) {
            Disp d = displace(qx, qy, qz, x[h], y[h], z[h]);
            const bool hb = rej[h] & (d.d2 > 1e-7f);
            if (hb) {
              fix_ties<true>(d, qx, qy, qz, x[h], y[h], z[h]);
              if (COMP)
                accumulate<false>(d, __int_as_float(qm), p.eps2, bx[h], by[h], bz[h], cx[h],
                                  cy[h], cz[h]);
              else
                accumulate<false>(d, __int_as_float(qm), p.eps2, ax[h], ay[h], az[h], cx[h],
                                  cy[h], cz[h]);
            }
            if (STATS) {
              w_vis += __builtin_popcountll(__ballot(rej[h]));
              w_int += __builtin_popcountll(__ballot(hb));
            }
          }
        }
        }
        if (COMP) {
#pragma unroll
          for (int h = 0; h < PPL; ++h) {
            if (rej[h]) {
              neumaier_add(ax[h], cx[h], bx[h]);
              neumaier_add(ay[h], cy[h], by[h]);
              neumaier_add(az[h], cz[h], bz[h]);
            }
          }
        }
      }
    } else if (MAC == COSMO_MAC_REF && om != 0) {
      // internal nodes have two children (build_tree): push c2, then c1
      // (physics_kernels.py:143-149); without room the reference drops them
      stk = write_lane(stk, __float_as_int(B.w), top);
      stk = write_lane(stk, __float_as_int(B.z), top + 1);
#pragma unroll
      for (int h = 0; h < PPL; ++h)
        depth[h] = __builtin_amdgcn_inverse_ballot_w64(openm[h]) ? top + 2 : depth[h];
      top += 2;
    }
  }
  // G multiplies the finished sums (exact for G = 1)
#pragma unroll
  for (int h = 0; h < PPL; ++h) {
    if (active[h]) {
      const int64_t i = base + h * kWave + lane;
      force[3 * i + 0] = p.G * (ax[h] + cx[h]);
      force[3 * i + 1] = p.G * (ay[h] + cy[h]);
      force[3 * i + 2] = p.G * (az[h] + cz[h]);
    }
  }
  if (STATS && lane == 0) {
    atomicAdd(&stats[0], static_cast<unsigned long long>(w_int));
    atomicAdd(&stats[1], static_cast<unsigned long long>(w_vis));
    atomicAdd(&stats[2], static_cast<unsigned long long>(w_pop));
    if (w_drop) atomicAdd(&stats[3], static_cast<unsigned long long>(w_drop));
    if (p.wave_stats) {
      unsigned long long *ws = stats + 4 + 3 * wave;
      ws[0] = w_int;
      ws[1] = w_vis;
      ws[2] = w_pop;
    }
  }
  // the group's cost for a later launch order (cosmo_order_groups_by_pops)
  if (p.group_pops && lane == 0) p.group_pops[wave] = w_pop;
}

// ------------------------------------------------------------ split walk
//
// A launch over few particles (the reference's 5000-particle batches,
// engine.py:91-111) has few wavefronts, and each walks the whole tree alone
// on its CU: the walk is bound by the latency of its chain of dependent
// record loads.  The split walk gives each group of 64 particles F
// wavefronts, one per subtree below a small top tree (the frontier, built
// by k_split_frontier).  Wavefront (g, q) first decides the top-tree nodes
// on the path to subtree q for its lanes exactly as the walk would (a lane
// that accepts an ancestor interacts with it in the wavefront of the
// leftmost subtree below that ancestor and goes no further; a massless
// ancestor is skipped with its subtree), then walks subtree q for the lanes
// that opened every ancestor, starting at the stack slot the subtree root
// has in the one-wavefront walk, so the 64-entry overflow rule applies at
// the same depths.  Every lane visits and accepts exactly the nodes it does
// in k_traverse; its terms are summed per subtree, and k_split_reduce adds
// the subtree sums in subtree order: the same terms, associated
// differently.  Reference test only (packed 32-byte records per node id).

// The frontier: a cut of the top tree into at most `target` subtrees, built
// by one block with one entry per thread, kept in depth-first order (c1
// before c2).  Children come from the records (c1, c2 at words 6 and 7).
// Each entry carries its leaf range: the root of a tree built by
// k_build_tree is internal node 0 (id n_tree) over [0, n_tree - 1], and an
// internal node's c1 ends its range at the split index (c1's leaf id, or its
// internal index = id - n_tree).  Subtrees are split largest-first by
// weight = leaves, times `focus` for a subtree that holds some of the
// launch's own particles [lo, hi): the near field of every group lies
// there, and its walks are the longest of the launch.  Each round expands
// the entries within a factor 2 of the heaviest one (heaviest first when
// they exceed the budget), so the frontier takes a few dependent record
// loads, not one per expansion.  Any cut gives the same visits and
// interactions per particle; the weight only shapes the work per
// wavefront.  With another root the ranges are unknown and the weight falls
// back to 2^-depth (breadth-first).
//
// Expanded nodes go to an append-only history (node, stack slot, history
// index of its parent, whether it is its parent's c1); an entry keeps the
// history index of its parent and whether it is a c1.  Paths and leftmost
// masks are read back from the history once, at the end.  The block's
// waves agree on the round's maximum weight and on each entry's new
// position (entries before it that expand) through LDS.
constexpr int kFrontierThreads = kSplitMaxFrontier;
constexpr int kFrontierWaves = kFrontierThreads / kWave;

//
// Block b builds the frontier of the range [lo + b stride, min(n_limit,
// lo + b stride + span)) into fr[b]: one block for a launch's own range, or
// one per window of a batch loop's plan.
__global__ __launch_bounds__(kFrontierThreads) void k_split_frontier(
    const void *__restrict__ tree, int n_tree, const int32_t *__restrict__ root_idx, int target,
    int64_t lo0, int64_t span, int64_t stride, int64_t n_limit, uint64_t focus,
    SplitFrontier *__restrict__ frontiers) {
  const int64_t lo = lo0 + int64_t(blockIdx.x) * stride;
  const int64_t hi = lo + span < n_limit ? lo + span : n_limit;
  SplitFrontier *__restrict__ fr = frontiers + blockIdx.x;
  constexpr int kT = kFrontierThreads;
  constexpr int kHist = kSplitMaxFrontier;  // at most target - 1 expansions
  __shared__ int s_node[2][kT], s_slot[2][kT], s_plen[2][kT], s_hpar[2][kT], s_left[2][kT];
  __shared__ int s_rlo[2][kT], s_rhi[2][kT];
  __shared__ int h_node[kHist], h_slot[kHist], h_par[kHist], h_left[kHist];
  __shared__ uint64_t s_w[kT];
  __shared__ uint64_t s_wmax[kFrontierWaves];
  __shared__ int s_nexp[kFrontierWaves];
  const int t = threadIdx.x;
  const int wid = t / kWave;
  const bool lead = (t & (kWave - 1)) == 0;
  const int32_t *words = reinterpret_cast<const int32_t *>(tree);
  const int root = root_idx[0];
  const bool ranged = root == n_tree;
  int cnt = root < 0 ? 0 : 1;
  int hcnt = 0;
  int cur = 0;
  // this thread's entry
  int node = root, slot = 0, plen = 0, hpar = -1, left = 0, rlo = 0, rhi = n_tree - 1;
  for (int round = 0; round < 2 * kSplitMaxPath && cnt < target; ++round) {
    const bool have = t < cnt;
    const bool expandable =
        have && node >= n_tree && slot + 1 < 64 && plen + 1 < kSplitMaxPath;
    uint64_t w = 0;
    if (expandable) {
      if (ranged) {
        w = static_cast<uint64_t>(rhi >= rlo ? rhi - rlo + 1 : 1);
        if (rlo < hi && rhi >= lo) w *= focus;
      } else {
        w = uint64_t(1) << (kSplitMaxPath - plen);
      }
    }
    uint64_t wm = w;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
      const uint64_t o = __shfl_xor(wm, m);
      wm = o > wm ? o : wm;
    }
    if (lead) s_wmax[wid] = wm;
    __syncthreads();
    uint64_t wmax = 0;
#pragma unroll
    for (int u = 0; u < kFrontierWaves; ++u) wmax = s_wmax[u] > wmax ? s_wmax[u] : wmax;
    if (wmax == 0) break;  // the same value in every thread
    bool expand = w != 0 && 2 * w > wmax;
    const int budget = target - cnt;
    {
      const int pc = __popcll(__ballot(expand));  // ballot with every lane active
      if (lead) s_nexp[wid] = pc;
    }
    s_w[t] = expand ? w : 0;
    __syncthreads();
    int total = 0;
#pragma unroll
    for (int u = 0; u < kFrontierWaves; ++u) total += s_nexp[u];
    if (total > budget) {  // heaviest first, ties by position
      int rank = 0;
      for (int u = 0; u < cnt; ++u) {
        const uint64_t wu = s_w[u];
        rank += (wu > w || (wu == w && u < t)) ? 1 : 0;
      }
      expand = expand && rank < budget;
      const int pc = __popcll(__ballot(expand));
      __syncthreads();
      if (lead) s_nexp[wid] = pc;
      __syncthreads();
    }
    // new position: this entry moves up by the number of entries before it
    // that expand (each adds one)
    int below = __popcll(__ballot(expand) & lanemask_lt());
    int nexp = 0;
#pragma unroll
    for (int u = 0; u < kFrontierWaves; ++u) {
      below += u < wid ? s_nexp[u] : 0;
      nexp += s_nexp[u];
    }
    const int pos = t + below;
    const int nxt = cur ^ 1;
    if (have) {
      if (expand) {
        const int h = hcnt + below;
        h_node[h] = node;
        h_slot[h] = slot;
        h_par[h] = hpar;
        h_left[h] = left;
        const int c1 = words[8 * int64_t(node) + 6], c2 = words[8 * int64_t(node) + 7];
        const int split = c1 < n_tree ? c1 : c1 - n_tree;
        const int mid = min(max(split, rlo), rhi - 1);
        s_node[nxt][pos] = c1;
        s_slot[nxt][pos] = slot + 1;
        s_plen[nxt][pos] = plen + 1;
        s_hpar[nxt][pos] = h;
        s_left[nxt][pos] = 1;
        s_rlo[nxt][pos] = rlo;
        s_rhi[nxt][pos] = mid;
        s_node[nxt][pos + 1] = c2;
        s_slot[nxt][pos + 1] = slot;
        s_plen[nxt][pos + 1] = plen + 1;
        s_hpar[nxt][pos + 1] = h;
        s_left[nxt][pos + 1] = 0;
        s_rlo[nxt][pos + 1] = mid + 1;
        s_rhi[nxt][pos + 1] = rhi;
      } else {
        s_node[nxt][pos] = node;
        s_slot[nxt][pos] = slot;
        s_plen[nxt][pos] = plen;
        s_hpar[nxt][pos] = hpar;
        s_left[nxt][pos] = left;
        s_rlo[nxt][pos] = rlo;
        s_rhi[nxt][pos] = rhi;
      }
    }
    hcnt += nexp;
    cnt += nexp;
    cur = nxt;
    __syncthreads();
    if (t < cnt) {
      node = s_node[cur][t];
      slot = s_slot[cur][t];
      plen = s_plen[cur][t];
      hpar = s_hpar[cur][t];
      left = s_left[cur][t];
      rlo = s_rlo[cur][t];
      rhi = s_rhi[cur][t];
    }
  }
  if (t < cnt) {
    fr->node[t] = node;
    fr->slot[t] = slot;
    fr->plen[t] = plen;
    // walk up the history: path[k] is the ancestor at depth k; the entry
    // is leftmost under path[k] while every step up from it is a c1
    uint32_t lmask = 0;
    bool chain = left != 0;
    int h = hpar;
    for (int k = plen - 1; k >= 0; --k) {
      fr->path[t][k] = h_node[h];
      fr->pslot[t][k] = h_slot[h];
      if (chain) lmask |= 1u << k;
      chain = chain && h_left[h] != 0;
      h = h_par[h];
    }
    fr->lmask[t] = lmask;
  }
  if (t == 0) fr->count = cnt;
}

template <bool STATS>
__global__ __launch_bounds__(256) void k_traverse_split(
    const float *__restrict__ px, const float *__restrict__ py, const float *__restrict__ pz,
    const void *__restrict__ tree, const SplitFrontier *__restrict__ fr, int F,
    float *__restrict__ partial, TravParams p, int64_t lo, int64_t hi,
    unsigned long long *__restrict__ stats) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(
      static_cast<int>((int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave));
  const int g = w / F, q = w % F;
  if (q >= fr->count) return;
  const int64_t base = lo + int64_t(g) * kWave;
  if (base >= hi) return;
  const int64_t i = base + lane;
  const bool active = i < hi;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (active) {
    x = px[i];
    y = py[i];
    z = pz[i];
  }
  const float th2 = p.th2;
  const double th2d = p.th2d;
  const float k1 = (48.0f * kU) * p.theta, k2 = (0.08f * kU) * th2;
  const int big_bits = p.big_bits;
  float ax = 0.0f, ay = 0.0f, az = 0.0f, unused0 = 0.0f, unused1 = 0.0f, unused2 = 0.0f;
  uint32_t w_int = 0, w_vis = 0;
  const char *tb = reinterpret_cast<const char *>(tree);
  const int plen = fr->plen[q];
  const uint32_t lmask = fr->lmask[q];
  bool reach = active;
  for (int k = 0; k < plen; ++k) {
    const int anc = fr->path[q][k];
    const ConstF *rec = (const ConstF *)(tb + int64_t(anc) * int64_t(sizeof(cosmo_inode_ref)));
    const float4 A = ld4(rec), B = ld4(rec + 4);
    if (__float_as_int(A.w) <= 0) {  // the walk skips a massless node and its subtree
      reach = false;
      break;
    }
    Disp d = displace(A.x, A.y, A.z, x, y, z);
    bool accept, near_ok;
    ref_decide(A, B, __float_as_int(B.x) < 0, x, y, z, d, th2, th2d, k1, k2, big_bits, reach,
               accept, near_ok);
    const bool hit = reach & accept & near_ok;
    if ((lmask >> k) & 1u) {
      if (hit) {
        fix_ties(d, A.x, A.y, A.z, x, y, z);
        accumulate<false>(d, A.w, p.eps2, ax, ay, az, unused0, unused1, unused2);
      }
      if (STATS) {
        w_vis += __builtin_popcountll(__ballot(reach));
        w_int += __builtin_popcountll(__ballot(hit));
      }
    }
    reach = reach & !accept & (fr->pslot[q][k] + 1 < 64);
  }
  const int floor_slot = fr->slot[q];
  if (__ballot(reach) != 0) {
    int depth = reach ? floor_slot + 1 : 0;
    int stk = (lane == floor_slot) ? fr->node[q] : 0;
    // The walk of k_traverse with the next record loaded ahead: while a node
    // is decided, the records of both possible next pops (its c1 if it is
    // opened, else the entry below it) are already on their way, so the
    // load latency overlaps the decision instead of following it.
    int top = floor_slot;  // slot of the current node
    int node = fr->node[q];
    float4 A = ld4((const ConstF *)(tb + int64_t(node) * int64_t(sizeof(cosmo_inode_ref))));
    float4 B = ld4((const ConstF *)(tb + int64_t(node) * int64_t(sizeof(cosmo_inode_ref))) + 4);
    for (;;) {
      const bool in = depth > top;
      depth = min(depth, top);
      const int below = top > floor_slot ? __builtin_amdgcn_readlane(stk, top - 1) : node;
      const int c1 = __float_as_int(B.z);  // a leaf's: the sentinel record
      const ConstF *rb = (const ConstF *)(tb + int64_t(below) * int64_t(sizeof(cosmo_inode_ref)));
      const ConstF *rc = (const ConstF *)(tb + int64_t(c1) * int64_t(sizeof(cosmo_inode_ref)));
      const float4 Ab = ld4(rb), Bb = ld4(rb + 4), Ac = ld4(rc), Bc = ld4(rc + 4);
      bool pushed = false;
      if (__float_as_int(A.w) > 0) {  // massless nodes are skipped
        const bool room = top + 1 < 64;
        Disp d = displace(A.x, A.y, A.z, x, y, z);
        bool accept, near_ok;
        ref_decide(A, B, __float_as_int(B.x) < 0, x, y, z, d, th2, th2d, k1, k2, big_bits, in,
                   accept, near_ok);
        const bool hit = in & accept & near_ok;
        const bool open = in & !accept & room;
        if (hit) {
          fix_ties(d, A.x, A.y, A.z, x, y, z);
          accumulate<false>(d, A.w, p.eps2, ax, ay, az, unused0, unused1, unused2);
        }
        if (STATS) {
          w_vis += __builtin_popcountll(__ballot(in));
          w_int += __builtin_popcountll(__ballot(hit));
        }
        // (a leaf opened on a non-finite distance pushes two sentinels)
        if (__ballot(open) != 0) {
          stk = write_lane(stk, __float_as_int(B.w), top);
          stk = write_lane(stk, c1, top + 1);
          depth = open ? top + 2 : depth;
          pushed = true;
        }
      }
      if (pushed) {  // next pop: c1, at slot top + 1
        top += 1;
        node = c1;
        A = Ac;
        B = Bc;
      } else if (top > floor_slot) {  // next pop: the entry below
        top -= 1;
        node = below;
        A = Ab;
        B = Bb;
      } else {
        break;
      }
    }
  }
  if (active) {
    float *o = partial + (int64_t(q) * (hi - lo) + (i - lo)) * 3;
    o[0] = ax;
    o[1] = ay;
    o[2] = az;
  }
  if (STATS && lane == 0) {
    atomicAdd(&stats[0], static_cast<unsigned long long>(w_int));
    atomicAdd(&stats[1], static_cast<unsigned long long>(w_vis));
  }
}

// force = G * (sum of the subtree partial sums).  A block takes 32
// consecutive force components; its 8 waves' halves each sum one contiguous
// chunk of the subtrees in subtree order, and the chunk sums are added in
// chunk order: a fixed association, independent of scheduling, with eight
// times the loads in flight of one thread per component.
constexpr int kReduceComps = 32;
constexpr int kReduceChunks = 8;

__global__ __launch_bounds__(kReduceComps * kReduceChunks) void k_split_reduce(
    const float *__restrict__ partial, const SplitFrontier *__restrict__ fr, int64_t lo,
    int64_t hi, float G, float *__restrict__ force) {
  __shared__ float s_sum[kReduceChunks][kReduceComps];
  const int c = threadIdx.x % kReduceComps, k = threadIdx.x / kReduceComps;
  const int64_t e = int64_t(blockIdx.x) * kReduceComps + c;  // component 3 j + axis
  const int64_t m3 = (hi - lo) * 3;
  const int nf = fr->count;
  const int per = (nf + kReduceChunks - 1) / kReduceChunks;
  const int q0 = k * per, q1 = min(nf, q0 + per);
  float sum = 0.0f;
  if (e < m3) {
    const float *o = partial + e;
    int q = q0;
    for (; q + 4 <= q1; q += 4) {  // four loads in flight, added in subtree order
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = o[int64_t(q + u) * m3];
#pragma unroll
      for (int u = 0; u < 4; ++u) sum += v[u];
    }
    for (; q < q1; ++q) sum += o[int64_t(q) * m3];
  }
  s_sum[k][c] = sum;
  __syncthreads();
  if (k == 0 && e < m3) {
    float tot = s_sum[0][c];
#pragma unroll
    for (int u = 1; u < kReduceChunks; ++u) tot += s_sum[u][c];
    force[3 * lo + e] = G * tot;
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

// Launch order for per-group theta: groups by increasing theta (smaller
// theta, more interactions, earlier), so the longest wavefronts start first
// and the tail of the launch holds short ones.  64 bins over log2(theta) in
// [-6, 1); order within a bin is arbitrary (results do not depend on it).
constexpr int kOrderThreads = 1024;
constexpr int kOrderBins = 64;

__device__ __forceinline__ int order_bin(float t) {
  const float b = (log2f(fmaxf(t, 1e-30f)) + 6.0f) * (kOrderBins / 7.0f);
  return b < 0.0f ? 0 : (b >= kOrderBins - 1 ? kOrderBins - 1 : static_cast<int>(b));
}

__global__ __launch_bounds__(kOrderThreads) void k_order_groups(const float *__restrict__ tg,
                                                                int64_t groups,
                                                                int32_t *__restrict__ order) {
  __shared__ int hist[kOrderBins];
  if (threadIdx.x < kOrderBins) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t g = threadIdx.x; g < groups; g += kOrderThreads) atomicAdd(&hist[order_bin(tg[g])], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int b = 0; b < kOrderBins; ++b) {
      const int c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int64_t g = threadIdx.x; g < groups; g += kOrderThreads)
    order[atomicAdd(&hist[order_bin(tg[g])], 1)] = static_cast<int32_t>(g);
}

// Launch order from measured costs: groups by decreasing pop count of an
// earlier launch (a group of the same Morton range walks nearly the same
// nodes one step later), so the longest walks start first.  When the
// groups' candidates of a per-group selection are given for both launches,
// a group's pops are scaled by cost[k_now] / cost[k_prev]: the few groups
// that switch candidates between steps switch between very different costs.
// 64 bins of a third of an octave each over log2(1 + estimate); order
// within a bin is arbitrary.
__device__ __forceinline__ int pops_bin(float est) {
  const float b = log2f(1.0f + est) * 3.0f;
  return kOrderBins - 1 - (b >= kOrderBins - 1 ? kOrderBins - 1 : static_cast<int>(b));
}

__device__ __forceinline__ float pops_estimate(const uint32_t *__restrict__ pops,
                                               const int32_t *__restrict__ k_now,
                                               const int32_t *__restrict__ k_prev,
                                               const float *__restrict__ cost, int K,
                                               int64_t g) {
  float est = static_cast<float>(pops[g]);
  if (k_now) {
    const int a = k_now[g], b = k_prev[g];
    if (a >= 0 && a < K && b >= 0 && b < K && cost[b] > 0.0f) est *= cost[a] / cost[b];
  }
  return est;
}

__global__ __launch_bounds__(kOrderThreads) void k_order_by_pops(
    const uint32_t *__restrict__ pops, const int32_t *__restrict__ k_now,
    const int32_t *__restrict__ k_prev, const float *__restrict__ cost, int K, int64_t groups,
    int32_t *__restrict__ order) {
  __shared__ int hist[kOrderBins];
  if (threadIdx.x < kOrderBins) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t g = threadIdx.x; g < groups; g += kOrderThreads)
    atomicAdd(&hist[pops_bin(pops_estimate(pops, k_now, k_prev, cost, K, g))], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int b = 0; b < kOrderBins; ++b) {
      const int c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int64_t g = threadIdx.x; g < groups; g += kOrderThreads)
    order[atomicAdd(&hist[pops_bin(pops_estimate(pops, k_now, k_prev, cost, K, g))], 1)] =
        static_cast<int32_t>(g);
}

__host__ __device__ constexpr bool pairs_of(int flags) { return (flags & COSMO_TRAV_PAIRS) != 0; }

}  // namespace cosmo

using namespace cosmo;

extern "C" {

static int traverse_impl(const float *px, const float *py, const float *pz, const void *nodes,
                         int64_t n_tree, const int32_t *root_idx, float *force, double theta,
                         const float *theta_group, double G, double soft, double box,
                         int64_t off, int64_t count, int64_t n_bodies, int mac, int flags,
                         uint64_t *stats, const int32_t *order, double straddle_mass,
                         const float *straddle_group, uint32_t *group_pops, void *stream) {
  if (!(straddle_mass >= 0.0) || (straddle_mass > 0.0 && mac != COSMO_MAC_AABB))
    return COSMO_E_INVALID;
  // pop counts are kept per 64-particle group, by the AABB walk
  if (group_pops && ((flags & COSMO_TRAV_PAIRS) || mac != COSMO_MAC_AABB))
    return COSMO_E_INVALID;
  // per-group straddle masses: AABB test, one 64-particle group per wavefront
  if (straddle_group && (mac != COSMO_MAC_AABB || (flags & COSMO_TRAV_PAIRS)))
    return COSMO_E_INVALID;
  if (n_bodies < 0 || n_bodies > kMaxBodies || off < 0 || count < 0 || n_tree < 1 ||
      n_tree > kMaxBodies)
    return COSMO_E_INVALID;
  if (mac != COSMO_MAC_REF && mac != COSMO_MAC_AABB) return COSMO_E_INVALID;
  // the walk addresses its records with 32-bit byte offsets
  if (cosmo_tree_bytes(n_tree, mac) > (int64_t(1) << 32)) return COSMO_E_TOO_LARGE;
  const bool stats_on = (flags & (COSMO_TRAV_STATS | COSMO_TRAV_GROUP_STATS)) != 0;
  if (stats_on && !stats) return COSMO_E_INVALID;
  const int64_t lo = off;
  const int64_t hi = (off + count < n_bodies) ? off + count : n_bodies;
  if (hi <= lo) return COSMO_OK;
  if (!px || !py || !pz || !nodes || !root_idx || !force) return COSMO_E_INVALID;
  TravParams p;
  const float thf = static_cast<float>(theta);
  p.th2d = theta * theta;
  p.th2 = thf * thf;
  {
    const float big = p.th2 * 2e-7f;
    memcpy(&p.big_bits, &big, sizeof big);
  }
  p.inv_theta = 1.0f / thf;
  p.theta = thf;
  p.G = static_cast<float>(G);
  const float sf = static_cast<float>(soft);
  p.eps2 = sf * sf;
  p.half_box = static_cast<float>(box) * 0.5f;
  p.straddle_mu = static_cast<float>(straddle_mass);
  p.straddle_group = straddle_group;
  p.wave_stats = (flags & COSMO_TRAV_GROUP_STATS) != 0;
  p.order = order;
  p.group_pops = group_pops;
  p.n_waves = (hi - lo + kWave * (pairs_of(flags) ? 2 : 1) - 1) / (kWave * (pairs_of(flags) ? 2 : 1));
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

int cosmo_traverse(const float *px, const float *py, const float *pz, const void *nodes,
                   int64_t n_tree, const int32_t *root_idx, float *force, double theta,
                   const float *theta_group, double G, double soft, double box, int64_t off,
                   int64_t count, int64_t n_bodies, int mac, int flags, uint64_t *stats,
                   void *stream) {
  return traverse_impl(px, py, pz, nodes, n_tree, root_idx, force, theta, theta_group, G, soft,
                       box, off, count, n_bodies, mac, flags, stats, nullptr, 0.0, nullptr,
                       nullptr, stream);
}

int cosmo_traverse_ordered(const float *px, const float *py, const float *pz, const void *nodes,
                           int64_t n_tree, const int32_t *root_idx, float *force, double theta,
                           const float *theta_group, double G, double soft, double box,
                           int64_t off, int64_t count, int64_t n_bodies, int mac, int flags,
                           uint64_t *stats, const int32_t *group_order, void *stream) {
  if (!group_order || (flags & COSMO_TRAV_PAIRS)) return COSMO_E_INVALID;
  return traverse_impl(px, py, pz, nodes, n_tree, root_idx, force, theta, theta_group, G, soft,
                       box, off, count, n_bodies, mac, flags, stats, group_order, 0.0, nullptr,
                       nullptr, stream);
}

int cosmo_traverse_ex(const float *px, const float *py, const float *pz, const void *nodes,
                      int64_t n_tree, const int32_t *root_idx, float *force, double theta,
                      const float *theta_group, double G, double soft, double box, int64_t off,
                      int64_t count, int64_t n_bodies, int mac, int flags, uint64_t *stats,
                      const int32_t *group_order, double straddle_mass,
                      const float *straddle_group, uint32_t *group_pops, void *stream) {
  if (group_order && (flags & COSMO_TRAV_PAIRS)) return COSMO_E_INVALID;
  return traverse_impl(px, py, pz, nodes, n_tree, root_idx, force, theta, theta_group, G, soft,
                       box, off, count, n_bodies, mac, flags, stats, group_order, straddle_mass,
                       straddle_group, group_pops, stream);
}

int cosmo_order_groups_by_pops(const uint32_t *group_pops, int64_t groups,
                               const int32_t *k_now, const int32_t *k_prev, const float *cost,
                               int K, int32_t *order, void *stream) {
  if (groups < 0 || groups > kMaxBodies) return COSMO_E_INVALID;
  if ((k_now != nullptr) != (k_prev != nullptr) || (k_now && (!cost || K < 1)))
    return COSMO_E_INVALID;
  if (groups == 0) return COSMO_OK;
  if (!group_pops || !order) return COSMO_E_INVALID;
  k_order_by_pops<<<1, kOrderThreads, 0, as_stream(stream)>>>(group_pops, k_now, k_prev, cost,
                                                               k_now ? K : 0, groups, order);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_order_groups(const float *theta_group, int64_t groups, int32_t *order, void *stream) {
  if (groups < 0 || groups > kMaxBodies) return COSMO_E_INVALID;
  if (groups == 0) return COSMO_OK;
  if (!theta_group || !order) return COSMO_E_INVALID;
  k_order_groups<<<1, kOrderThreads, 0, as_stream(stream)>>>(theta_group, groups, order);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

// The walk of compute_forces over the reference-test records in ctx->nodes:
// a launch of at most 1024 groups cannot fill the device with one wavefront
// per group, so it splits each group's walk over the top tree.
//
// Batch loops (engine.py:91-111: one call per 5000 particles on one tree)
// with a nonzero epoch use planned frontiers: on the first call for an
// (epoch, root, n, fan-out) one launch builds the frontiers of every window
// [j W/2, j W/2 + W) of the tree (W = the power of two >= 2 * count, so every
// range of `count` particles lies inside window floor(off / (W/2))), each
// weighted towards its window's leaves; later calls of the loop take their
// window's frontier and skip the per-launch build.  The cut only decides
// which wavefront walks which subtree: visits and interactions per particle
// are those of the walk either way.  The caller changes the epoch whenever
// the records or the root change.
static int split_walk(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                      const void *nodes, int64_t n_tree, const int32_t *root_idx, float *force,
                      double theta, double G, double soft, int64_t lo, int64_t hi,
                      const SplitFrontier *planned, int F_planned, uint64_t *stats,
                      hipStream_t s);

static int forces_from_ctx_records(cosmo_ctx *ctx, const float *px, const float *py,
                                   const float *pz, float *force, double theta, double G,
                                   double soft, double box, const int32_t *root_idx,
                                   int64_t off, int64_t count, int64_t n_bodies, uint64_t epoch,
                                   void *stream) {
  const int64_t lo = off, hi = std::min(off + count, n_bodies);
  if (hi <= lo) return COSMO_OK;
  if (!px || !py || !pz || !root_idx || !force) return COSMO_E_INVALID;
  if (hi - lo > std::min<int64_t>(ctx->n_max, 1024 * kWave))
    return cosmo_traverse(px, py, pz, ctx->nodes, n_bodies, root_idx, force, theta, nullptr, G,
                          soft, box, off, count, n_bodies, COSMO_MAC_REF, 0, nullptr, stream);
  hipStream_t s = as_stream(stream);
  const SplitFrontier *planned = nullptr;
  int F = 0;
  if (epoch != 0) {
    int64_t W = 128;
    while (W < 2 * (hi - lo)) W <<= 1;
    const int64_t windows = (n_bodies + W / 2 - 1) / (W / 2);
    if (windows <= ctx->plan_cap) {
      // a plan built for launches of plan_span particles serves any shorter
      // launch with the same window (the loop's last batch) at its fan-out
      if (ctx->plan_epoch != epoch || ctx->plan_root != root_idx || ctx->plan_n != n_bodies ||
          ctx->plan_win != W || hi - lo > ctx->plan_span) {
        const int Fp = split_fanout(hi - lo);
        const uint64_t focus = static_cast<uint64_t>(std::max<int64_t>(8, n_bodies / W));
        k_split_frontier<<<static_cast<unsigned>(windows), kFrontierThreads, 0, s>>>(
            ctx->nodes, static_cast<int>(n_bodies), root_idx, Fp, 0, W, W / 2, n_bodies, focus,
            reinterpret_cast<SplitFrontier *>(ctx->split_plans));
        COSMO_CHECK_LAUNCH();
        ctx->plan_epoch = epoch;
        ctx->plan_root = root_idx;
        ctx->plan_n = n_bodies;
        ctx->plan_win = W;
        ctx->plan_span = hi - lo;
        ctx->plan_fanout = Fp;
      }
      planned = reinterpret_cast<const SplitFrontier *>(ctx->split_plans) + lo / (W / 2);
      F = ctx->plan_fanout;
    }
  }
  return split_walk(ctx, px, py, pz, ctx->nodes, n_bodies, root_idx, force, theta, G, soft, lo,
                    hi, planned, F, nullptr, s);
}

int cosmo_compute_forces(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                         const int32_t *children, const double *n_mass, const double *n_com,
                         const float *n_min, const float *n_max, float *force, double theta,
                         double G, double soft, double box, const int32_t *root_idx,
                         const int32_t *top_nodes, const int32_t *node_to_cache, int64_t off,
                         int64_t count, int64_t n_bodies, void *stream) {
  (void)top_nodes;
  (void)node_to_cache;
  if (!ctx || off < 0 || count < 0) return COSMO_E_INVALID;
  if (n_bodies > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (n_bodies < 1) return COSMO_OK;
  const int rc = cosmo_pack_nodes(children, n_mass, n_com, n_min, n_max, 2 * n_bodies - 1, box,
                                  COSMO_MAC_REF, ctx->nodes, stream);
  if (rc != COSMO_OK) return rc;
  ctx->plan_epoch = 0;  // the records were just rewritten
  return forces_from_ctx_records(ctx, px, py, pz, force, theta, G, soft, box, root_idx, off,
                                 count, n_bodies, 0, stream);
}

int cosmo_compute_forces_f32(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                             const int32_t *children, const float *n_mass, const float *n_com,
                             const float *n_min, const float *n_max, float *force,
                             double theta, double G, double soft, double box,
                             const int32_t *root_idx, const int32_t *top_nodes,
                             const int32_t *node_to_cache, int64_t off, int64_t count,
                             int64_t n_bodies, void *stream) {
  (void)top_nodes;
  (void)node_to_cache;
  if (!ctx || off < 0 || count < 0) return COSMO_E_INVALID;
  if (n_bodies > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (n_bodies < 1) return COSMO_OK;
  const int rc = cosmo_pack_nodes_f32(children, n_mass, n_com, n_min, n_max, 2 * n_bodies - 1,
                                      box, COSMO_MAC_REF, ctx->nodes, stream);
  if (rc != COSMO_OK) return rc;
  ctx->plan_epoch = 0;  // the records were just rewritten
  return forces_from_ctx_records(ctx, px, py, pz, force, theta, G, soft, box, root_idx, off,
                                 count, n_bodies, 0, stream);
}

int cosmo_compute_forces_packed(cosmo_ctx *ctx, const float *px, const float *py,
                                const float *pz, float *force, double theta, double G,
                                double soft, double box, const int32_t *root_idx, int64_t off,
                                int64_t count, int64_t n_bodies, uint64_t epoch, void *stream) {
  if (!ctx || off < 0 || count < 0) return COSMO_E_INVALID;
  if (n_bodies > ctx->n_max) return COSMO_E_TOO_LARGE;
  if (n_bodies < 1) return COSMO_OK;
  return forces_from_ctx_records(ctx, px, py, pz, force, theta, G, soft, box, root_idx, off,
                                 count, n_bodies, epoch, stream);
}

// The split walk of particles [lo, hi) over the reference-test records
// `nodes` of an n_tree-leaf tree.  planned: a frontier built with target
// F_planned for a range that contains [lo, hi), where F_planned * (hi - lo)
// fits the partial buffer (nullptr: build the launch's own).
static int split_walk(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                      const void *nodes, int64_t n_tree, const int32_t *root_idx, float *force,
                      double theta, double G, double soft, int64_t lo, int64_t hi,
                      const SplitFrontier *planned, int F_planned, uint64_t *stats,
                      hipStream_t s) {
  TravParams p;
  memset(&p, 0, sizeof p);
  const float thf = static_cast<float>(theta);
  p.th2d = theta * theta;
  p.th2 = thf * thf;
  {
    const float big = p.th2 * 2e-7f;
    memcpy(&p.big_bits, &big, sizeof big);
  }
  p.theta = thf;
  p.G = static_cast<float>(G);
  const float sf = static_cast<float>(soft);
  p.eps2 = sf * sf;
  const int64_t groups = (hi - lo + kWave - 1) / kWave;
  // F subtrees per group, at most kSplitWaveBudget wavefronts in all: the
  // heaviest subtree walks (the near field of the groups at the batch's
  // spatial boundary) bound the launch, and a finer cut shortens them
  const int F = planned ? F_planned : split_fanout(hi - lo);
  const SplitFrontier *fr = planned;
  if (!fr) {
    SplitFrontier *own = reinterpret_cast<SplitFrontier *>(ctx->split_frontier);
    // weight of the subtrees that hold the launch's particles: half the
    // number of launches of this size that would cover the tree, at least 8
    const uint64_t focus = static_cast<uint64_t>(std::max<int64_t>(8, n_tree / (hi - lo) / 2));
    k_split_frontier<<<1, kFrontierThreads, 0, s>>>(nodes, static_cast<int>(n_tree), root_idx,
                                                     F, lo, hi - lo, 0, hi, focus, own);
    fr = own;
  }
  const unsigned grid = blocks_for(groups * F * kWave, 256);
  if (stats)
    k_traverse_split<true><<<grid, 256, 0, s>>>(px, py, pz, nodes, fr, F, ctx->split_partial, p,
                                                lo, hi, reinterpret_cast<unsigned long long *>(stats));
  else
    k_traverse_split<false><<<grid, 256, 0, s>>>(px, py, pz, nodes, fr, F, ctx->split_partial, p,
                                                 lo, hi, nullptr);
  k_split_reduce<<<blocks_for(3 * (hi - lo), kReduceComps), kReduceComps * kReduceChunks, 0, s>>>(
      ctx->split_partial, fr, lo, hi, p.G, force);
  COSMO_CHECK_LAUNCH();
  return COSMO_OK;
}

int cosmo_traverse_split(cosmo_ctx *ctx, const float *px, const float *py, const float *pz,
                         const void *nodes, int64_t n_tree, const int32_t *root_idx,
                         float *force, double theta, double G, double soft, double box,
                         int64_t off, int64_t count, int64_t n_bodies, uint64_t *stats,
                         void *stream) {
  (void)box;
  if (!ctx) return COSMO_E_INVALID;
  if (n_bodies < 0 || n_bodies > kMaxBodies || off < 0 || count < 0 || n_tree < 1 ||
      n_tree > kMaxBodies)
    return COSMO_E_INVALID;
  const int64_t lo = off;
  const int64_t hi = (off + count < n_bodies) ? off + count : n_bodies;
  if (hi <= lo) return COSMO_OK;
  if (!px || !py || !pz || !nodes || !root_idx || !force) return COSMO_E_INVALID;
  if (hi - lo > std::min<int64_t>(ctx->n_max, kSplitMaxCount)) return COSMO_E_TOO_LARGE;
  return split_walk(ctx, px, py, pz, nodes, n_tree, root_idx, force, theta, G, soft, lo, hi,
                    nullptr, 0, stats, as_stream(stream));
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
