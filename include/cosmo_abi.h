/*
 * cosmo_abi.h -- C ABI of libcosmo_hip.so, the gfx950 (MI355X) kernels of
 * the per-step gravity path of a cosmological Barnes-Hut N-body code.
 *
 * Plain C: device pointers are ordinary pointers obtained from any HIP
 * allocator (torch tensors' data_ptr(), hipMalloc), sizes are int64_t and
 * `stream` is a hipStream_t passed as void* (NULL = the default stream).
 * Every call only enqueues work on `stream` and returns a status:
 * COSMO_OK, or an error whose text cosmo_status_string() gives.  Device
 * memory is allocated only by cosmo_ctx_create() and by the first
 * cosmo_zeldovich_ics() call per device and grid size (its hipFFT plan,
 * kept until the process ends); every other call allocates nothing.
 * Calls that share a cosmo_ctx must be issued on one stream (the context's
 * sort scratch, packed records and split-walk buffers are reused by every
 * call); use one context per stream for concurrent work.
 *
 * Each entry point names the reference (Cosmo-NBody) function it replaces.
 * The reference launches Numba kernels as kern[blocks, threads](args) on
 * CuPy arrays; the Python drop-ins in cosmo-nbody_amd/src/kernels keep those
 * names and call the functions below through ctypes.
 *
 * Layouts are the reference's: positions/velocities SoA float[n];
 * force float[n][3]; children int32[M][2]; node CoM/min/max [M][3];
 * M = 2n - 1 nodes, leaves 0..n-1, internal node i at n+i, root n.
 * Node mass and CoM are FP64 here (the reference stores FP32): the upward
 * pass accumulates in FP64 and the traversal rounds them to FP32.
 */
#ifndef COSMO_ABI_H
#define COSMO_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- statuses */
#define COSMO_OK 0
#define COSMO_E_INVALID 1        /* bad argument (null pointer, size)      */
#define COSMO_E_HIP 2            /* HIP runtime error                      */
#define COSMO_E_TOO_LARGE 3      /* n exceeds the context's capacity       */
#define COSMO_E_MALFORMED_TREE 4 /* tree validator found a defect          */
#define COSMO_E_FFT 5            /* hipFFT plan or execution failed        */

/* Morton mask sets: FULL21 = 21 bits per axis (63-bit code);
 * REF10 = the masks of the reference's spread_bits as written, which keep
 * each axis' low 10 bits (src/kernels/morton.py:33-36). */
#define COSMO_MASK_FULL21 0
#define COSMO_MASK_REF10 1

/* Acceptance tests of the traversal: REF = the reference's
 * (physics_kernels.py:128-133: x-extent of the AABB vs theta * distance to
 * the CoM); AABB = largest extent + CoM offset, and the node's box must
 * lie on one side of the minimum-image boundary in every axis. */
#define COSMO_MAC_REF 0
#define COSMO_MAC_AABB 1

/* traversal flags */
#define COSMO_TRAV_COMPENSATED 1 /* Neumaier-compensated accumulation     */
#define COSMO_TRAV_STATS 2       /* count interactions / node visits      */
#define COSMO_TRAV_PAIRS 4       /* two particles per lane (128-particle
                                    groups per wavefront)                 */
#define COSMO_TRAV_GROUP_STATS 8 /* STATS, plus the counts of each
                                    wavefront group                       */

/* Fixed-point fraction bits of the density bins used by the drop-ins. */
#define COSMO_RENDER_FRAC_BITS 20

/* Packed FP32 tree records read by the traversal.  One buffer of
 * cosmo_tree_bytes(n, mac) bytes holds, for the test `mac`:
 *   REF  2n records cosmo_inode_ref (32 B), record k for node id k:
 *          internal nodes (ids n..2n-2): sx = x-extent, replaced by box
 *          when > box/2 (physics_kernels.py:129-130), and sx2 = sx*sx
 *          rounded to FP32; leaves (ids 0..n-1, and any node lacking a
 *          child): CoM and mass with sx2 = -1, sx = 0 and c1 = c2 = 2n-1,
 *          the massless sentinel record 2n-1 (its own children), which a
 *          walk that opens a leaf (only on a non-finite distance) pops and
 *          skips;
 *   AABB n leaf records (cosmo_leaf, 16 B; node ids 0..n-1): CoM and mass,
 *        then n-1 internal records for node ids n..2n-2:
 *        cosmo_inode_aabb (64 B): s = largest extent of the node's box,
 *          dl = |CoM - box centre|, lo[k] = CoM_k - min_k, hi[k] = max_k - CoM_k,
 *          and, for a node of at most `bucket` leaves packed by
 *          cosmo_normalize_pack_buckets, its leaf range [first,
 *          first + count) (count = 0 otherwise).  A lane that opens such a
 *          node sums the leaves of the range directly, in leaf order, instead
 *          of walking the node's subtree.
 * Leaves are the node ids below n, as build_tree produces them (a leaf is
 * what the reference recognises by children == -1). */
typedef struct cosmo_leaf {
  float x, y, z, m;
} cosmo_leaf;

typedef struct cosmo_inode_ref {
  float cx, cy, cz, m;
  float sx2, sx;
  int32_t c1, c2;
} cosmo_inode_ref;

typedef struct cosmo_inode_aabb {
  float cx, cy, cz, m;
  float s, dl;
  int32_t c1, c2;
  float lo[3], hi[3];
  int32_t first, count;
} cosmo_inode_aabb;

/* Size in bytes of the packed records of an n-leaf tree for test `mac`:
 * REF: the 2n-1 records; AABB: the n leaf and n-1 internal records plus one
 * internal record of tail room (the traversal loads a whole internal-record
 * size at every node, leaves included, and keeps only the leaf's 16 bytes). */
int64_t cosmo_tree_bytes(int64_t n, int mac);

typedef struct cosmo_ctx cosmo_ctx;

const char *cosmo_status_string(int status);

/* Device-side scratch for the sort, the packed node records and the
 * validator, sized for up to n_max particles on the current device. */
int cosmo_ctx_create(int64_t n_max, cosmo_ctx **out);
int cosmo_ctx_destroy(cosmo_ctx *ctx);
/* Packed-record buffer inside the context (room for either test at n_max). */
void *cosmo_ctx_nodes(cosmo_ctx *ctx);

/* src/kernels/morton.py:65-91 compute_codes(pos_x, pos_y, pos_z, codes,
 * indices, n_bodies).  indices may be NULL. */
int cosmo_compute_codes(const float *pos_x, const float *pos_y,
                        const float *pos_z, uint64_t *codes, uint32_t *indices,
                        int64_t n_bodies, int mask_set, void *stream);

/* src/core/state.py:35 sort_idx = cp.argsort(codes): stable LSD radix sort
 * on the key (code, index).  codes_out = sorted codes, perm[k] = index of the
 * k-th smallest.  key_bits (1..64) = how many low bits of the code vary
 * (63 for FULL21, 30 for REF10).  codes_in may equal codes_out. */
int cosmo_sort_by_code(cosmo_ctx *ctx, const uint64_t *codes_in,
                       uint64_t *codes_out, uint32_t *perm, int64_t n,
                       int key_bits, void *stream);

/* cosmo_sort_by_code without a context and without any allocation (no
 * reference counterpart; for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is one static device array of 24.5 MiB per
 * device, held until the process ends; n above COSMO_SORT_NOALLOC_MAX
 * (2^21 = 128^3 codes) returns COSMO_E_TOO_LARGE.  Every call on a device
 * shares that array: calls that overlap on one device (other streams,
 * other host threads) race.  The first call on a device looks up the
 * array's address (hipGetSymbolAddress); make it outside graph capture. */
#define COSMO_SORT_NOALLOC_MAX (1 << 21)
int cosmo_sort_by_code_noalloc(const uint64_t *codes_in, uint64_t *codes_out,
                               uint32_t *perm, int64_t n, int key_bits,
                               void *stream);

/* src/core/state.py:37-40 the gathers x[sort_idx]: dst[a][k] = src[a][perm[k]]
 * for n_arrays (<= 8) float arrays; src/dst are host arrays of device
 * pointers; dst must not alias src. */
int cosmo_permute(const uint32_t *perm, int n_arrays, const float *const *src,
                  float *const *dst, int64_t n, void *stream);

/* src/kernels/tree_kernels.py:236-262 reset_tree_buffers(child, parent,
 * counters, root_idx). */
int cosmo_reset_tree_buffers(int32_t *child, int32_t *parent,
                             int32_t *counters, int32_t *root_idx,
                             int64_t n_nodes, void *stream);

/* src/kernels/tree_kernels.py:60-129 build_tree(codes, children, parents). */
int cosmo_build_tree(const uint64_t *codes, int32_t *children,
                     int32_t *parents, int64_t n, void *stream);

/* cosmo_build_tree that also writes far[i] (int32[n-1]) = j, the other end
 * of internal node n+i's leaf range [min(i,j), max(i,j)] (no reference
 * counterpart: the leaf buckets of cosmo_normalize_pack_buckets read it). */
int cosmo_build_tree_spans(const uint64_t *codes, int32_t *children,
                           int32_t *parents, int32_t *far, int64_t n,
                           void *stream);

/* src/kernels/tree_kernels.py:131-191 compute_multipoles(pos_x, pos_y, pos_z,
 * mass, children, parents, n_mass, n_com, n_min, n_max, counts); FP64 mass
 * and mass-weighted CoM sums.  counts must be zero (reset_tree_buffers). */
int cosmo_compute_multipoles(const float *pos_x, const float *pos_y,
                             const float *pos_z, const float *mass,
                             const int32_t *children, const int32_t *parents,
                             double *n_mass, double *n_com, float *n_min,
                             float *n_max, int32_t *counts, int64_t n,
                             void *stream);

/* compute_multipoles on the reference's own float32 node arrays
 * (octree.py:37-38): the same FP64 sums, rounded once to FP32 on store.
 * The FP64 sums are formed in the context's packed-record buffer, so records
 * packed there before (cosmo_ctx_nodes) are overwritten.  Requires n <= the
 * context's capacity. */
int cosmo_compute_multipoles_f32(cosmo_ctx *ctx, const float *pos_x, const float *pos_y,
                                 const float *pos_z, const float *mass,
                                 const int32_t *children, const int32_t *parents,
                                 float *n_mass, float *n_com, float *n_min,
                                 float *n_max, int32_t *counts, int64_t n,
                                 void *stream);

/* src/kernels/tree_kernels.py:193-214 normalize_com(n_mass, n_com, n_nodes)
 * on the reference's float32 arrays, in the reference's arithmetic:
 * inv = 1.0 / m in FP64, com = fl32(com * inv). */
int cosmo_normalize_com_f32(const float *n_mass, float *n_com, int64_t n_nodes,
                            void *stream);

/* src/kernels/tree_kernels.py:193-214 normalize_com(n_mass, n_com, n_nodes). */
int cosmo_normalize_com(const double *n_mass, double *n_com, int64_t n_nodes,
                        void *stream);

/* src/kernels/tree_kernels.py:216-234 find_root(parents, root_idx): the
 * largest parentless index (the unique one in a well-formed tree); root_idx
 * must hold 0 on entry (reset_tree_buffers). */
int cosmo_find_root(const int32_t *parents, int32_t *root_idx,
                    int64_t n_nodes, void *stream);

/* src/kernels/tree_kernels.py:264-314 build_top_tree_cache(child, root_idx,
 * top_nodes, node_to_cache, cache_size); node_to_cache pre-filled with -1. */
int cosmo_build_top_tree_cache(const int32_t *child, const int32_t *root_idx,
                               int32_t *top_nodes, int32_t *node_to_cache,
                               int64_t cache_size, void *stream);

/* Packs the reference-layout node arrays (n_nodes = 2n-1) into the records
 * for the acceptance test `mac` (no reference counterpart: the reference
 * reads its arrays directly, physics_kernels.py:96-113). */
int cosmo_pack_nodes(const int32_t *children, const double *n_mass,
                     const double *n_com, const float *n_min,
                     const float *n_max, int64_t n_nodes, double box, int mac,
                     void *nodes, void *stream);

/* cosmo_pack_nodes from the reference's float32 node arrays (mass, CoM
 * already normalised by normalize_com). */
int cosmo_pack_nodes_f32(const int32_t *children, const float *n_mass,
                         const float *n_com, const float *n_min,
                         const float *n_max, int64_t n_nodes, double box,
                         int mac, void *nodes, void *stream);

/* normalize_com and cosmo_pack_nodes fused in one pass. */
int cosmo_normalize_pack(const int32_t *children, const double *n_mass,
                         double *n_com, const float *n_min, const float *n_max,
                         int64_t n_nodes, double box, int mac,
                         void *nodes, void *stream);

/* cosmo_normalize_pack for the AABB test with leaf buckets (no reference
 * counterpart): internal nodes of at most `bucket` (1..64; 0 = none) leaves
 * carry their leaf range, read from far (cosmo_build_tree_spans). */
int cosmo_normalize_pack_buckets(const int32_t *children, const int32_t *far,
                                 const double *n_mass, double *n_com,
                                 const float *n_min, const float *n_max,
                                 int64_t n_nodes, double box, int mac,
                                 int bucket, void *nodes, void *stream);

/* Barnes-Hut traversal over the packed records of an n_tree-leaf tree for
 * the particles [off, off+count) ∩ [0, n_bodies) of pos_* (the tree's own
 * particles, or any other targets) in Morton-contiguous groups of 64 (one
 * wavefront each),
 * writing force[i][0..2] (accelerations).  theta_group (nullable, device
 * float[ceil(count/64)]) gives each group's opening angle in place of
 * theta.  stats (nullable, device uint64[4]) accumulates interactions,
 * node visits (per particle), stack pops (per wavefront) and overflow drops
 * (per particle: nodes it would have opened while its stack held 63
 * entries, whose children the reference drops, physics_kernels.py:143) when
 * flags has COSMO_TRAV_STATS.  With COSMO_TRAV_GROUP_STATS stats holds
 * 4 + 3*W entries for W = ceil(count / (64 * particles per lane))
 * wavefront groups: the totals accumulate in stats[0..3] and group w's own
 * interactions, visits and pops are written to stats[4+3w .. 4+3w+2].
 * The walk addresses records with 32-bit byte offsets: a tree whose
 * records (cosmo_tree_bytes) pass 4 GiB returns COSMO_E_TOO_LARGE (about
 * 2^26 leaves for the reference test, 5.4e7 for the AABB test). */
int cosmo_traverse(const float *pos_x, const float *pos_y, const float *pos_z,
                   const void *nodes, int64_t n_tree, const int32_t *root_idx,
                   float *force, double theta, const float *theta_group,
                   double G, double soft, double box, int64_t off,
                   int64_t count, int64_t n_bodies, int mac, int flags,
                   uint64_t *stats, void *stream);

/* The same walk with the reference test for a launch of at most
 * min(ctx capacity, 65536) particles (no reference counterpart; what
 * compute_forces uses for the 5000-particle batches of engine.py:91-111):
 * each group of 64 particles gets one wavefront per subtree below a small
 * top tree, which decides the top-tree nodes on its path exactly as the
 * walk does and walks its subtree from the stack depth the walk reaches it
 * at.  Each particle visits and accepts the same nodes as with
 * cosmo_traverse; its subtree sums are added in subtree order.  nodes: the
 * reference test's packed records (they already hold the extents clamped to
 * the box, so box is not read; it is kept for the argument order of
 * cosmo_traverse).  stats (nullable, device uint64[2]):
 * interactions and node visits are added.  Returns COSMO_E_TOO_LARGE for
 * larger launches. */
int cosmo_traverse_split(cosmo_ctx *ctx, const float *pos_x, const float *pos_y,
                         const float *pos_z, const void *nodes, int64_t n_tree,
                         const int32_t *root_idx, float *force, double theta, double G,
                         double soft, double box, int64_t off, int64_t count,
                         int64_t n_bodies, uint64_t *stats, void *stream);

/* cosmo_traverse with the wavefront groups launched in the order
 * group_order[0..W) (a permutation of the W = ceil(count/64) groups; no
 * reference counterpart).  Results do not depend on the order; a good
 * order starts the most expensive groups first.  Not with
 * COSMO_TRAV_PAIRS. */
int cosmo_traverse_ordered(const float *pos_x, const float *pos_y,
                           const float *pos_z, const void *nodes, int64_t n_tree,
                           const int32_t *root_idx, float *force, double theta,
                           const float *theta_group, double G, double soft,
                           double box, int64_t off, int64_t count,
                           int64_t n_bodies, int mac, int flags, uint64_t *stats,
                           const int32_t *group_order, void *stream);

/* cosmo_traverse with every option (no reference counterpart):
 * group_order (nullable) as for cosmo_traverse_ordered, and, for the AABB
 * test, straddle_mass >= 0: a node whose box crosses a particle's
 * minimum-image boundary still passes the test when its FP32 mass is below
 * fl(straddle_mass) (0 = the exact guard of cosmo_traverse).  Such a node
 * puts at most its mass on the wrong image, so the limit bounds the error
 * it adds.  straddle_group (nullable, device float[ceil(count/64)], AABB
 * test, not with COSMO_TRAV_PAIRS) gives each group its own straddle mass
 * in place of straddle_mass.  group_pops (nullable, device
 * uint32[ceil(count/64)], AABB test, not with COSMO_TRAV_PAIRS) receives
 * each group's number of stack pops, the cost measure of
 * cosmo_order_groups_by_pops. */
int cosmo_traverse_ex(const float *pos_x, const float *pos_y, const float *pos_z,
                      const void *nodes, int64_t n_tree, const int32_t *root_idx,
                      float *force, double theta, const float *theta_group,
                      double G, double soft, double box, int64_t off,
                      int64_t count, int64_t n_bodies, int mac, int flags,
                      uint64_t *stats, const int32_t *group_order,
                      double straddle_mass, const float *straddle_group,
                      uint32_t *group_pops, void *stream);

/* order[0..groups) = the groups sorted by increasing theta_group (64
 * logarithmic bins; arbitrary order inside a bin): the launch order for
 * cosmo_traverse_ordered with per-group theta. */
int cosmo_order_groups(const float *theta_group, int64_t groups, int32_t *order,
                       void *stream);

/* order[0..groups) = the groups sorted by decreasing group_pops (the pop
 * counts an earlier cosmo_traverse_ex stored; 64 bins of a third of an
 * octave; arbitrary order inside a bin): a longest-first launch order for
 * the next launch over the same groups (no reference counterpart).  With a
 * per-group selection (cosmo_select_theta's k_group) k_now and k_prev
 * (nullable together) give each group's candidate now and at the launch
 * that counted the pops; a group's pops are then scaled by cost[k_now] /
 * cost[k_prev] (cost: device float[K], both indices in [0, K)). */
int cosmo_order_groups_by_pops(const uint32_t *group_pops, int64_t groups,
                               const int32_t *k_now, const int32_t *k_prev, const float *cost,
                               int K, int32_t *order, void *stream);

/* src/kernels/physics_kernels.py:14-153 compute_forces(pos_x, pos_y, pos_z,
 * children, n_mass, n_com, n_min, n_max, force, theta, G, soft, box, root,
 * top_nodes, node_to_cache, off) for particles [off, off+count): packs the
 * nodes into the context and runs cosmo_traverse with the reference test.
 * top_nodes/node_to_cache are accepted for signature compatibility; the
 * reference only uses them as an on-chip copy of the same node data. */
int cosmo_compute_forces(cosmo_ctx *ctx, const float *pos_x,
                         const float *pos_y, const float *pos_z,
                         const int32_t *children, const double *n_mass,
                         const double *n_com, const float *n_min,
                         const float *n_max, float *force, double theta,
                         double G, double soft, double box,
                         const int32_t *root_idx, const int32_t *top_nodes,
                         const int32_t *node_to_cache, int64_t off,
                         int64_t count, int64_t n_bodies, void *stream);

/* cosmo_compute_forces on the reference's float32 node arrays
 * (octree.py:37-40; n_mass, n_com as normalize_com leaves them). */
int cosmo_compute_forces_f32(cosmo_ctx *ctx, const float *pos_x,
                             const float *pos_y, const float *pos_z,
                             const int32_t *children, const float *n_mass,
                             const float *n_com, const float *n_min,
                             const float *n_max, float *force, double theta,
                             double G, double soft, double box,
                             const int32_t *root_idx, const int32_t *top_nodes,
                             const int32_t *node_to_cache, int64_t off,
                             int64_t count, int64_t n_bodies, void *stream);

/* The traversal half of compute_forces for a batch loop that packs once:
 * the same walk as cosmo_compute_forces over the reference-test records
 * already in cosmo_ctx_nodes(ctx) (written by cosmo_pack_nodes or
 * cosmo_pack_nodes_f32 with COSMO_MAC_REF for this n_bodies-leaf tree), for
 * particles [off, off+count).  engine.py:91-111 calls compute_forces once
 * per 5000-particle batch on one tree; packing once per tree and calling
 * this per batch gives the same forces.
 * epoch: 0, or a value the caller changes whenever the records in the
 * context or the root value change (e.g. a count of packs).  With a nonzero
 * epoch the first call builds the split walk's frontiers for every window
 * of the tree and later calls under the same epoch, root_idx pointer and
 * n_bodies reuse them (one launch less per batch).  Forces are those of
 * the walk in every case; only the association of the per-subtree partial
 * sums follows the frontier. */
int cosmo_compute_forces_packed(cosmo_ctx *ctx, const float *pos_x,
                                const float *pos_y, const float *pos_z,
                                float *force, double theta, double G,
                                double soft, double box,
                                const int32_t *root_idx, int64_t off,
                                int64_t count, int64_t n_bodies, uint64_t epoch,
                                void *stream);

/* src/kernels/physics_kernels.py:155-203 integrate(pos_x, pos_y, pos_z,
 * vel_x, vel_y, vel_z, force, a, H, dt): c_drag = 1 - H*dt and
 * c_kick = c_drift = dt/a are formed in FP64 and rounded to FP32, then
 * cosmo_kick_drift runs. */
int cosmo_integrate(float *pos_x, float *pos_y, float *pos_z, float *vel_x,
                    float *vel_y, float *vel_z, const float *force, double a,
                    double H, double dt, int64_t n, void *stream);

/* v = v*c_drag + F*c_kick;  x = x + v*c_drift;  x -= floor(x)  (FP32). */
int cosmo_kick_drift(float *pos_x, float *pos_y, float *pos_z, float *vel_x,
                     float *vel_y, float *vel_z, const float *force,
                     float c_drag, float c_kick, float c_drift, int64_t n,
                     void *stream);

/* cosmo_kick_drift with the three coefficients read from device memory
 * (coefs[0..2] = c_drag, c_kick, c_drift), so that a captured HIP graph can
 * be replayed with new coefficients written before each launch. */
int cosmo_kick_drift_dev(float *pos_x, float *pos_y, float *pos_z,
                         float *vel_x, float *vel_y, float *vel_z,
                         const float *force, const float *coefs, int64_t n,
                         void *stream);

/* src/kernels/render_kernels.py:12-47 render_density(pos_x, pos_y, pos_z,
 * grid, mvp, res) into fixed-point bins: grid[res][res] uint64 gains
 * rint(2.5/w * 2^frac_bits) per visible particle.  mvp: host float[16],
 * row-major.  The caller clears grid. */
int cosmo_render_density(const float *pos_x, const float *pos_y,
                         const float *pos_z, uint64_t *grid, const float *mvp,
                         int32_t res, int64_t n, int frac_bits, void *stream);

/* The float32 density grid of the reference (camera.py:26) from fixed-point
 * bins: grid[p] += fl32(bins[p] * 2^-frac_bits) for p < cells.  The
 * render_density drop-in splats into bins with cosmo_render_density and
 * adds them to a float32 grid with this, so the float grid does not depend
 * on arrival order either. */
int cosmo_grid_add_fixed(const uint64_t *bins, float *grid, int64_t cells,
                         int frac_bits, void *stream);

/* cosmo_render_density with the MVP matrix read from device memory
 * (mvp_dev: float[16], row-major), for captured HIP graphs. */
int cosmo_render_density_dev(const float *pos_x, const float *pos_y,
                             const float *pos_z, uint64_t *grid,
                             const float *mvp_dev, int32_t res, int64_t n,
                             int frac_bits, void *stream);

/* src/utils/visualizer.py:115-171 colorize_frame(grid_d), fused after the
 * splat: img = log1p(120 x)/(max img + 1e-9), val = img^2.2, the four-band
 * palette, x1.3, clip to [0, 1], with x = grid * 2^-frac_bits in FP32.
 * frame (nullable): float[res][res][3]; rgb (nullable): the bytes
 * (frame * 255).astype(uint8) that engine.py:139 stores.  max_scratch:
 * device uint64[1] (the image maximum; overwritten). */
int cosmo_colorize_frame(const uint64_t *grid, int32_t res, int frac_bits,
                         uint64_t *max_scratch, float *frame, uint8_t *rgb,
                         void *stream);

/* Failure detection (no reference counterpart; the reference has no checks):
 * count[0] += the number of non-finite values (NaN, +-inf) among v[0..n),
 * e.g. the 3n components of the accelerations after a traversal.  count:
 * device uint64[1]. */
int cosmo_count_nonfinite(const float *v, int64_t n, uint64_t *count, void *stream);

/* FP64 direct sum (no reference counterpart; the error yardstick): for each
 * target t, acc[t][k] = sum_j G m_j (d^2+eps^2)^-3/2 d_k over all n sources
 * with the traversal's minimum image and d^2 > 1e-7 cut. */
int cosmo_direct_sum(const float *pos_x, const float *pos_y,
                     const float *pos_z, const float *mass, int64_t n,
                     const int64_t *targets, int64_t n_targets, double G,
                     double soft, double *acc, void *stream);

/* src/utils/cosmology.py:12-73 generate_zeldovich_ics after the phase draw
 * (:35), on the device: phase is double[n^3] in [0, 2 pi), flat index
 * (i n + j) n + l with i along x (numpy's ij meshgrid order).  P(k) = k^-2.5
 * with k = fftfreq(n) n, displacement d_c = Re ifftn(i k_c / (k^2 + 1e-10)
 * sqrt(P) e^{i phase}) (three in-place double-complex inverse FFTs through
 * hipFFT), scaled by (1/n) amplitude / std(d_x); positions = (lattice + d)
 * mod 1 and velocities = d * coupling, rounded to float.  work: device
 * memory of cosmo_zeldovich_work_bytes(n_grid) bytes (the FFT buffer, the
 * three components and the reduction scratch), overwritten.  n_grid in
 * [1, 1024].  Returns COSMO_E_FFT if hipFFT fails. */
int64_t cosmo_zeldovich_work_bytes(int64_t n_grid);
int cosmo_zeldovich_ics(int64_t n_grid, const double *phase, double amplitude,
                        double coupling, void *work, float *pos_x, float *pos_y,
                        float *pos_z, float *vel_x, float *vel_y, float *vel_z,
                        void *stream);

/* Tree validator (no reference counterpart): writes a bit set to
 * flags_out (device int32[1]): 1 = parentless-node count != 1,
 * 2 = root != n, 4 = inconsistent child/parent link, 8 = not every leaf
 * reachable from the root exactly once. */
int cosmo_validate_tree(cosmo_ctx *ctx, const int32_t *children,
                        const int32_t *parents, int64_t n, int32_t *flags_out,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif
