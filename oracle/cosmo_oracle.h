/*
 * cosmo_oracle.h -- CPU oracle for the per-step gravity path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the product path
 * (cosmo-nbody_amd/) never does.  It is the checker the HIP kernels are
 * compared against, never the thing measured or shipped.
 *
 * Every function restates one stage of the reference (Cosmo-NBody,
 * src/kernels and src/core) on host pointers.  The reference itself
 * cannot run here (Numba/CuPy absent), so each function is pinned by the
 * checks in tests/test_oracle_*.py: closed forms, brute force on tiny
 * inputs and invariants.  Functions whose exact values no independent check
 * fixes say "parity unpinned" below.
 */
#ifndef COSMO_ORACLE_H
#define COSMO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Morton mask sets.  FULL21 = 21 bits per axis (the documented 63-bit code,
 * README.md:44, notebooks/final.ipynb:428-434).  REF10 = the masks literally
 * written in src/kernels/morton.py:33-36, which keep each axis' low 10 bits. */
#define ORC_MASK_FULL21 0
#define ORC_MASK_REF10 1

/* Acceptance tests for compute_forces (see cosmo_oracle.c). */
#define ORC_MAC_REF 0
#define ORC_MAC_AABB 1

/* src/kernels/morton.py:40-91  (compute_codes). indices may be NULL. */
void orc_compute_codes(const float *px, const float *py, const float *pz,
                       uint64_t *codes, uint32_t *indices, int64_t n,
                       int mask_set);

/* src/core/state.py:35  cp.argsort(codes), taken as a stable sort, i.e. a
 * sort on the key (code, index).  perm[k] = index of the k-th smallest. */
void orc_argsort_stable(const uint64_t *codes, uint32_t *perm, int64_t n);

/* src/kernels/tree_kernels.py:236-262 */
void orc_reset_tree_buffers(int32_t *child, int32_t *parent, int32_t *counters,
                            int32_t *root_idx, int64_t n_nodes);

/* src/kernels/tree_kernels.py:17-129 (clz64, delta_fn, build_tree) */
void orc_build_tree(const uint64_t *codes, int32_t *children, int32_t *parents,
                    int64_t n);

/* src/kernels/tree_kernels.py:131-191, accumulated in FP64 in the fixed
 * left+right order of :175-187.  n_mass[M], n_com[M*3] (mass-weighted sums),
 * n_min/n_max[M*3] FP32.  counts[M] must be zero on entry. */
void orc_compute_multipoles(const float *px, const float *py, const float *pz,
                            const float *mass, const int32_t *children,
                            const int32_t *parents, double *n_mass,
                            double *n_com, float *n_min, float *n_max,
                            int32_t *counts, int64_t n);

/* src/kernels/tree_kernels.py:193-214 (FP64) */
void orc_normalize_com(const double *n_mass, double *n_com, int64_t n_nodes);

/* src/kernels/tree_kernels.py:216-234.  The reference keeps whichever
 * parentless node wins an atomic exchange; this restatement keeps the
 * largest parentless index (0 if none), which is the same node whenever the
 * tree is well formed (exactly one parentless node). */
void orc_find_root(const int32_t *parents, int32_t *root_idx, int64_t n_nodes);

/* src/kernels/tree_kernels.py:264-314.  node_to_cache must be pre-filled
 * with -1 by the caller (src/core/engine.py:81). */
void orc_build_top_tree_cache(const int32_t *child, const int32_t *root_idx,
                              int32_t *top_nodes, int32_t *node_to_cache,
                              int64_t cache_size);

/* Tree validator (no reference counterpart; the reference's equal-code
 * tie-break can produce malformed trees).  Returns a bit set:
 *   1 = number of parentless nodes != 1
 *   2 = root is not node n (n >= 2)
 *   4 = a child/parent link is inconsistent (child claimed twice, bad id)
 *   8 = not every leaf is reachable exactly once from the root (cycle)   */
int orc_validate_tree(const int32_t *children, const int32_t *parents,
                      int64_t n);

/* src/kernels/physics_kernels.py:14-153 (compute_forces) for particles
 * [off, off+count) ∩ [0, n).  Node data in the reference layout, FP32.
 * precision = 64: arithmetic in FP64 on the FP32 inputs (the reference's
 *                 Numba promotion, physics_kernels.py:118-141);
 * precision = 32: the FP32 arithmetic the HIP traversal uses (explicit fmaf
 *                 for d2 and the accumulation, IEEE sqrt and divide).
 * mac: ORC_MAC_REF (the reference's test) or ORC_MAC_AABB (corrected).
 * theta_g: optional per-particle opening angle (NULL = use theta).
 * straddle_mass: ORC_MAC_AABB only: a node whose box crosses the particle's
 *   minimum-image boundary is still accepted (if the size test passes) when
 *   its FP32 mass is below fl(straddle_mass); 0 = never.
 * bucket: ORC_MAC_AABB only (0 = off): a rejected internal node of at most
 *   `bucket` leaves is not opened; its leaves are visited in leaf order
 *   (no reference counterpart).
 * straddle_g: optional per-particle straddle mass (in place of straddle_mass).
 * targets: optional list of count particle indices (off is then ignored);
 *   output row t (force, counts, force64) belongs to particle targets[t].
 * n_inter / n_visit: optional per-particle counts of accepted interactions
 * (d2 > 1e-7) and of node visits (MAC evaluations and bucket leaves).
 * n_drop: optional per-particle count of rejected internal nodes whose
 * children the full 64-entry stack drops (physics_kernels.py:143).       */
void orc_compute_forces(const float *px, const float *py, const float *pz,
                        const int32_t *children, const float *n_mass,
                        const float *n_com, const float *n_min,
                        const float *n_max, float *force, double theta,
                        double G, double soft, double box, int32_t root,
                        int64_t off, int64_t count, int64_t n, int precision,
                        int mac, const float *theta_g, int bucket,
                        double straddle_mass, const float *straddle_g,
                        const int64_t *targets, int64_t *n_inter, int64_t *n_visit,
                        int64_t *n_drop, double *force64);

/* FP64 direct sum over all n sources for the listed targets, with the same
 * minimum image, d2 > 1e-7 cut and Plummer law as compute_forces.  Output
 * acc[t*3+k] in FP64.  No reference counterpart (error yardstick). */
void orc_direct_sum(const float *px, const float *py, const float *pz,
                    const float *mass, int64_t n, const int64_t *targets,
                    int64_t n_targets, double G, double soft, double *acc);

/* src/kernels/physics_kernels.py:155-203 as a coefficient kick-drift in FP32:
 * v = v*c_drag + F*c_kick;  x = x + v*c_drift;  x -= floor(x).          */
void orc_kick_drift(float *px, float *py, float *pz, float *vx, float *vy,
                    float *vz, const float *force, float c_drag, float c_kick,
                    float c_drift, int64_t n);

/* The coefficients the integrate() entry point derives from (a, H, dt):
 * c_drag = 1 - H*dt, c_kick = c_drift = dt/a, computed in FP64 and rounded
 * to FP32 (physics_kernels.py:180-195). */
void orc_integrate_coeffs(double a, double H, double dt, float *c_drag,
                          float *c_kick, float *c_drift);

/* src/kernels/render_kernels.py:12-47 with the float atomic replaced by
 * fixed-point u64 bins: q = rint(2.5/w * 2^frac_bits). mvp row-major 4x4. */
void orc_render_density(const float *px, const float *py, const float *pz,
                        uint64_t *grid, const float *mvp, int32_t res,
                        int64_t n, int frac_bits);

int orc_num_threads(void);
void orc_set_num_threads(int t);

#ifdef __cplusplus
}
#endif
#endif
