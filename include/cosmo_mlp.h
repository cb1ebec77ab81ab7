/*
 * cosmo_mlp.h -- C ABI of the learned opening angle (part of
 * libcosmo_hip.so).  No reference counterpart: the reference uses one fixed
 * theta (src/utils/config.py:31, physics_kernels.py:87).
 *
 * A group is 64 Morton-consecutive particles (one wavefront of the
 * traversal).  For every group, cosmo_group_features computes nine
 * features; the MLP input row is [theta, features..., mu, 0-padding to 16]
 * for a candidate (theta, mu) of opening angle and straddle mass code.
 * The network 11 -> 64 -> 64 -> 1 (ReLU, BF16 weights, FP32 accumulation on
 * MFMA) predicts log2 of the group's mean squared acceleration error
 * relative to the mean squared acceleration of all particles.
 * cosmo_select_theta picks, per group, the candidate minimising
 * cost[k] + lambda * 2^prediction, with fixed values as fallback.
 *
 * Parameters: one FP32 vector of COSMO_MLP_PARAMS values laid out as
 * W1[16][64] (rows >= 11 zero), b1[64], W2[64][64], b2[64], W3[64], b3, and
 * a BF16 copy of W1|W2|W3 (COSMO_MLP_BF16_WEIGHTS values) read by the
 * kernels.  Weights are row-major [inputs][outputs].
 */
#ifndef COSMO_MLP_H
#define COSMO_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSMO_MLP_IN 11
#define COSMO_MLP_IN_PAD 16
#define COSMO_MLP_HIDDEN 64
#define COSMO_MLP_GROUP_FEATURES 9
#define COSMO_MLP_PARAMS (16 * 64 + 64 + 64 * 64 + 64 + 64 + 1)
#define COSMO_MLP_BF16_WEIGHTS (16 * 64 + 64 * 64 + 64)

#define COSMO_MLP_PROFILE_MAX_BINS 65536
#define COSMO_MLP_MAX_PROFILES 8

/* Straddle profile for the AABB test's straddle limit: the straddle-level
 * nodes are the internal nodes lighter than straddle_mass whose parent is
 * not; along each axis, bin b (of nbins over [0, 1)) holds the sums of
 * (M/straddle_mass)^2 and of M/straddle_mass over the straddle-level nodes
 * whose box spans the bin, as integers in units of 2^-32.  profile and
 * scratch: device uint64[2 * 3 * nbins] each, layout [sum][axis][bin]. */
int cosmo_straddle_profile(const int32_t *parents, const double *n_mass,
                           const float *n_min, const float *n_max, int64_t n,
                           double straddle_mass, uint64_t *profile,
                           uint64_t *scratch, int nbins, void *stream);

/* Per-group features feat[g][9] for groups g = 0..ceil(n/64)-1 from the
 * sorted positions and the tree (reference layouts; FP64 normalised CoM):
 *  0 log2(group extent * n_grid)        1 log2(relative number density)
 *  2 scale factor a
 *  3,5 extent/distance of the sibling of the group's node and of the next
 *      ancestor's sibling               4,6 log2 of their mass * n / cnt
 *  7,8 log2(1 + S2), log2(1 + S1): the straddle profile (nullable: 0)
 *      summed over the group's three minimum-image boundaries, the bins at
 *      the group centre + 1/2 in each axis. */
int cosmo_group_features(const float *pos_x, const float *pos_y,
                         const float *pos_z, const int32_t *children,
                         const int32_t *parents, const double *n_mass,
                         const double *n_com, const float *n_min,
                         const float *n_max, int64_t n, double a,
                         double n_grid, const uint64_t *profile, int nbins,
                         float *feat, void *stream);

/* cosmo_group_features with the scale factor read from device memory
 * (a[0]), for callers that replay a captured step. */
int cosmo_group_features_dev(const float *pos_x, const float *pos_y,
                             const float *pos_z, const int32_t *children,
                             const int32_t *parents, const double *n_mass,
                             const double *n_com, const float *n_min,
                             const float *n_max, int64_t n, const float *a,
                             double n_grid, const uint64_t *profile, int nbins,
                             float *feat, void *stream);

/* cosmo_straddle_profile for count (1..COSMO_MLP_MAX_PROFILES) straddle
 * masses at once (host array straddle_masses[count]), reading each node
 * once: profiles and scratch are device uint64[count][2 * 3 * nbins], block
 * j equal to cosmo_straddle_profile with straddle_masses[j]. */
int cosmo_straddle_profiles(const int32_t *parents, const double *n_mass,
                            const float *n_min, const float *n_max, int64_t n,
                            const double *straddle_masses, int count,
                            uint64_t *profiles, uint64_t *scratch, int nbins,
                            void *stream);

/* cosmo_group_features for count profiles (cosmo_straddle_profiles) at
 * once: feat is float[count][ceil(n/64)][9], block j equal to
 * cosmo_group_features with profile j (features 0-6 are the same in every
 * block).  a_dev (nullable): the scale factor read from device memory in
 * place of a, as cosmo_group_features_dev.  One wavefront per group. */
int cosmo_group_features_multi(const float *pos_x, const float *pos_y,
                               const float *pos_z, const int32_t *children,
                               const int32_t *parents, const double *n_mass,
                               const double *n_com, const float *n_min,
                               const float *n_max, int64_t n, double a,
                               const float *a_dev, double n_grid,
                               const uint64_t *profiles, int count, int nbins,
                               float *feat, void *stream);

/* X[g*K + k][0..15] = BF16 [thetas[k], F[g][0..8], cand2[k], 0...] with
 * F = feat (float[groups][9]), or, when fidx is given, block fidx[k] of feat
 * (float[blocks][groups][9]: the features computed with each candidate's
 * straddle profile).  cand2 nullable: 0. */
int cosmo_assemble_rows(const float *feat, int64_t groups,
                        const float *thetas, const float *cand2,
                        const int32_t *fidx, int K, uint16_t *X, void *stream);

/* Y[r] = MLP(X[r]) for rows r < rows; X is BF16 [rows][16]. */
int cosmo_mlp_forward(const uint16_t *X, int64_t rows, const uint16_t *wbf,
                      const float *params, float *Y, void *stream);

/* Gradient of mean((Y - T)^2) over `rows` samples: grad[COSMO_MLP_PARAMS]
 * plus grad[COSMO_MLP_PARAMS] = sum of squared errors.  partial is device
 * scratch of nblocks * (COSMO_MLP_PARAMS + 1) floats; blocks are reduced
 * in a fixed order. */
int cosmo_mlp_grad(const uint16_t *X, const float *T, int64_t rows,
                   const uint16_t *wbf, const float *params, float *partial,
                   int nblocks, float *grad, void *stream);

/* One Adam step (bias-corrected, step >= 1) on the FP32 parameters; also
 * refreshes the BF16 weights. */
int cosmo_mlp_adam(float *params, float *m, float *v, const float *grad,
                   float lr, float beta1, float beta2, float eps, int step,
                   uint16_t *wbf, void *stream);

/* BF16 copy of the weights of an FP32 parameter vector. */
int cosmo_mlp_to_bf16(const float *params, uint16_t *wbf, void *stream);

/* With k* = argmin_k cost[k] + lambda * 2^pred[g*K+k]: theta_group[g] =
 * thetas[k*] and, when mu_group is given, mu_group[g] = mus[k*] (the
 * candidates' straddle masses); when k_group is given, k_group[g] = k*.  A
 * group with any non-finite prediction gets `fallback` and `fallback_mu`
 * (k_group -1) and increments *n_fallback (nullable, device uint64). */
int cosmo_select_theta(const float *pred, int64_t groups, const float *thetas,
                       const float *mus, const float *cost, int K, double lambda,
                       double fallback, double fallback_mu, float *theta_group,
                       float *mu_group, int32_t *k_group, uint64_t *n_fallback,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif
