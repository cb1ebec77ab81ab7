/* ThreadSanitizer driver for the oracle's OpenMP loops (test infrastructure,
 * built and run by `make -C oracle tsan`).
 *
 * Runs every parallel entry point of cosmo_oracle.c on a seeded random
 * state -- codes, tree build, upward pass, traversal (both acceptance tests,
 * buckets, per-particle theta and straddle mass, counts, drops, targets),
 * direct sum, kick-drift -- once with one thread and once with several, and
 * requires bit-identical outputs: every parallel loop writes only its own
 * rows, so the thread count may not change a single bit.  Under
 * -fsanitize=thread with the OpenMP runtime's TSan annotations (LLVM
 * libomp + libarcher) any unsynchronised shared access is reported and the
 * process exits non-zero. */
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cosmo_oracle.h"

typedef struct {
  uint64_t *codes;
  uint32_t *perm;
  int32_t *child, *parent, *counts;
  double *nm64, *nc64;
  float *nm, *nc, *lo, *hi;
  float *f_ref, *f_aabb, *f_tg;
  int64_t *ni, *nv, *nd;
  double *acc;
  float *px, *py, *pz, *vx, *vy, *vz;
} Out;

static uint64_t rng_state = 0x853c49e6748fea9bull;
static double urand(void) { /* splitmix64 -> [0, 1) */
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

#define ALLOC(p, k) ((p) = calloc((size_t)(k), sizeof(*(p))))

static void run(const float *x0, const float *y0, const float *z0, const float *mass,
                const float *theta_g, const float *straddle_g, const int64_t *targets,
                int64_t n_targets, int64_t n, int threads, Out *o) {
  const int64_t M = 2 * n - 1;
  orc_set_num_threads(threads);
  ALLOC(o->codes, n); ALLOC(o->perm, n);
  ALLOC(o->child, 2 * M); ALLOC(o->parent, M); ALLOC(o->counts, M);
  ALLOC(o->nm64, M); ALLOC(o->nc64, 3 * M);
  ALLOC(o->nm, M); ALLOC(o->nc, 3 * M); ALLOC(o->lo, 3 * M); ALLOC(o->hi, 3 * M);
  ALLOC(o->f_ref, 3 * n); ALLOC(o->f_aabb, 3 * n); ALLOC(o->f_tg, 3 * n_targets);
  ALLOC(o->ni, n); ALLOC(o->nv, n); ALLOC(o->nd, n);
  ALLOC(o->acc, 3 * n_targets);
  float *p[6];
  for (int k = 0; k < 6; ++k) ALLOC(p[k], n);
  orc_compute_codes(x0, y0, z0, o->codes, NULL, n, ORC_MASK_FULL21);
  orc_argsort_stable(o->codes, o->perm, n);
  uint64_t *sc = malloc(sizeof(uint64_t) * n);
  float *ms = malloc(sizeof(float) * n);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t j = o->perm[i];
    sc[i] = o->codes[j];
    p[0][i] = x0[j]; p[1][i] = y0[j]; p[2][i] = z0[j];
    p[3][i] = 0.01f * (float)(j % 7); p[4][i] = -0.02f; p[5][i] = 0.005f;
    ms[i] = mass[j];
  }
  int32_t root_idx[1];
  orc_reset_tree_buffers(o->child, o->parent, o->counts, root_idx, M);
  orc_build_tree(sc, o->child, o->parent, n);
  if (orc_validate_tree(o->child, o->parent, n) != 0) {
    fprintf(stderr, "malformed tree\n");
    exit(2);
  }
  orc_compute_multipoles(p[0], p[1], p[2], ms, o->child, o->parent, o->nm64, o->nc64, o->lo,
                         o->hi, o->counts, n);
  orc_normalize_com(o->nm64, o->nc64, M);
  for (int64_t k = 0; k < M; ++k) o->nm[k] = (float)o->nm64[k];
  for (int64_t k = 0; k < 3 * M; ++k) o->nc[k] = (float)o->nc64[k];
  const int32_t root = (int32_t)n;
  orc_compute_forces(p[0], p[1], p[2], o->child, o->nm, o->nc, o->lo, o->hi, o->f_ref, 0.5,
                     1.0, 1e-3, 1.0, root, 0, n, n, 32, ORC_MAC_REF, theta_g, 0, 0.0, NULL,
                     NULL, o->ni, o->nv, o->nd, NULL);
  orc_compute_forces(p[0], p[1], p[2], o->child, o->nm, o->nc, o->lo, o->hi, o->f_aabb, 0.6,
                     1.0, 1e-3, 1.0, root, 0, n, n, 32, ORC_MAC_AABB, NULL, 16, 0.0,
                     straddle_g, NULL, NULL, NULL, NULL, NULL);
  orc_compute_forces(p[0], p[1], p[2], o->child, o->nm, o->nc, o->lo, o->hi, o->f_tg, 0.4, 1.0,
                     1e-3, 1.0, root, 0, 0, n, 64, ORC_MAC_AABB, NULL, 0, 1e-4, NULL, targets,
                     NULL, NULL, NULL, NULL);
  orc_direct_sum(p[0], p[1], p[2], ms, n, targets, n_targets, 1.0, 1e-3, o->acc);
  orc_kick_drift(p[0], p[1], p[2], p[3], p[4], p[5], o->f_ref, 0.99f, 0.01f, 0.01f, n);
  o->px = p[0]; o->py = p[1]; o->pz = p[2]; o->vx = p[3]; o->vy = p[4]; o->vz = p[5];
  free(sc);
  free(ms);
}

static int same(const char *name, const void *a, const void *b, size_t bytes) {
  if (memcmp(a, b, bytes) == 0) return 1;
  fprintf(stderr, "%s differs between thread counts\n", name);
  return 0;
}

int main(int argc, char **argv) {
  const int64_t n = argc > 1 ? atoll(argv[1]) : 20000;
  const int threads = argc > 2 ? atoi(argv[2]) : 8;
  const int64_t T = 257;
  float *x = malloc(sizeof(float) * n), *y = malloc(sizeof(float) * n),
        *z = malloc(sizeof(float) * n), *m = malloc(sizeof(float) * n),
        *tg = malloc(sizeof(float) * n), *sg = malloc(sizeof(float) * n);
  int64_t *targets = malloc(sizeof(int64_t) * T);
  for (int64_t i = 0; i < n; ++i) {
    /* a quarter of the points in a small clump: deep subtrees */
    const double s = (i % 4 == 0) ? 0.01 : 1.0;
    x[i] = (float)(0.3 + s * (urand() - 0.3));
    y[i] = (float)(0.6 + s * (urand() - 0.6));
    z[i] = (float)(s * urand());
    m[i] = (float)(1.0 / n);
    tg[i] = (float)(0.3 + 0.5 * urand());
    sg[i] = (float)(1e-4 * urand());
  }
  for (int64_t t = 0; t < T; ++t) targets[t] = (t * 7919) % n;
  Out a, b;
  run(x, y, z, m, tg, sg, targets, T, n, 1, &a);
  run(x, y, z, m, tg, sg, targets, T, n, threads, &b);
  const int64_t M = 2 * n - 1;
  int ok = 1;
  ok &= same("codes", a.codes, b.codes, sizeof(uint64_t) * n);
  ok &= same("perm", a.perm, b.perm, sizeof(uint32_t) * n);
  ok &= same("children", a.child, b.child, sizeof(int32_t) * 2 * M);
  ok &= same("parents", a.parent, b.parent, sizeof(int32_t) * M);
  ok &= same("node mass", a.nm64, b.nm64, sizeof(double) * M);
  ok &= same("node com", a.nc64, b.nc64, sizeof(double) * 3 * M);
  ok &= same("aabb min", a.lo, b.lo, sizeof(float) * 3 * M);
  ok &= same("aabb max", a.hi, b.hi, sizeof(float) * 3 * M);
  ok &= same("forces (reference test)", a.f_ref, b.f_ref, sizeof(float) * 3 * n);
  ok &= same("forces (aabb, buckets)", a.f_aabb, b.f_aabb, sizeof(float) * 3 * n);
  ok &= same("forces (targets)", a.f_tg, b.f_tg, sizeof(float) * 3 * T);
  ok &= same("interactions", a.ni, b.ni, sizeof(int64_t) * n);
  ok &= same("visits", a.nv, b.nv, sizeof(int64_t) * n);
  ok &= same("drops", a.nd, b.nd, sizeof(int64_t) * n);
  ok &= same("direct sum", a.acc, b.acc, sizeof(double) * 3 * T);
  ok &= same("positions", a.px, b.px, sizeof(float) * n);
  ok &= same("velocities", a.vz, b.vz, sizeof(float) * n);
  printf("tsan_check: n=%lld threads=1 vs %d: %s\n", (long long)n, threads,
         ok ? "bit-identical" : "DIFFERENT");
  return ok ? 0 : 1;
}
