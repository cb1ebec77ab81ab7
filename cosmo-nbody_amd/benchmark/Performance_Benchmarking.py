"""Per-kernel benchmark of the gravity-path drop-ins, laid out like the
reference harness (benchmark/Performance_Benchmarking.py:44-419) so that the
two CSV files line up column by column.

Same N sweep, warm-up/iteration counts, TPB, cache size, resolution, box,
softening, G, theta list and pie theta (:44-74); same uniform random FP32
positions with mass 1/N and zero velocities (:134-176); same kernels with the
same arguments, launched through the drop-ins with the reference's
``kern[blocks, threads](...)`` syntax (:184-351); same CSV header (:407-417).
Each figure is the mean over ITERATIONS of a HIP-event interval on the
launch stream (the reference's ``timed_kernel_GPU``, :95-117).

Deliberate differences, each printed in the CSV's companion columns:
- compute_multipoles: the reference never resets the arrival counters
  between runs (:244-255), so its later runs climb over stale counters; here
  the counters are cleared before each timed run, outside the interval.
- render_density: the grid is the fixed-point int64 grid of the drop-in,
  cleared inside the interval as the reference clears its float grid.
- compute_forces keeps the reference's 5000-particle batch loop (:281-296);
  the drop-in packs the node records from the reference-layout arrays on
  the first call for a tree and reuses them while the arrays are unchanged,
  and plans the split walk's frontiers once per tree.  Every timed loop
  starts a new tree generation, so it pays the pack and the plan once, as a
  simulation step does.  The extra columns ``traverse_theta_*`` time the
  step path: one pack and one launch over all particles.
- positions are drawn from a seeded generator (the reference's are
  unseeded).

Run from the package directory: ``python -m benchmark.Performance_Benchmarking
[--n-values 250000,2000000] [--thetas 0.5] [--out benchmark_results.csv]``.
"""
from __future__ import annotations

import argparse
import csv
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

N_VALUES = [10_000, 15_000, 25_000, 37_500, 50_000, 75_000, 100_000, 175_000, 250_000,
            375_000, 500_000, 750_000, 1_000_000, 1_500_000, 2_000_000, 3_000_000]
WARMUP = 5
ITERATIONS = 10
TPB = 256
CACHE_SIZE = 256
RES = 1024
BOX_SIZE = 1.0
SOFTENING = 0.001
G = 1.0
THETA_LIST = [0.3, 0.5, 0.7, 0.9, 1.1]
THETA_PIE = 0.7
BATCH = 5_000


def bpg(n: int, tpb: int = TPB) -> int:
    return (n + tpb - 1) // tpb


def timed_gpu(fn, before=None, warmup=WARMUP, iters=ITERATIONS) -> float:
    """Mean HIP-event interval (ms) of fn() over iters runs after warmup runs;
    before() runs ahead of each interval, outside it."""
    import torch
    for _ in range(warmup):
        if before:
            before()
        fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(iters):
        if before:
            before()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.mean(times))


class BenchState:
    """Buffers of one N, built once through the drop-ins (:131-224)."""

    def __init__(self, n: int, seed: int = 0):
        import torch
        from src.kernels.morton import compute_codes
        from src.kernels.tree_kernels import node_buffer, MAC_REF
        self.n = n
        self.n_nodes = 2 * n - 1
        dev = "cuda"
        pos = np.random.default_rng(seed).random((3, n)).astype(np.float32)
        self.pos_x, self.pos_y, self.pos_z = (torch.from_numpy(p).to(dev) for p in pos)
        self.mass = torch.full((n,), 1.0 / n, dtype=torch.float32, device=dev)
        self.vel_x, self.vel_y, self.vel_z = (torch.zeros(n, dtype=torch.float32, device=dev)
                                              for _ in range(3))
        self.force = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        self.codes = torch.zeros(n, dtype=torch.int64, device=dev)
        self.indices = torch.zeros(n, dtype=torch.int32, device=dev)
        M = self.n_nodes
        self.child = torch.full((M, 2), -1, dtype=torch.int32, device=dev)
        self.parent = torch.full((M,), -1, dtype=torch.int32, device=dev)
        self.root_idx = torch.zeros(1, dtype=torch.int32, device=dev)
        # float32 node arrays, as the reference allocates them (octree.py:37-38)
        self.n_mass = torch.zeros(M, dtype=torch.float32, device=dev)
        self.n_com = torch.zeros((M, 3), dtype=torch.float32, device=dev)
        self.n_min = torch.zeros((M, 3), dtype=torch.float32, device=dev)
        self.n_max = torch.zeros((M, 3), dtype=torch.float32, device=dev)
        self.counters = torch.zeros(M, dtype=torch.int32, device=dev)
        self.top_nodes = torch.full((CACHE_SIZE,), -1, dtype=torch.int32, device=dev)
        self.node_to_cache = torch.full((M,), -1, dtype=torch.int32, device=dev)
        self.grid = torch.zeros((RES, RES), dtype=torch.int64, device=dev)
        self.nodes = node_buffer(M, dev, MAC_REF)
        compute_codes[bpg(n), TPB](self.pos_x, self.pos_y, self.pos_z, self.codes, self.indices,
                                   n)
        order = torch.sort(self.codes, stable=True).indices
        self.pos_x, self.pos_y, self.pos_z = (p[order].contiguous() for p in
                                              (self.pos_x, self.pos_y, self.pos_z))
        self.mass = self.mass[order].contiguous()
        self.codes = self.codes[order].contiguous()
        self._build_tree()

    def _build_tree(self):
        from src.kernels.tree_kernels import (build_top_tree_cache, build_tree,
                                              compute_multipoles, find_root, normalize_com,
                                              reset_tree_buffers)
        import torch
        n, M = self.n, self.n_nodes
        reset_tree_buffers[bpg(M), TPB](self.child, self.parent, self.counters, self.root_idx)
        build_tree[bpg(n), TPB](self.codes, self.child, self.parent)
        compute_multipoles[bpg(n), TPB](self.pos_x, self.pos_y, self.pos_z, self.mass,
                                        self.child, self.parent, self.n_mass, self.n_com,
                                        self.n_min, self.n_max, self.counters)
        normalize_com[bpg(M, 256), 256](self.n_mass, self.n_com, M)
        find_root[bpg(M), TPB](self.parent, self.root_idx)
        self.node_to_cache.fill_(-1)
        build_top_tree_cache[1, 1](self.child, self.root_idx, self.top_nodes, self.node_to_cache,
                                   CACHE_SIZE)
        torch.cuda.synchronize()


def run_kernels(s: BenchState, thetas=None) -> dict:
    from src.kernels.morton import compute_codes
    from src.kernels._hip import bump_tree_generation
    from src.kernels.physics_kernels import compute_forces, integrate, traverse
    from src.kernels.render_kernels import render_density
    from src.kernels.tree_kernels import (MAC_REF, build_top_tree_cache, build_tree,
                                          compute_multipoles, find_root, normalize_com,
                                          pack_nodes, reset_tree_buffers)
    n, M = s.n, s.n_nodes
    out = {}
    out["compute_codes"] = timed_gpu(lambda: compute_codes[bpg(n), TPB](
        s.pos_x, s.pos_y, s.pos_z, s.codes, s.indices, n))
    # the topology buffers are rebuilt after the reset timing (the reset
    # clears them; build_tree leaves leaf children and the root's parent)
    out["reset_tree"] = timed_gpu(lambda: reset_tree_buffers[bpg(M), TPB](
        s.child, s.parent, s.counters, s.root_idx))
    out["build_tree"] = timed_gpu(lambda: build_tree[bpg(n), TPB](s.codes, s.child, s.parent))
    out["compute_multipoles"] = timed_gpu(
        lambda: compute_multipoles[bpg(n), TPB](s.pos_x, s.pos_y, s.pos_z, s.mass, s.child,
                                               s.parent, s.n_mass, s.n_com, s.n_min, s.n_max,
                                               s.counters),
        before=lambda: s.counters.zero_())
    out["normalize_com"] = timed_gpu(lambda: normalize_com[bpg(M, 256), 256](s.n_mass, s.n_com,
                                                                          M))
    s._build_tree()
    out["find_root"] = timed_gpu(lambda: find_root[bpg(M), TPB](s.parent, s.root_idx))

    def cache():
        s.node_to_cache.fill_(-1)
        build_top_tree_cache[1, 1](s.child, s.root_idx, s.top_nodes, s.node_to_cache, CACHE_SIZE)
    out["build_cache"] = timed_gpu(cache)
    out["integrate"] = timed_gpu(lambda: integrate[bpg(n), TPB](
        s.pos_x, s.pos_y, s.pos_z, s.vel_x, s.vel_y, s.vel_z, s.force, 0.5, 67.8, 1e-4))
    out["render_density"] = timed_gpu(lambda: (s.grid.zero_(), render_density[bpg(n), TPB](
        s.pos_x, s.pos_y, s.pos_z, s.grid, _bench_mvp(), RES)))
    # integrate moved the particles; rebuild so the traversals see a valid tree
    s._build_tree()
    thetas = THETA_LIST if thetas is None else thetas
    for theta in thetas:
        def batched(theta=theta):
            # a new tree generation: the loop packs the records and plans the
            # split walk's frontiers again, as each simulation step does
            bump_tree_generation()
            s.force.zero_()
            for off in range(0, n, BATCH):
                c = min(BATCH, n - off)
                compute_forces[bpg(c), TPB](s.pos_x, s.pos_y, s.pos_z, s.child, s.n_mass,
                                            s.n_com, s.n_min, s.n_max, s.force, theta, G,
                                            SOFTENING, BOX_SIZE, s.root_idx, s.top_nodes,
                                            s.node_to_cache, off)

        def single(theta=theta):
            pack_nodes(s.child, s.n_mass, s.n_com, s.n_min, s.n_max, BOX_SIZE, MAC_REF, s.nodes)
            traverse(s.pos_x, s.pos_y, s.pos_z, s.nodes, s.root_idx, s.force, theta, G,
                     SOFTENING, BOX_SIZE)
        out[f"compute_forces_theta_{theta}"] = timed_gpu(batched)
        out[f"traverse_theta_{theta}"] = timed_gpu(single)
    pie = THETA_PIE if THETA_PIE in thetas else thetas[0]
    out["compute_forces"] = out[f"compute_forces_theta_{pie}"]
    return out


_MVP = None


def _bench_mvp():
    """The harness's fixed camera (:303-321) in float32."""
    global _MVP
    if _MVP is None:
        eye = np.array([1.5, 0.2, 1.5], dtype=np.float32)
        target = np.array([0.5, 0.5, 0.5], dtype=np.float32)
        z = eye - target
        z /= np.linalg.norm(z)
        x = np.cross([0, 1, 0], z).astype(np.float32)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        view = np.eye(4, dtype=np.float32)
        view[0, :3] = x
        view[0, 3] = -np.dot(x, eye)
        view[1, :3] = y
        view[1, 3] = -np.dot(y, eye)
        view[2, :3] = z
        view[2, 3] = -np.dot(z, eye)
        f = 1.0 / np.tan(np.radians(60) / 2.0)
        proj = np.zeros((4, 4), dtype=np.float32)
        proj[0, 0] = f
        proj[1, 1] = f
        proj[2, 2] = -(100 + 0.1) / (100 - 0.1)
        proj[2, 3] = -2 * 100 * 0.1 / (100 - 0.1)
        proj[3, 2] = -1.0
        _MVP = np.ascontiguousarray(proj @ view, dtype=np.float32)
    return _MVP


KERNEL_COLUMNS = ["compute_codes", "reset_tree", "build_tree", "compute_multipoles",
                  "normalize_com", "find_root", "build_cache", "integrate", "render_density"]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n-values", default=None,
                    help="comma-separated N list (default: the reference's 16 sizes)")
    ap.add_argument("--out", default="benchmark_results.csv")
    ap.add_argument("--thetas", default=None,
                    help="comma-separated theta list for the force columns (default: the "
                         "reference's 0.3,0.5,0.7,0.9,1.1)")
    args = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    ns = N_VALUES if args.n_values is None else [int(x) for x in args.n_values.split(",")]
    thetas = THETA_LIST if args.thetas is None else [float(x) for x in args.thetas.split(",")]
    header = (["N"] + KERNEL_COLUMNS + ["compute_forces"]
              + [f"compute_forces_theta_{t}" for t in thetas]
              + [f"traverse_theta_{t}" for t in thetas])
    rows = []
    for n in ns:
        s = BenchState(n)
        r = run_kernels(s, thetas)
        rows.append([n] + [r[k] for k in header[1:]])
        print(f"N={n:,}: " + ", ".join(f"{k} {r[k]:.3f}" for k in header[1:]), flush=True)
        del s
        torch.cuda.empty_cache()
    with open(args.out, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)
    print(f"results saved to {args.out}")


if __name__ == "__main__":
    main()
