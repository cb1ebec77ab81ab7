"""Throughput of the per-step gravity path on one MI355X.

Workload (default): N = 128^3 Zel'dovich ICs (seed 42), theta = 0.5,
eps = 0.02/128, reference time schedule; one step = Morton encode + radix
sort + permute, tree build, FP64 upward pass + node packing, Barnes-Hut
traversal (reference acceptance test, FP32), kick-drift update, and the
1024^2 fixed-point density splat.  W untimed steps, then K timed steps
between device synchronisations.  Prints one JSON line.  It also carries
the learned-theta half of the metric ("learned_vs_fixed"): the full-step
rate of the learned per-group opening angle against the fixed opening
angle at equal 1e-3 relative RMS force error (FP64 direct sum), with the
network trained on the GPU in the same run.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cosmo-nbody_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 published peak (packed FP32 / dense MFMA FP32)
# FP32 VALU without packed instructions (256 CUs x 64 lanes x 2 flop per FMA x
# 2.4 GHz): the ceiling of the traversal, which is built without packed FP32
PEAK_VALU_FP32_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0         # MI355X HBM3E
FLOP_PER_INTERACTION = 20     # standard particle-node interaction count
# HBM bytes per traversal launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (scaled by a 256 MiB copy: x2.0 and x1.0), per workload; the
# summary with its commands is profiles/step_kernels_hbm_traffic_128.md.
TRAVERSAL_TRAFFIC_BYTES = {"repo_128": 332581 * 1024 * 2 + 24576 * 1024}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--preset", default="repo_128")
    ap.add_argument("--theta", type=float, default=None,
                    help="opening angle in place of the preset's")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--pairs", action="store_true",
                    help="128 particles per wavefront in the traversal instead of 64")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="short run for hardware counters: 3 traversal launches on a warmed "
                         "state plus a 256 MiB device copy (counter scale)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the whole step from HIP graphs (per-stage times not split)")
    ap.add_argument("--learned", action="store_true",
                    help="learned-theta vs fixed-theta step rate at equal 1e-3 relative RMS "
                         "force error (FP64 direct sum) on the 128^3 halo state: trains the "
                         "network on three training states, evaluates on a held-out one")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-theta vs fixed-theta comparison of the default line")
    ap.add_argument("--error-targets", type=int, default=0,
                    help="also report the relative RMS force error of the final state "
                         "against the FP64 direct sum on this many sampled particles")
    ap.add_argument("--shards", type=int, default=8,
                    help="also time the traversal of each of this many contiguous Morton "
                         "slices of the targets against the whole tree (0: skip)")
    return ap.parse_args()


def shard_slices(eng, c, parts, repeats=3):
    """Per-slice traversal time of a step sharded over `parts` GPUs with a
    replicated tree: the targets are cut into `parts` contiguous Morton ranges
    (multiples of 64 particles) and each range is traversed against the whole
    tree, as one GPU of such a step would.  HIP events on the stream the
    traversal is launched on; the tree is the one the timed steps ended on."""
    import torch

    from src.kernels.physics_kernels import traverse
    s = eng.state
    n = s.n
    cuts = [min(n, ((k * n) // parts + 63) // 64 * 64) for k in range(parts)] + [n]
    ms = []
    for k in range(parts):
        off, cnt = cuts[k], cuts[k + 1] - cuts[k]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            traverse(*s.pos, eng.tree_manager.nodes, eng.tree_manager.root_idx, s.force,
                     c.THETA, c.G, c.SOFTENING, c.BOX_SIZE, off=off, count=cnt, mac=eng.mac,
                     pairs=eng.pairs, straddle_mass=eng.straddle_mass)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1) / repeats)
    mean = sum(ms) / parts
    return {"parts": parts, "slice_ms": ms, "max_ms": max(ms), "mean_ms": mean,
            "imbalance": max(ms) / mean}


def learned_summary(log):
    """The learned-theta half of the metric: the network is trained on the
    GPU from three training halo states (128^3, every group sampled) and
    judged on three held-out ones: per state, the fastest fixed (theta,
    straddle mass) pair meeting 1e-3 relative RMS error (FP64 direct sum;
    one theta per straddle mass of the grid, bisected) against the learned
    per-group (theta, straddle mass) at the same error, both errors confirmed
    on a disjoint sample of groups, both full-step times.  ratio = geometric
    mean of the step-rate ratios."""
    import numpy as np

    from src.core import learned_theta as L
    from src.utils import config as cfg
    t0 = time.perf_counter()
    c = cfg.preset("halos_128")
    vals = [(c.SEED, 0), (c.SEED + 4, 0), (c.SEED + 5, 0)]
    out = L.experiment("halos_128", 1e-3, "aabb", n_groups=32768, log=log, val_states=vals)
    wall = time.perf_counter() - t0
    states = []
    for res, _ in out:
        if not res.get("reachable"):
            states.append({"held_out_state": res["val_state"], "reachable": False})
            continue
        states.append({
            "held_out_state": f"halos_128 seed {res['val_state'][0]}",
            "ratio": res["speedup_step"],
            "theta_fixed": res["theta_fixed"], "straddle_fixed": res["straddle_fixed"],
            "fixed_by_straddle": res["fixed_by_straddle"], "lambda": res["lambda"],
            "err_fixed": res["err_fixed"], "err_learned": res["err_learned"],
            "err_fixed_confirm": res["err_fixed_confirm"],
            "err_learned_confirm": res["err_learned_confirm"],
            "step_ms_fixed": res["step_ms_fixed"], "step_ms_learned": res["step_ms_learned"],
            "traversal_ms_fixed": res["traversal_ms_fixed"],
            "traversal_ms_learned": res["traversal_ms_learned"], "select_ms": res["select_ms"],
            "hindsight_bound": res["hindsight_bound_speedup"],
            "interactions_fixed": res["inter_fixed"],
            "interactions_learned": res["inter_learned"],
            "straddle_group_histogram": res["straddle_group_histogram"]})
    ratios = [s_["ratio"] for s_ in states if "ratio" in s_]
    first = out[0][0]
    return {
        "ratio": (float(np.exp(np.mean(np.log(ratios)))) if len(ratios) == len(states)
                  else None),
        "ratio_kind": "geometric mean over the held-out states of fixed step ms / learned "
                      "step ms",
        "target_rel_rms": 1e-3,
        "test": f"AABB, bucket {c.BUCKET}, compensated sums; fixed: the fastest pair of one "
                f"theta and one straddle mass out of "
                f"{[c.STRADDLE_MASS * f for f in L.MU_FACTORS]} meeting the target; learned: "
                f"per-group theta x straddle mass from the same grid",
        "state": f"half of N={first.get('n')} in 64 Plummer spheres, half uniform; training "
                 f"seeds {[s_ for s_, _ in first.get('train_states', [])]}",
        "tuning_groups": first.get("val_groups"), "confirm_groups": first.get("confirm_groups"),
        "held_out": states,
        "train_rows": first.get("train_rows"), "fit_s": first.get("fit_s"), "wall_s": wall,
    }


def learned_line(args):
    """One JSON line with only the learned-theta half of the metric (see
    learned_summary)."""
    res = learned_summary(lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps({
        "metric": "learned-theta / fixed-theta step rate at equal 1e-3 relative RMS force error",
        "value": res["ratio"], "unit": "x", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "halos_128: " + res["state"] + "; " + res["test"]},
        "detail": res}, default=float))


def timed_step(eng, ev=None):
    """One engine step (the same stages as SimulationEngine.step plus the
    density splat) with HIP events between the stages when ev (six events)
    is given: sort | tree (build, upward pass, packing, and the learned
    per-group selection of a learned preset) | forces | update | splat."""
    from src.kernels.physics_kernels import integrate
    st, c = eng.state, eng.config
    H, dt = eng.timestep()
    if ev: ev[0].record()
    st.sort_buffer_by_morton(c)
    if ev: ev[1].record()
    eng.tree_manager.build_packed(st.pos, st.mass, st.codes, c.BOX_SIZE, eng.mac, eng.bucket)
    eng.select_theta()
    if ev: ev[2].record()
    eng.compute_forces()
    if ev: ev[3].record()
    integrate(*st.pos, *st.vel, st.force, eng.a, H, dt)
    if ev: ev[4].record()
    eng.camera.capture(eng.steps_done, *st.pos)
    if ev: ev[5].record()
    eng.a += eng.da
    eng.steps_done += 1


def cpu_baseline(pos, vel, c, threads):
    """One full step of the CPU oracle (C/OpenMP restatement of the reference
    kernels, FP64 force arithmetic) on the host cores."""
    import numpy as np

    from oracle import oracle as orc
    from src.utils.cosmology import get_hubble_param
    if threads > 0:
        orc.set_num_threads(threads)
    used = orc.num_threads()
    n = len(pos[0])
    t0 = time.perf_counter()
    codes, _ = orc.compute_codes(*pos)
    perm = orc.argsort_stable(codes)
    p = [x[perm] for x in pos]
    v = [x[perm] for x in vel]
    mass = np.full(n, 1.0 / n, np.float32)
    children, parents = orc.build_tree(codes[perm])
    nm, nc, lo, hi = orc.compute_multipoles(*p, mass, children, parents)
    com = orc.normalize_com(nm, nc)
    force = orc.compute_forces(*p, children, nm, com, lo, hi, c.THETA, c.G, c.SOFTENING,
                               c.BOX_SIZE, n)
    H = get_hubble_param(c.A_START, c)
    orc.integrate(*p, *v, force, c.A_START, H, c.DA / (c.A_START * H) * 0.8)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "particles/s", "cores": used, "kind": "port",
            "sample": f"one full step (codes, sort, tree, FP64 upward pass, FP64 BH forces, "
                      f"update) at N={n}, {dt:.2f} s on {used} host threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this build runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    from src.utils.cosmology import generate_ics

    if args.learned:
        learned_line(args)
        return
    c = cfg.preset(args.preset, TRAVERSAL_PAIRS=args.pairs)
    if args.theta is not None:
        c = c.replace(THETA=args.theta)
    pos, vel = generate_ics(c)
    ics_name = ("halo ICs" if getattr(c, "IC_KIND", "zeldovich") == "halos"
                else "Zel'dovich ICs")
    n = len(pos[0])
    torch.cuda.set_device(0)
    st = SimulationState(n, pos, vel)
    eng = SimulationEngine(c, st)

    def one_step(ev=None):
        timed_step(eng, ev)

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    if args.kernel_probe:
        eng.build_tree()
        for _ in range(3):
            eng.compute_forces()
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "traversal x3 + copy 256 MiB", "n": n}))
        return
    names = ["sort", "tree", "forces", "update", "splat"]
    if args.graph:
        eng.capture_graphs(args.steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            eng.step_graph()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        # traversal launch time on the final state, same stream, HIP events
        eng.build_tree()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            eng.compute_forces()
        e1.record()
        torch.cuda.synchronize()
        stage = {"forces": e0.elapsed_time(e1) / 3}
    else:
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(6)]
               for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            one_step(evs[k])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        stage = {nm: float(np.mean([evs[k][j].elapsed_time(evs[k][j + 1])
                                    for k in range(args.steps)])) for j, nm in enumerate(names)}
    # Work counters come from one extra traversal of the final state (the
    # counting kernel variant is not the one timed above).
    eng.build_tree()
    eng.stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    eng.compute_forces()
    inter, visits, pops, drops = (int(x) for x in eng.stats.cpu().tolist())
    eng.stats = None
    shards = None
    if args.shards > 1 and eng.theta_group is None:
        shards = shard_slices(eng, c, args.shards)
    ms = elapsed * 1e3 / args.steps
    trav_s = stage["forces"] * 1e-3
    flops = FLOP_PER_INTERACTION * inter
    achieved = flops / trav_s / 1e12
    traffic = TRAVERSAL_TRAFFIC_BYTES.get(c.NAME) if not eng.pairs else None
    out = {
        "metric": f"ms/step + particles/s at N={c.N_GRID}^3 (full step)",
        "value": n * args.steps / elapsed,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{c.NAME}: N={n} {ics_name} seed {c.SEED}, theta={c.THETA}, "
                               f"eps={c.SOFTENING:.4g}, MAC={c.MAC}, {c.RES}^2 splat, "
                               f"reference schedule da={c.DA:.3g}",
                   "n_bodies": n, "theta": c.THETA, "res": c.RES, "hip_graph": args.graph},
        "stages_ms": stage,
        "interactions_per_particle": inter / n,
        "visits_per_particle": visits / n,
        "wave_pops_per_particle": pops / n,
        "traversal": "pairs (128 particles per wavefront)" if eng.pairs else "64 per wavefront",
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_VALU_FP32_TFLOPS,
                     "peak_fp32_published": PEAK_FP32_TFLOPS,
                     "frac_of_published": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "hbm_frac": (traffic / trav_s / 1e9 / PEAK_HBM_GBS) if traffic else None,
                     "kernel": "k_traverse (FP32 VALU, no packed FP32; 20 flop per accepted "
                               "interaction)"},
    }
    if shards is not None:
        out["shard_slices"] = shards
    if eng.theta_group is not None:
        g = eng.theta_group
        out["config"]["learned_theta"] = {
            "network": c.THETA_NETWORK, "theta_group_mean": float(g.mean().item()),
            "theta_group_min": float(g.min().item()), "theta_group_max": float(g.max().item()),
            "fallbacks": int(eng.theta_selector.n_fallback)}
    if args.error_targets > 0:
        from src.kernels.physics_kernels import direct_sum
        k = min(args.error_targets, n)
        tg = torch.from_numpy(np.sort(np.random.default_rng(1234).choice(n, k, replace=False)))
        tg = tg.to(st.force.device)
        ref = direct_sum(*st.pos, st.mass, tg, c.G, c.SOFTENING)
        err = float(((st.force.index_select(0, tg).double() - ref) ** 2).sum().sqrt() /
                    (ref ** 2).sum().sqrt())
        out["force_error"] = {"rel_rms_vs_fp64_direct_sum": err, "targets": k,
                              "sample": "PCG64 seed 1234, final state"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos, vel, c, args.cpu_threads)
    if not args.no_learned and c.NAME == "repo_128":
        del eng, st
        torch.cuda.empty_cache()
        out["learned_vs_fixed"] = learned_summary(
            lambda m: print(f"[learned] {m}", file=sys.stderr, flush=True))
    print(json.dumps(out, default=float))


if __name__ == "__main__":
    main()
