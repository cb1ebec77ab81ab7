"""Force-accuracy study (no reference counterpart): Barnes-Hut accelerations
of sampled targets against the FP64 direct sum, swept over theta, the
acceptance test (reference / AABB) and plain vs compensated FP32
accumulation.

Error metrics over the targets s:
  rel_rms   = sqrt(sum_s |a_s - d_s|^2 / sum_s |d_s|^2)   (primary)
  mean_rel  = sqrt(mean_s |a_s - d_s|^2 / |d_s|^2)

Run: python -m src.core.precision --preset clustered_128 --evolve 300
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from src.kernels.physics_kernels import MAC_AABB, MAC_REF, direct_sum, traverse


def sample_targets(n, n_targets, seed=1234):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.sort(rng.choice(n, size=min(n_targets, n), replace=False))


def errors(force, ref):
    d = force.double() - ref
    num = (d ** 2).sum(1)
    den = (ref ** 2).sum(1)
    return (float(torch.sqrt(num.sum() / den.sum())),
            float(torch.sqrt((num / den.clamp_min(1e-300)).mean())))


def sweep(engine, thetas, n_targets=16384, seed=1234,
          variants=(("ref", 0, 0.0), ("aabb", 0, 0.0), ("aabb", 32, 0.0)),
          compensated=(False, True)):
    """engine: a SimulationEngine whose tree is built for the current state.
    variants: (acceptance test, leaf bucket, straddle mass) triples; the
    engine's tree is re-packed for each and restored at the end."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    idx = sample_targets(st.n, n_targets, seed)
    it = torch.from_numpy(idx).to(st.pos[0].device)
    tpos = [p.index_select(0, it).contiguous() for p in st.pos]
    t0 = time.perf_counter()
    ref = direct_sum(*st.pos, st.mass, it, c.G, c.SOFTENING)
    torch.cuda.synchronize()
    t_direct = time.perf_counter() - t0
    out = []
    force = torch.empty((len(idx), 3), dtype=torch.float32, device=ref.device)
    stats = torch.zeros(4, dtype=torch.int64, device=ref.device)
    own = (engine.mac, engine.bucket)
    for name, bucket, mu in variants:
        mac = MAC_REF if name == "ref" else MAC_AABB
        # the particles are already in Morton order, so the rebuild keeps them
        engine.mac, engine.bucket = mac, (bucket if mac == MAC_AABB else 0)
        engine.build_tree()
        for th in thetas:
            for comp in compensated:
                stats.zero_()
                traverse(*tpos, tm.nodes, tm.root_idx, force, th, c.G, c.SOFTENING, c.BOX_SIZE,
                         mac=mac, compensated=comp, stats=stats, n_nodes=tm.n_nodes,
                         straddle_mass=mu if mac == MAC_AABB else 0.0)
                r, m = errors(force, ref)
                out.append({"mac": name, "bucket": engine.bucket, "straddle_mass": mu,
                            "theta": th, "compensated": comp, "rel_rms": r, "mean_rel": m,
                            "interactions": float(stats[0].item()) / len(idx),
                            "visits": float(stats[1].item()) / len(idx)})
    engine.mac, engine.bucket = own
    engine.build_tree()
    return out, t_direct


def clustering(engine, cells=128):
    """max/mean and variance of the CIC-free counts-in-cells density."""
    pos = torch.stack(engine.state.pos, 1)
    ijk = torch.clamp((pos * cells).long(), 0, cells - 1)
    flat = (ijk[:, 0] * cells + ijk[:, 1]) * cells + ijk[:, 2]
    cnt = torch.bincount(flat, minlength=cells ** 3).double()
    mean = cnt.mean()
    return {"cells": cells, "max_over_mean": float(cnt.max() / mean),
            "delta_rms": float(((cnt / mean - 1) ** 2).mean().sqrt())}


def main():
    import sys
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils import config as cfg
    from src.utils.cosmology import generate_ics
    ap = argparse.ArgumentParser()
    ap.add_argument("--preset", default="repo_128")
    ap.add_argument("--evolve", type=int, default=0)
    ap.add_argument("--targets", type=int, default=16384)
    ap.add_argument("--thetas", default="0.1,0.2,0.3,0.5,0.7,0.9")
    args = ap.parse_args()
    c = cfg.preset(args.preset)
    pos, vel = generate_ics(c)
    st = SimulationState(len(pos[0]), pos, vel)
    eng = SimulationEngine(c, st)
    info = {"preset": c.NAME, "n": st.n, "evolve_steps": args.evolve,
            "ics": clustering(eng)}
    for _ in range(args.evolve):
        eng.step()
    torch.cuda.synchronize()
    info["evolved"] = clustering(eng)
    info["a"] = eng.a
    eng.build_tree()
    thetas = [float(x) for x in args.thetas.split(",")]
    rows, t_direct = sweep(eng, thetas, args.targets)
    info["direct_sum_s"] = t_direct
    print(json.dumps(info), flush=True)
    for r in rows:
        print(json.dumps(r), flush=True)
    sys.stdout.flush()


if __name__ == "__main__":
    main()
