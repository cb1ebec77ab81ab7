"""Per-step metrics as JSON lines (the reference prints a tqdm bar and polls
NVML every 20 steps, engine.py:130-135,153-158; this records what a run did
instead).

``SimulationEngine.run(n, metrics=StepMetrics("run.jsonl"))`` writes one line
per step:

  step, a (scale factor the step started from), n,
  ms: sort, tree (build, upward pass, records, per-group selection), forces,
      update, splat (when the density is captured) and their sum, total --
      HIP events on the launch stream between the stages (the force checks
      below run between forces and update and are in no stage),
  particles_per_s = n / total,
  counts=True adds the walk's interactions, visits and overflow drops per
      particle (the counting walk, about a third slower than the plain one),
      the traversal's FP32 rate at 20 flop per interaction, and the number of
      non-finite acceleration components,
  error_targets > 0 adds, every error_every steps, the relative RMS and the
      largest per-target relative error of the accelerations against the
      FP64 direct sum on that many sampled particles (fixed PCG64 sample),
  learned selection: mean / min / max of the group opening angles and the
      selection's fallbacks so far.

Each line is written after the step's work has finished (one device
synchronisation per step), so metrics slow a run down; eager steps only.
"""
from __future__ import annotations

import json

import numpy as np
import torch

STAGES = ("sort", "tree", "forces", "update", "splat")
FLOP_PER_INTERACTION = 20


class StepMetrics:
    def __init__(self, out, counts=False, error_targets=0, error_every=1, seed=1234):
        """out: a path (opened for appending) or a writable text stream."""
        if error_every < 1:
            raise ValueError("error_every must be >= 1")
        self._own = isinstance(out, str)
        self._out = open(out, "a") if self._own else out
        self.counts = bool(counts)
        self.error_targets = int(error_targets)
        self.error_every = int(error_every)
        self.seed = seed
        self.records = []
        self._targets = None
        self._events = {}
        self._stats = None
        self._nonfinite = None
        self._error = None

    def close(self):
        if self._own:
            self._out.close()

    # hooks called by SimulationEngine.step / run
    def begin(self, engine):
        dev = engine.state.pos[0].device
        self._a = float(engine.a)
        self._error = None
        self._events = {}
        if self.counts:
            if self._stats is None:
                self._stats = torch.zeros(4, dtype=torch.int64, device=dev)
                self._nonfinite = torch.zeros(1, dtype=torch.int64, device=dev)
            self._nonfinite.zero_()
            if engine.health and engine.stats is None:
                # the engine's own counters run anyway: this step's share is
                # their difference
                self._stats.copy_(engine._health_stats)
                self._swapped = False
            else:
                self._stats.zero_()
                self._saved_stats = engine.stats
                engine.stats = self._stats
                self._swapped = True
        self.mark("start")

    def mark(self, name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self._events[name] = ev

    def after_forces(self, engine):
        st = engine.state
        if self.counts:
            from src.kernels.physics_kernels import count_nonfinite
            count_nonfinite(st.force, self._nonfinite)
            if self._swapped:
                engine.stats = self._saved_stats
            else:
                torch.sub(engine._health_stats, self._stats, out=self._stats)
        k = engine.steps_done
        if self.error_targets > 0 and k % self.error_every == 0:
            from src.kernels.physics_kernels import direct_sum
            if self._targets is None:
                n = st.n
                m = min(self.error_targets, n)
                idx = np.sort(np.random.default_rng(self.seed).choice(n, m, replace=False))
                self._targets = torch.from_numpy(idx).to(st.force.device)
            c = engine.config
            ref = direct_sum(*st.pos, st.mass, self._targets, c.G, c.SOFTENING)
            got = st.force.index_select(0, self._targets).double()
            d2 = ((got - ref) ** 2).sum(1)
            r2 = (ref ** 2).sum(1)
            self._error = (d2, r2)
        # the checks above are not part of any stage: the update is timed
        # from here
        self.mark("checked")

    def end(self, engine):
        """Writes the step's line (synchronises) and returns it as a dict."""
        self.mark("end")
        self._events["end"].synchronize()
        ev = self._events
        ms, prev = {}, "start"
        for name in STAGES:
            if name in ev:
                # the update follows the force checks, which are not timed
                a = "checked" if name == "update" and "checked" in ev else prev
                ms[name] = ev[a].elapsed_time(ev[name])
                prev = name
        ms["total"] = sum(ms.values())
        n = engine.state.n
        rec = {"step": engine.steps_done - 1, "a": self._a, "n": n, "ms": ms,
               "particles_per_s": n / (ms["total"] * 1e-3) if ms["total"] > 0 else None}
        if self.counts:
            s = [int(x) for x in self._stats.cpu().tolist()]
            rec["interactions_per_particle"] = s[0] / n
            rec["visits_per_particle"] = s[1] / n
            rec["overflow_drops"] = s[3]
            rec["traversal_tflops"] = (FLOP_PER_INTERACTION * s[0] / (ms["forces"] * 1e-3) / 1e12
                                       if ms.get("forces") else None)
            rec["nonfinite_force_components"] = int(self._nonfinite.item())
        if self._error is not None:
            d2, r2 = self._error
            rec["force_error"] = {
                "rel_rms_vs_fp64_direct_sum": float(d2.sum().sqrt() / r2.sum().sqrt()),
                "max_rel": float((d2 / r2.clamp_min(1e-300)).sqrt().max()),
                "targets": int(d2.shape[0])}
        if engine.theta_group is not None:
            g = engine.theta_group
            rec["theta_group"] = {"mean": float(g.mean().item()), "min": float(g.min().item()),
                                  "max": float(g.max().item()),
                                  "fallbacks": int(engine.theta_selector.n_fallback)}
        self._out.write(json.dumps(rec) + "\n")
        self._out.flush()
        self.records.append(rec)
        return rec
