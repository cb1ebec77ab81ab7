"""Per-step driver (src/core/engine.py:27-123).

``SimulationEngine(config, state).step()`` follows the reference's stage
order: H(a) and dt = 0.8 da / (a H) on the host (engine.py:63-64), Morton
sort (state.sort_buffer_by_morton), tree build with the FP64 upward pass,
Barnes-Hut accelerations for every particle, the kick-drift update and
a += da.  Differences from the reference, all without effect on results:
one traversal launch instead of 5000-particle batches (engine.py:91-111), no
top-of-tree cache (build_top_tree_cache only feeds an on-chip copy of node
data, tree_kernels.py:264-314), no host synchronisation inside the step,
and normalize_com fused with packing of the node records.  ``run()``
steps and captures the density grid, and with a StepMetrics writes one JSON
line per step (src/core/metrics.py); video output and NVML polling of the
reference are not part of this package.  With LEARNED_THETA the engine
installs the trained per-group theta network named by THETA_NETWORK.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from src.core.camera import Camera
from src.core.octree import OctreeManager
from src.kernels import _hip
from src.kernels.physics_kernels import (MAC_AABB, MAC_REF, count_nonfinite, integrate,
                                        order_groups_by_pops, traverse)
from src.utils.cosmology import get_hubble_param
from src.utils.visualizer import get_mvp_matrix


NETWORK_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))), "networks")


def mac_of(config) -> int:
    return MAC_AABB if str(getattr(config, "MAC", "ref")).lower() == "aabb" else MAC_REF


def network_path(name: str) -> str:
    """A trained theta network by file path, or by name in the package's
    networks/ directory (e.g. "halos_128.npz")."""
    if os.path.exists(name):
        return name
    p = os.path.join(NETWORK_DIR, name)
    if os.path.exists(p):
        return p
    raise FileNotFoundError(f"trained theta network {name!r} not found (nor in {NETWORK_DIR})")


class SimulationEngine:
    def __init__(self, config, state, device="cuda", theta_selector=None, health=False):
        # LEARNED_THETA installs the network named by THETA_NETWORK (or the one
        # given); it never falls back to fixed theta silently
        if getattr(config, "LEARNED_THETA", False):
            if mac_of(config) != MAC_AABB:
                raise ValueError("LEARNED_THETA needs MAC='aabb': the network predicts the "
                                 "errors of the AABB acceptance test")
            if theta_selector is None:
                name = str(getattr(config, "THETA_NETWORK", "") or "")
                if not name:
                    raise ValueError("LEARNED_THETA is set but no trained network is named: "
                                     "set THETA_NETWORK (a file written by LearnedTheta.save) "
                                     "or pass theta_selector")
                path = network_path(name)
                from src.core.learned_theta import LearnedTheta
                theta_selector = LearnedTheta.load(path, device=device)
        self.config = config
        self.state = state
        self.tree_manager = OctreeManager(state.n, config.TPB, device=device)
        self.camera = Camera(config.N_STEPS, config.RES, config, device=device)
        self.cache_size = config.CACHE_SIZE
        self.tpb = config.TPB
        self.a = config.A_START
        self.da = (config.A_END - config.A_START) / config.N_STEPS
        self.top_nodes = torch.full((self.cache_size,), -1, dtype=torch.int32, device=device)
        self.node_to_cache = torch.full((self.tree_manager.n_nodes,), -1, dtype=torch.int32,
                                        device=device)
        self.mac = mac_of(config)
        # leaf buckets of the AABB records (0 = walk every leaf)
        self.bucket = int(getattr(config, "BUCKET", 0)) if self.mac == MAC_AABB else 0
        # AABB test: nodes lighter than straddle_mass pass the minimum-image guard
        self.straddle_mass = (float(getattr(config, "STRADDLE_MASS", 0.0))
                              if self.mac == MAC_AABB else 0.0)
        self.theta_group = None     # per-group opening angles (learned theta), or None
        # LearnedTheta: recomputes theta_group after each build
        self.theta_selector = theta_selector
        self.group_order = None     # launch order of the groups under per-group theta
        self.straddle_group = None  # per-group straddle masses chosen with theta_group
        self._a_dev = None          # device scale factor while a step is being captured
        self.stats = None           # int64[4] interaction / visit / pop / drop counters, or None
        self.pairs = bool(getattr(config, "TRAVERSAL_PAIRS", False))
        # LONGEST_FIRST: each traversal stores its groups' stack pops, and the
        # next one launches the groups by decreasing pops (zeros at first)
        self.longest_first = (bool(getattr(config, "LONGEST_FIRST", False)) and not self.pairs
                              and self.mac == MAC_AABB)
        if self.longest_first:
            groups = (state.n + 63) // 64
            self.group_pops = torch.zeros(groups, dtype=torch.int32, device=device)
            self._pops_order = torch.zeros(groups, dtype=torch.int32, device=device)
        # a per-group selection since the last launch (its pops then belong to
        # the previous candidates)
        self._fresh_selection = False
        self.steps_done = 0
        # failure detection (health=True): the traversal's counters (its
        # overflow drops among them) and the non-finite accelerations,
        # accumulated on the device over the steps, read by health_report()
        self.health = bool(health)
        self._health_stats = (torch.zeros(4, dtype=torch.int64, device=device) if health
                              else None)
        self._nonfinite = torch.zeros(1, dtype=torch.int64, device=device) if health else None
        self._health_steps = 0

    def timestep(self):
        """(H, dt) at the current scale factor (engine.py:63-64)."""
        H = get_hubble_param(self.a, self.config)
        return H, self.da / (self.a * H) * 0.8

    def build_tree(self, probe=None):
        s = self.state
        s.sort_buffer_by_morton(self.config)
        if probe is not None:
            probe.mark("sort")
        self.tree_manager.build_packed(s.pos, s.mass, s.codes, self.config.BOX_SIZE, self.mac,
                                       self.bucket)
        self.select_theta()

    def select_theta(self):
        """Per-group (theta, straddle mass) of the learned selection for the
        current tree, and the traversal's launch order (no-op without a
        selector)."""
        if self.theta_selector is not None:
            self.theta_group = self.theta_selector.theta_groups(self, a=self._a_dev)
            self.straddle_group = getattr(self.theta_selector, "mu_group", None)
            self._fresh_selection = True
            if not self.longest_first:
                self.group_order = self.theta_selector.launch_order(self.theta_group)

    def launch_order(self, scale=True):
        """The traversal's launch order of the groups: by the previous
        launch's pops (LONGEST_FIRST; with scale after a new learned
        selection, each group's pops scaled by the change of its candidate's
        cost since that launch), else by the learned thetas, else none."""
        if self.longest_first:
            sel = self.theta_selector
            if scale and sel is not None and self.theta_group is not None:
                return order_groups_by_pops(self.group_pops, self._pops_order, sel.k_group,
                                            sel.k_prev, sel.cost_dev)
            return order_groups_by_pops(self.group_pops, self._pops_order)
        if self.theta_group is not None and not self.pairs:
            return self.group_order
        return None

    def compute_forces(self):
        s, c = self.state, self.config
        stats = self._health_stats if self.health and self.stats is None else self.stats
        traverse(*s.pos, self.tree_manager.nodes, self.tree_manager.root_idx, s.force, c.THETA,
                 c.G, c.SOFTENING, c.BOX_SIZE, mac=self.mac, theta_group=self.theta_group,
                 compensated=bool(getattr(c, "COMPENSATED", False)), stats=stats,
                 pairs=self.pairs, straddle_mass=self.straddle_mass,
                 straddle_group=self.straddle_group if self.theta_group is not None else None,
                 group_order=self.launch_order(scale=self._fresh_selection),
                 group_pops=self.group_pops if self.longest_first else None)
        self._fresh_selection = False
        if self.health:
            count_nonfinite(s.force, self._nonfinite)
            self._health_steps += 1

    def health_report(self):
        """Counters of the steps run with health=True (synchronises): overflow
        drops (nodes a particle's full 64-entry stack could not open, whose
        children the reference drops), non-finite acceleration components,
        per-group theta fallbacks of the learned selection."""
        if not self.health:
            raise RuntimeError("the engine was created without health=True")
        st = [int(x) for x in self._health_stats.cpu().tolist()]
        return {"steps": self._health_steps, "interactions": st[0], "visits": st[1],
                "overflow_drops": st[3], "nonfinite_force_components": int(self._nonfinite.item()),
                "theta_fallbacks": (self.theta_selector.n_fallback if self.theta_selector
                                    is not None else 0)}

    def step(self, probe=None):
        """One step.  probe: a src.core.metrics.StepMetrics that records HIP
        events between the stages (and samples the forces), or None."""
        H, dt = self.timestep()
        if probe is not None:
            probe.begin(self)
        self.build_tree(probe)
        if probe is not None:
            probe.mark("tree")
        self.compute_forces()
        if probe is not None:
            probe.mark("forces")
            probe.after_forces(self)
        s = self.state
        integrate(*s.pos, *s.vel, s.force, self.a, H, dt)
        if probe is not None:
            probe.mark("update")
        self.a += self.da
        self.steps_done += 1

    def run(self, n_steps=None, capture=True, metrics=None):
        """n_steps steps (the configuration's RUN_STEPS by default), the
        density captured after each; metrics: a StepMetrics that writes one
        JSON line per step."""
        n_steps = self.config.RUN_STEPS if n_steps is None else n_steps
        for i in range(n_steps):
            self.step(metrics)
            if capture:
                self.camera.capture(self.steps_done - 1, *self.state.pos)
                if metrics is not None:
                    metrics.mark("splat")
            if metrics is not None:
                metrics.end(self)

    # ------------------------------------------------------------ HIP graphs
    #
    # The whole step (sort, tree, traversal, update, density splat) is
    # captured once per ping-pong parity of the particle buffers.  Per-step
    # scalars (kick-drift coefficients from H(a), dt and a; the orbiting MVP)
    # are precomputed on the host for the next `steps` steps, exactly as
    # step() would compute them, and a device step counter selects the row
    # inside the graph, so a replay needs no host-to-device copy.

    def _graph_body(self):
        s, c = self.state, self.config
        self._coefs.copy_(torch.index_select(self._coef_table, 0, self._step_dev).reshape(-1))
        self._mvp.copy_(torch.index_select(self._mvp_table, 0, self._step_dev).reshape(-1))
        self.build_tree()
        self.compute_forces()
        L = _hip.load()
        P = [_hip.ptr(t, torch.float32, "pos") for t in s.pos]
        V = [_hip.ptr(t, torch.float32, "vel") for t in s.vel]
        _hip.check(L.cosmo_kick_drift_dev(*P, *V, _hip.ptr(s.force, torch.float32, "force"),
                                          _hip.ptr(self._coefs, torch.float32, "coefs"), s.n,
                                          _hip.stream_handle()), "kick_drift_dev")
        grid = self.camera.grid_d
        grid.zero_()
        _hip.check(L.cosmo_render_density_dev(
            *[_hip.ptr(t, torch.float32, "pos") for t in s.pos], _hip.ptr(grid, _hip.U64, "grid"),
            _hip.ptr(self._mvp, torch.float32, "mvp"), self.camera.res, s.n,
            self.camera.frac_bits, _hip.stream_handle()), "render_density_dev")
        self._step_dev += 1

    def capture_graphs(self, steps):
        """Prepares graph replays for the next `steps` steps (density each step)."""
        c, dev = self.config, self.state.pos[0].device
        coefs, mvps, a = [], [], self.a
        for k in range(steps):
            H = get_hubble_param(a, c)
            dt = self.da / (a * H) * 0.8
            cd = np.float32(1.0 - H * dt)
            ck = np.float32(dt / a)
            coefs.append([cd, ck, ck, np.float32(a)])
            mvps.append(get_mvp_matrix(self.steps_done + k, self.camera.box_center, c.BOX_SIZE,
                                       c).reshape(-1))
            a += self.da
        self._coef_table = torch.tensor(np.array(coefs, np.float32), device=dev)
        self._mvp_table = torch.tensor(np.array(mvps, np.float32), device=dev)
        self._step_dev = torch.zeros(1, dtype=torch.int64, device=dev)
        self._coefs = torch.zeros(4, dtype=torch.float32, device=dev)
        self._mvp = torch.zeros(16, dtype=torch.float32, device=dev)
        self._graph_steps = steps
        self._graph_done = 0
        st = self.state
        _ = st.ctx  # lazily created scratch must exist before capture
        if self.theta_selector is not None:
            self.theta_selector.theta_groups(self)  # its buffers too
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        self._graphs, self._views = [], []
        self._a_dev = self._coefs[3:4]  # features read a from the step table
        try:
            for _parity in range(2):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    self._graph_body()
                self._graphs.append(g)
                self._views.append((st.pos, st.vel, st.mass, st._alt))
        finally:
            self._a_dev = None
        torch.cuda.current_stream().wait_stream(side)
        self._step_dev.zero_()
        self._parity = 0

    def step_graph(self):
        """One step by graph replay (after capture_graphs)."""
        if self._graph_done >= self._graph_steps:
            raise RuntimeError("captured step tables exhausted; call capture_graphs again")
        self._graphs[self._parity].replay()
        st = self.state
        st.pos, st.vel, st.mass, st._alt = self._views[self._parity]
        self._parity ^= 1
        self._graph_done += 1
        self.a += self.da
        self.steps_done += 1
