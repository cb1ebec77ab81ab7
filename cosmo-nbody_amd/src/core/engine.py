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
steps and captures the density grid; video output and NVML polling of the
reference are not part of this package.
"""
from __future__ import annotations

import torch

from src.core.camera import Camera
from src.core.octree import OctreeManager
from src.kernels.physics_kernels import MAC_AABB, MAC_REF, integrate, traverse
from src.utils.cosmology import get_hubble_param


def mac_of(config) -> int:
    return MAC_AABB if str(getattr(config, "MAC", "ref")).lower() == "aabb" else MAC_REF


class SimulationEngine:
    def __init__(self, config, state, device="cuda"):
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
        self.theta_group = None     # per-group opening angles (learned theta), or None
        self.stats = None           # int64[2] interaction / visit counters, or None
        self.steps_done = 0

    def timestep(self):
        """(H, dt) at the current scale factor (engine.py:63-64)."""
        H = get_hubble_param(self.a, self.config)
        return H, self.da / (self.a * H) * 0.8

    def build_tree(self):
        s = self.state
        s.sort_buffer_by_morton(self.config)
        self.tree_manager.build_packed(s.pos, s.mass, s.codes, self.config.BOX_SIZE, self.mac)

    def compute_forces(self):
        s, c = self.state, self.config
        traverse(*s.pos, self.tree_manager.nodes, self.tree_manager.root_idx, s.force, c.THETA,
                 c.G, c.SOFTENING, c.BOX_SIZE, mac=self.mac, theta_group=self.theta_group,
                 compensated=bool(getattr(c, "COMPENSATED", False)), stats=self.stats)

    def step(self):
        H, dt = self.timestep()
        self.build_tree()
        self.compute_forces()
        s = self.state
        integrate(*s.pos, *s.vel, s.force, self.a, H, dt)
        self.a += self.da
        self.steps_done += 1

    def run(self, n_steps=None, capture=True):
        n_steps = self.config.RUN_STEPS if n_steps is None else n_steps
        for i in range(n_steps):
            self.step()
            if capture:
                self.camera.capture(self.steps_done - 1, *self.state.pos)
