"""Tree buffers and the build sequence (src/core/octree.py).

``OctreeManager(n_bodies, tpb)`` allocates the 2n-1 node buffers once
(octree.py:27-43) and ``build(pos, mass, codes)`` runs reset -> build_tree ->
compute_multipoles -> normalize_com -> find_root (octree.py:45-83) with the
gfx950 drop-ins.  ``get_buffers()`` returns the same keys as the reference.
Node mass and CoM are float64 (FP64 upward pass).

``build_packed(pos, mass, codes, box, mac, bucket)`` is the step path: it
fuses normalize_com with packing of the FP32 node records the traversal
reads; with the AABB test and bucket > 0, nodes of at most `bucket` leaves
carry their leaf range (the build also records every internal node's range).
"""
from __future__ import annotations

import torch

from src.kernels.tree_kernels import (MAC_AABB, MAC_REF, build_tree, build_tree_spans,
                                      compute_multipoles, find_root, node_buffer, normalize_com,
                                      normalize_pack, normalize_pack_buckets,
                                      reset_tree_buffers, validate_tree)


class OctreeManager:
    def __init__(self, n_bodies, tpb=256, device="cuda"):
        self.n = int(n_bodies)
        self.n_nodes = 2 * self.n - 1
        self.tpb = tpb
        self.bpg = (self.n + tpb - 1) // tpb
        dev = torch.device(device)
        M = self.n_nodes
        self.child = torch.full((M, 2), -1, dtype=torch.int32, device=dev)
        self.parent = torch.full((M,), -1, dtype=torch.int32, device=dev)
        self.root_idx = torch.zeros(1, dtype=torch.int32, device=dev)
        self.mass = torch.zeros(M, dtype=torch.float64, device=dev)
        self.com = torch.zeros((M, 3), dtype=torch.float64, device=dev)
        self.aabb_min = torch.zeros((M, 3), dtype=torch.float32, device=dev)
        self.aabb_max = torch.zeros((M, 3), dtype=torch.float32, device=dev)
        self.counters = torch.zeros(M, dtype=torch.int32, device=dev)
        self.far = torch.zeros(max(self.n - 1, 1), dtype=torch.int32, device=dev)
        self.nodes = node_buffer(M, dev)

    def _topology_and_moments(self, pos_d, mass_d, codes_d, spans=False):
        bpg_nodes = (self.n_nodes + self.tpb - 1) // self.tpb
        reset_tree_buffers[bpg_nodes, self.tpb](self.child, self.parent, self.counters,
                                                self.root_idx)
        if spans:
            build_tree_spans(codes_d, self.child, self.parent, self.far)
        else:
            build_tree[self.bpg, self.tpb](codes_d, self.child, self.parent)
        px, py, pz = pos_d
        compute_multipoles[self.bpg, self.tpb](px, py, pz, mass_d, self.child, self.parent,
                                               self.mass, self.com, self.aabb_min,
                                               self.aabb_max, self.counters)

    def build(self, pos_d, mass_d, codes_d):
        self._topology_and_moments(pos_d, mass_d, codes_d)
        normalize_com[(self.n_nodes + 255) // 256, 256](self.mass, self.com, self.n_nodes)
        bpg_nodes = (self.n_nodes + self.tpb - 1) // self.tpb
        find_root[bpg_nodes, self.tpb](self.parent, self.root_idx)
        return self.root_idx

    def build_packed(self, pos_d, mass_d, codes_d, box=1.0, mac=MAC_REF, bucket=0):
        buckets = mac == MAC_AABB and bucket > 0
        self._topology_and_moments(pos_d, mass_d, codes_d, spans=buckets)
        if buckets:
            normalize_pack_buckets(self.child, self.far, self.mass, self.com, self.aabb_min,
                                   self.aabb_max, box, bucket, self.nodes)
        else:
            normalize_pack(self.child, self.mass, self.com, self.aabb_min, self.aabb_max, box,
                           mac, self.nodes)
        bpg_nodes = (self.n_nodes + self.tpb - 1) // self.tpb
        find_root[bpg_nodes, self.tpb](self.parent, self.root_idx)
        return self.root_idx

    def validate(self) -> int:
        return validate_tree(self.child, self.parent, self.n)

    def get_buffers(self):
        return {
            "child": self.child,
            "mass": self.mass,
            "com": self.com,
            "min": self.aabb_min,
            "max": self.aabb_max,
            "root": self.root_idx,
        }
