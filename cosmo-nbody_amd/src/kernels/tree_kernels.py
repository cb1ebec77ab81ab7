"""Tree construction kernels (src/kernels/tree_kernels.py).

Drop-ins with the reference's names and arguments, each running one gfx950
kernel through the C ABI:

  reset_tree_buffers(child, parent, counters, root_idx)            :236-262
  build_tree(codes, children, parents)                               :60-129
  compute_multipoles(pos_x, pos_y, pos_z, mass, children, parents,
                     n_mass, n_com, n_min, n_max, counts)           :131-191
  normalize_com(n_mass, n_com, n_nodes)                             :193-214
  find_root(parents, root_idx)                                      :216-234
  build_top_tree_cache(child, root_idx, top_nodes, node_to_cache,
                       cache_size)                                  :264-314

Node mass and CoM tensors may be float64 (this build's engine: the upward
pass accumulates in FP64 and keeps it) or float32, as the reference
allocates them (octree.py:37-38): then the FP64 sums are rounded once to
FP32 on store, and normalize_com works in the reference's arithmetic on the
float32 arrays.  build_tree and compute_multipoles need a launch that covers
every node, as all of the reference's callers use.

Additions without a reference counterpart: pack_nodes / normalize_pack (the
packed FP32 leaf and internal-node records the traversal reads, a uint8
buffer of tree_bytes(n_nodes, mac) bytes), build_tree_spans and
normalize_pack_buckets (AABB records of small nodes carry their leaf range,
which the traversal sums directly when it opens them), and validate_tree.
"""
from __future__ import annotations

import torch

from ._hip import (MAC_AABB, MAC_REF, U32, U64, bump_tree_generation, check, kernel,
                   launch_extent, load, ptr, shared_context, stream_handle)

__all__ = ["reset_tree_buffers", "build_tree", "build_tree_spans", "compute_multipoles",
           "normalize_com", "find_root", "build_top_tree_cache", "pack_nodes", "normalize_pack",
           "normalize_pack_buckets", "validate_tree", "node_buffer", "MAC_REF", "MAC_AABB"]


def _full(launch, need, name):
    if launch is not None and launch[0] * launch[1] < need:
        raise ValueError(f"{name}: the launch must cover all {need} work items "
                         f"(got {launch[0]}x{launch[1]})")


@kernel
def reset_tree_buffers(child, parent, counters, root_idx, _launch=None):
    m = launch_extent(_launch, child.shape[0])
    bump_tree_generation()
    check(load().cosmo_reset_tree_buffers(
        ptr(child, torch.int32, "child"), ptr(parent, torch.int32, "parent", numel=m),
        ptr(counters, torch.int32, "counters", numel=m), ptr(root_idx, torch.int32, "root_idx"),
        m, stream_handle()), "reset_tree_buffers")


@kernel
def build_tree(codes, children, parents, _launch=None):
    n = codes.shape[0]
    _full(_launch, n - 1, "build_tree")
    m = 2 * n - 1
    bump_tree_generation()
    check(load().cosmo_build_tree(
        ptr(codes, U64, "codes"), ptr(children, torch.int32, "children", numel=2 * m),
        ptr(parents, torch.int32, "parents", numel=m), n, stream_handle()), "build_tree")


def build_tree_spans(codes, children, parents, far):
    """build_tree that also writes far[i] (int32[n-1]): the other end of
    internal node n+i's leaf range."""
    n = codes.shape[0]
    m = 2 * n - 1
    bump_tree_generation()
    check(load().cosmo_build_tree_spans(
        ptr(codes, U64, "codes"), ptr(children, torch.int32, "children", numel=2 * m),
        ptr(parents, torch.int32, "parents", numel=m),
        ptr(far, torch.int32, "far", numel=max(n - 1, 0)), n, stream_handle()),
        "build_tree_spans")


@kernel
def compute_multipoles(pos_x, pos_y, pos_z, mass, children, parents, n_mass, n_com, n_min,
                       n_max, counts, _launch=None):
    n = pos_x.shape[0]
    _full(_launch, n, "compute_multipoles")
    m = 2 * n - 1
    f32 = torch.float32
    bump_tree_generation()
    if n_mass.dtype == f32:  # the reference's float32 node arrays
        ctx = shared_context(n)
        check(load().cosmo_compute_multipoles_f32(
            ctx.handle, ptr(pos_x, f32, "pos_x", numel=n), ptr(pos_y, f32, "pos_y", numel=n),
            ptr(pos_z, f32, "pos_z", numel=n), ptr(mass, f32, "mass", numel=n),
            ptr(children, torch.int32, "children", numel=2 * m),
            ptr(parents, torch.int32, "parents", numel=m),
            ptr(n_mass, f32, "n_mass", numel=m), ptr(n_com, f32, "n_com", numel=3 * m),
            ptr(n_min, f32, "n_min", numel=3 * m), ptr(n_max, f32, "n_max", numel=3 * m),
            ptr(counts, torch.int32, "counts", numel=m), n, stream_handle()),
            "compute_multipoles")
        ctx.packed_key = None  # the FP64 sums were formed in the record buffer
        return
    check(load().cosmo_compute_multipoles(
        ptr(pos_x, f32, "pos_x", numel=n), ptr(pos_y, f32, "pos_y", numel=n),
        ptr(pos_z, f32, "pos_z", numel=n), ptr(mass, f32, "mass", numel=n),
        ptr(children, torch.int32, "children", numel=2 * m),
        ptr(parents, torch.int32, "parents", numel=m),
        ptr(n_mass, torch.float64, "n_mass", numel=m),
        ptr(n_com, torch.float64, "n_com", numel=3 * m),
        ptr(n_min, f32, "n_min", numel=3 * m), ptr(n_max, f32, "n_max", numel=3 * m),
        ptr(counts, torch.int32, "counts", numel=m), n, stream_handle()), "compute_multipoles")


@kernel
def normalize_com(n_mass, n_com, n_nodes, _launch=None):
    m = launch_extent(_launch, int(n_nodes))
    bump_tree_generation()
    if n_mass.dtype == torch.float32:
        check(load().cosmo_normalize_com_f32(
            ptr(n_mass, torch.float32, "n_mass", numel=m),
            ptr(n_com, torch.float32, "n_com", numel=3 * m), m, stream_handle()),
            "normalize_com")
        return
    check(load().cosmo_normalize_com(
        ptr(n_mass, torch.float64, "n_mass", numel=m),
        ptr(n_com, torch.float64, "n_com", numel=3 * m), m, stream_handle()), "normalize_com")


@kernel
def find_root(parents, root_idx, _launch=None):
    m = launch_extent(_launch, parents.shape[0])
    bump_tree_generation()
    check(load().cosmo_find_root(ptr(parents, torch.int32, "parents", numel=m),
                                 ptr(root_idx, torch.int32, "root_idx"), m, stream_handle()),
          "find_root")


@kernel
def build_top_tree_cache(child, root_idx, top_nodes, node_to_cache, cache_size, _launch=None):
    cs = int(cache_size)
    check(load().cosmo_build_top_tree_cache(
        ptr(child, torch.int32, "child"), ptr(root_idx, torch.int32, "root_idx"),
        ptr(top_nodes, torch.int32, "top_nodes", numel=cs),
        ptr(node_to_cache, torch.int32, "node_to_cache"), cs, stream_handle()),
        "build_top_tree_cache")


def tree_bytes(n_nodes: int, mac=MAC_AABB) -> int:
    """Bytes of the packed records of a tree with n_nodes = 2n-1 nodes."""
    return int(load().cosmo_tree_bytes((n_nodes + 1) // 2, int(mac)))


def node_buffer(n_nodes: int, device=None, mac=MAC_AABB) -> torch.Tensor:
    """Uninitialised byte storage for the packed records (room for `mac`)."""
    return torch.empty(tree_bytes(n_nodes, mac), dtype=torch.uint8, device=device or "cuda")


def _pack(fn, children, n_mass, n_com, n_min, n_max, box, mac, nodes, what):
    m = children.shape[0]
    bump_tree_generation()
    check(fn(ptr(children, torch.int32, "children", numel=2 * m),
             ptr(n_mass, torch.float64, "n_mass", numel=m),
             ptr(n_com, torch.float64, "n_com", numel=3 * m),
             ptr(n_min, torch.float32, "n_min", numel=3 * m),
             ptr(n_max, torch.float32, "n_max", numel=3 * m), m, float(box), int(mac),
             ptr(nodes, torch.uint8, "nodes", numel=tree_bytes(m, mac)), stream_handle()), what)
    return nodes


def pack_nodes(children, n_mass, n_com, n_min, n_max, box, mac=MAC_REF, nodes=None):
    """Packs normalised node arrays (float64, or the reference's float32)
    into FP32 records for acceptance test mac."""
    nodes = node_buffer(children.shape[0], children.device, mac) if nodes is None else nodes
    if n_mass.dtype == torch.float32:
        m = children.shape[0]
        f32 = torch.float32
        check(load().cosmo_pack_nodes_f32(
            ptr(children, torch.int32, "children", numel=2 * m),
            ptr(n_mass, f32, "n_mass", numel=m), ptr(n_com, f32, "n_com", numel=3 * m),
            ptr(n_min, f32, "n_min", numel=3 * m), ptr(n_max, f32, "n_max", numel=3 * m), m,
            float(box), int(mac), ptr(nodes, torch.uint8, "nodes", numel=tree_bytes(m, mac)),
            stream_handle()), "pack_nodes")
        return nodes
    return _pack(load().cosmo_pack_nodes, children, n_mass, n_com, n_min, n_max, box, mac,
                 nodes, "pack_nodes")


def normalize_pack(children, n_mass, n_com, n_min, n_max, box, mac=MAC_REF, nodes=None):
    """normalize_com fused with pack_nodes (n_com is normalised in place)."""
    nodes = node_buffer(children.shape[0], children.device, mac) if nodes is None else nodes
    return _pack(load().cosmo_normalize_pack, children, n_mass, n_com, n_min, n_max, box, mac,
                 nodes, "normalize_pack")


def normalize_pack_buckets(children, far, n_mass, n_com, n_min, n_max, box, bucket, nodes=None):
    """normalize_pack for the AABB test; nodes of at most `bucket` leaves
    (1..64) also carry their leaf range (from build_tree_spans' far)."""
    m = children.shape[0]
    n = (m + 1) // 2
    nodes = node_buffer(m, children.device, MAC_AABB) if nodes is None else nodes
    bump_tree_generation()
    check(load().cosmo_normalize_pack_buckets(
        ptr(children, torch.int32, "children", numel=2 * m),
        ptr(far, torch.int32, "far", numel=max(n - 1, 0)),
        ptr(n_mass, torch.float64, "n_mass", numel=m),
        ptr(n_com, torch.float64, "n_com", numel=3 * m),
        ptr(n_min, torch.float32, "n_min", numel=3 * m),
        ptr(n_max, torch.float32, "n_max", numel=3 * m), m, float(box), MAC_AABB, int(bucket),
        ptr(nodes, torch.uint8, "nodes", numel=tree_bytes(m, MAC_AABB)), stream_handle()),
        "normalize_pack_buckets")
    return nodes


def validate_tree(children, parents, n: int, ctx=None) -> int:
    """Defect bits of the tree (0 = well formed); synchronises the stream."""
    ctx = ctx or shared_context(n)
    flags = torch.zeros(1, dtype=torch.int32, device=children.device)
    check(load().cosmo_validate_tree(ctx.handle, ptr(children, torch.int32, "children"),
                                     ptr(parents, torch.int32, "parents"), int(n),
                                     ptr(flags, torch.int32, "flags"), stream_handle()),
          "validate_tree")
    return int(flags.item())
