"""Force and update kernels (src/kernels/physics_kernels.py).

``compute_forces[blocks, threads](pos_x, pos_y, pos_z, children, n_mass,
n_com, n_min, n_max, force, theta, G, soft, box, root, top_nodes,
node_to_cache, off)`` computes Barnes-Hut accelerations for particles
[off, off + blocks*threads) with the reference's acceptance test and walk
(physics_kernels.py:14-153) on the gfx950 wavefront-shared-stack traversal.
``integrate[blocks, threads](pos_x, pos_y, pos_z, vel_x, vel_y, vel_z, force,
a, H, dt)`` is the kick-drift update (:155-203) in FP32.

``traverse`` (packed records, either acceptance test, per-group theta,
compensated sums, interaction counts) and ``direct_sum`` (FP64 yardstick)
have no reference counterpart.
"""
from __future__ import annotations

import torch

from ._hip import (MAC_AABB, MAC_REF, TRAV_COMPENSATED, TRAV_GROUP_STATS, TRAV_PAIRS, TRAV_STATS,
                   U32, check, kernel, load, ptr, shared_context, stream_handle, tree_generation)

__all__ = ["compute_forces", "integrate", "kick_drift", "traverse", "order_groups",
           "order_groups_by_pops", "direct_sum",
           "count_nonfinite", "MAC_REF", "MAC_AABB"]

_F32 = torch.float32


def _xyz(pos_x, pos_y, pos_z, n):
    return [ptr(t, _F32, nm, numel=n) for t, nm in ((pos_x, "pos_x"), (pos_y, "pos_y"),
                                                     (pos_z, "pos_z"))]


def _packed_key(ctx, tensors, n, box):
    """What decides the packed records: the arrays (storage, torch version),
    the tree generation of the drop-ins that write them, n and the box."""
    return (tree_generation(), n, float(box),
            tuple((t.data_ptr(), t._version, t.dtype) for t in tensors))


@kernel
def compute_forces(pos_x, pos_y, pos_z, children, n_mass, n_com, n_min, n_max, force, theta, G,
                   soft, box, root, top_nodes, node_to_cache, off, _launch=None):
    """Node arrays as the reference allocates them (float32, octree.py:37-40)
    or this build's float64 mass and CoM.  The reference calls this once per
    5000-particle batch on one tree (engine.py:91-111): the FP32 records are
    packed into the shared context on the first call for a tree and reused
    while the arrays (and the root) are unchanged, so a batch only runs the
    walk; each pack starts a new epoch, under which the library plans the
    split walk's frontiers for the whole batch loop once."""
    n = pos_x.shape[0]
    off = int(off)
    count = (n - off) if _launch is None else _launch[0] * _launch[1]
    # a launch that covers no particle writes nothing (every thread of the
    # reference's leaves at i = tid + off >= n, physics_kernels.py:48-50,76-77);
    # a negative offset is left to the library, which refuses it
    if off >= 0 and (n == 0 or count <= 0 or off >= n):
        return
    m = 2 * n - 1
    ctx = shared_context(n)
    L = load()
    key = _packed_key(ctx, (children, n_mass, n_com, n_min, n_max, root), n, box)
    if ctx.packed_key != key:
        f = n_mass.dtype
        if f not in (torch.float32, torch.float64) or n_com.dtype != f:
            raise TypeError(f"compute_forces: n_mass/n_com must both be float32 or float64, "
                            f"got {n_mass.dtype}/{n_com.dtype}")
        pack = L.cosmo_pack_nodes_f32 if f == torch.float32 else L.cosmo_pack_nodes
        check(pack(ptr(children, torch.int32, "children", numel=2 * m),
                   ptr(n_mass, f, "n_mass", numel=m), ptr(n_com, f, "n_com", numel=3 * m),
                   ptr(n_min, _F32, "n_min", numel=3 * m), ptr(n_max, _F32, "n_max", numel=3 * m),
                   m, float(box), MAC_REF, ctx.nodes_ptr, stream_handle()), "compute_forces")
        ctx.packed_key = key
        ctx.packed_epoch += 1
    check(L.cosmo_compute_forces_packed(
        ctx.handle, *_xyz(pos_x, pos_y, pos_z, n), ptr(force, _F32, "force", numel=3 * n),
        float(theta), float(G), float(soft), float(box), ptr(root, torch.int32, "root"), off,
        count, n, ctx.packed_epoch, stream_handle()), "compute_forces")


def traverse(pos_x, pos_y, pos_z, nodes, root, force, theta, G, soft, box, off=0, count=None,
             mac=MAC_REF, theta_group=None, compensated=False, stats=None, n_nodes=None,
             pairs=False, group_stats=False, group_order=None, straddle_mass=0.0,
             straddle_group=None, group_pops=None):
    """Traversal over packed node records; stats: int64[4] tensor accumulating
    (interactions, node visits, wavefront stack pops, overflow drops) when
    given; with group_stats it has 4 + 3*W entries and entries 4+3w.. hold
    wavefront group w's own interactions, visits and pops.  The targets (pos_*) need not be the tree's particles:
    n_nodes gives the tree size when they differ.  pairs: two particles per
    lane (128-particle groups).  group_order: int32 launch order of the
    64-particle groups (a permutation, e.g. from order_groups).
    straddle_mass (AABB test): nodes lighter than straddle_mass pass the
    minimum-image guard (0 = exact guard); straddle_group (float32 per
    64-particle group) gives each group its own.  group_pops (int32/uint32
    per 64-particle group) receives each group's stack pops, from which
    order_groups_by_pops makes a longest-first order for a later launch."""
    n = pos_x.shape[0]
    n_nodes = 2 * n - 1 if n_nodes is None else int(n_nodes)
    n_tree = (n_nodes + 1) // 2
    count = n - off if count is None else count
    flags = ((TRAV_COMPENSATED if compensated else 0) | (TRAV_STATS if stats is not None else 0) |
             (TRAV_PAIRS if pairs else 0) | (TRAV_GROUP_STATS if group_stats else 0))
    groups = (min(n, off + count) - off + 63) // 64
    waves = (min(n, off + count) - off + (127 if pairs else 63)) // (128 if pairs else 64)
    if group_stats and stats is None:
        raise ValueError("group_stats needs a stats tensor")
    tg = None if theta_group is None else ptr(theta_group, _F32, "theta_group",
                                              numel=max(groups, 0))
    st = None if stats is None else ptr(stats, (torch.int64, torch.uint64), "stats",
                                        numel=4 + (3 * max(waves, 0) if group_stats else 0))
    args = (*_xyz(pos_x, pos_y, pos_z, n), ptr(nodes, torch.uint8, "nodes"), n_tree,
            ptr(root, torch.int32, "root"), ptr(force, _F32, "force", numel=3 * n),
            float(theta), tg, float(G), float(soft), float(box), int(off), int(count), n,
            int(mac), flags, st)
    order = None if group_order is None else ptr(group_order, torch.int32, "group_order",
                                                 numel=max(groups, 0))
    sg = None if straddle_group is None else ptr(straddle_group, _F32, "straddle_group",
                                                 numel=max(groups, 0))
    gp = None if group_pops is None else ptr(group_pops, U32, "group_pops",
                                             numel=max(groups, 0))
    check(load().cosmo_traverse_ex(*args, order, float(straddle_mass), sg, gp, stream_handle()),
          "traverse")


def traverse_split(pos_x, pos_y, pos_z, nodes, root, force, theta, G, soft, box, off=0,
                   count=None, stats=None):
    """The reference-test walk for a small launch (at most 65536 particles):
    every 64-particle group gets one wavefront per subtree below a small top
    tree (cosmo_traverse_split).  Same visited and accepted nodes as
    traverse(); the per-subtree sums are added in subtree order.  stats:
    int64[2] accumulating (interactions, node visits)."""
    n = pos_x.shape[0]
    count = n - off if count is None else count
    ctx = shared_context(n)
    st = None if stats is None else ptr(stats, (torch.int64, torch.uint64), "stats", numel=2)
    check(load().cosmo_traverse_split(
        ctx.handle, *_xyz(pos_x, pos_y, pos_z, n), ptr(nodes, torch.uint8, "nodes"), n,
        ptr(root, torch.int32, "root"), ptr(force, _F32, "force", numel=3 * n), float(theta),
        float(G), float(soft), float(box), int(off), int(count), n, st, stream_handle()),
        "traverse_split")


def order_groups_by_pops(group_pops, out=None, k_now=None, k_prev=None, cost=None):
    """Launch order of the groups by decreasing pop count (longest first).
    With a per-group selection, k_now / k_prev (int32 per group: the chosen
    candidate now and at the counting launch) and cost (float32 per
    candidate) scale each group's pops by cost[k_now] / cost[k_prev]."""
    g = group_pops.shape[0]
    out = torch.empty(g, dtype=torch.int32, device=group_pops.device) if out is None else out
    if (k_now is None) != (k_prev is None) or (k_now is not None and cost is None):
        raise ValueError("order_groups_by_pops: k_now, k_prev and cost go together")
    K = 0 if cost is None else cost.shape[0]
    check(load().cosmo_order_groups_by_pops(
        ptr(group_pops, U32, "group_pops", numel=g), g,
        None if k_now is None else ptr(k_now, torch.int32, "k_now", numel=g),
        None if k_prev is None else ptr(k_prev, torch.int32, "k_prev", numel=g),
        None if cost is None else ptr(cost, _F32, "cost", numel=K), K,
        ptr(out, torch.int32, "order", numel=g), stream_handle()), "order_groups_by_pops")
    return out


def order_groups(theta_group, out=None):
    """Launch order of the groups by increasing theta (longest first)."""
    g = theta_group.shape[0]
    out = torch.empty(g, dtype=torch.int32, device=theta_group.device) if out is None else out
    check(load().cosmo_order_groups(ptr(theta_group, _F32, "theta_group"), g,
                                    ptr(out, torch.int32, "order", numel=g), stream_handle()),
          "order_groups")
    return out


@kernel
def integrate(pos_x, pos_y, pos_z, vel_x, vel_y, vel_z, force, a, H, dt, _launch=None):
    n = pos_x.shape[0] if _launch is None else min(pos_x.shape[0], _launch[0] * _launch[1])
    v = [ptr(t, _F32, nm, numel=n) for t, nm in ((vel_x, "vel_x"), (vel_y, "vel_y"),
                                                  (vel_z, "vel_z"))]
    check(load().cosmo_integrate(*_xyz(pos_x, pos_y, pos_z, n), *v,
                                 ptr(force, _F32, "force", numel=3 * n), float(a), float(H),
                                 float(dt), n, stream_handle()), "integrate")


def kick_drift(pos, vel, force, c_drag, c_kick, c_drift):
    n = pos[0].shape[0]
    v = [ptr(t, _F32, "vel", numel=n) for t in vel]
    check(load().cosmo_kick_drift(*_xyz(*pos, n), *v, ptr(force, _F32, "force", numel=3 * n),
                                  float(c_drag), float(c_kick), float(c_drift), n,
                                  stream_handle()), "kick_drift")


def direct_sum(pos_x, pos_y, pos_z, mass, targets, G, soft, out=None):
    """FP64 direct-sum accelerations float64[t, 3] for the target indices."""
    n = pos_x.shape[0]
    targets = targets.to(device=pos_x.device, dtype=torch.int64).contiguous()
    nt = targets.shape[0]
    out = torch.empty((nt, 3), dtype=torch.float64, device=pos_x.device) if out is None else out
    check(load().cosmo_direct_sum(*_xyz(pos_x, pos_y, pos_z, n), ptr(mass, _F32, "mass", numel=n),
                                  n, ptr(targets, torch.int64, "targets"), nt, float(G),
                                  float(soft), ptr(out, torch.float64, "out", numel=3 * nt),
                                  stream_handle()), "direct_sum")
    return out


def count_nonfinite(values, count):
    """count[0] += number of non-finite entries of the float32 tensor values
    (count: int64[1] on the device; no host synchronisation)."""
    check(load().cosmo_count_nonfinite(ptr(values, _F32, "values"), values.numel(),
                                       ptr(count, (torch.int64, torch.uint64), "count", numel=1),
                                       stream_handle()), "count_nonfinite")
