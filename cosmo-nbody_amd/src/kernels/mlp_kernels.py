"""Learned opening angle: bindings of the MFMA MLP kernels and of the group
features / theta selection (include/cosmo_mlp.h).  No reference counterpart.

Tensors: X int16 [rows, 16] (BF16 bit patterns); parameters float32
[PARAMS]; BF16 weights int16 [BF16_WEIGHTS]; features float32 [groups, 9].
"""
from __future__ import annotations

import torch

from ._hip import check, load, ptr, stream_handle

IN, IN_PAD, HIDDEN, GROUP_FEATURES = 11, 16, 64, 9
PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1
BF16_WEIGHTS = 16 * 64 + 64 * 64 + 64
GRAD_BLOCKS = 64

_F32, _I16 = torch.float32, torch.int16


PROFILE_BINS = 1024
MAX_PROFILES = 8


def straddle_profile(tree, straddle_mass, profile=None, scratch=None, nbins=PROFILE_BINS):
    """Straddle profile (uint64 [2, 3, nbins]) of the tree for the AABB
    test's straddle limit; see cosmo_straddle_profile."""
    n = tree.n
    M = 2 * n - 1
    dev = tree.child.device
    profile = (torch.empty(6 * nbins, dtype=torch.int64, device=dev) if profile is None
               else profile)
    scratch = (torch.empty(6 * nbins, dtype=torch.int64, device=dev) if scratch is None
               else scratch)
    check(load().cosmo_straddle_profile(
        ptr(tree.parent, torch.int32, "parents", numel=M),
        ptr(tree.mass, torch.float64, "n_mass", numel=M),
        ptr(tree.aabb_min, _F32, "n_min", numel=3 * M),
        ptr(tree.aabb_max, _F32, "n_max", numel=3 * M), n, float(straddle_mass),
        ptr(profile, (torch.int64, torch.uint64), "profile", numel=6 * nbins),
        ptr(scratch, (torch.int64, torch.uint64), "scratch", numel=6 * nbins), int(nbins),
        stream_handle()), "straddle_profile")
    return profile


def group_features(pos, tree, a, n_grid, out=None, profile=None, nbins=PROFILE_BINS):
    """Nine features per group of 64 sorted particles (tree: OctreeManager).
    a: the scale factor, a float or a device float tensor (read at run time,
    as a captured step needs).  profile: the straddle profile (features 7
    and 8), or None (both 0)."""
    n = pos[0].shape[0]
    G = (n + 63) // 64
    out = torch.empty((G, GROUP_FEATURES), dtype=_F32, device=pos[0].device) if out is None else out
    M = 2 * n - 1
    on_dev = isinstance(a, torch.Tensor)
    fn = load().cosmo_group_features_dev if on_dev else load().cosmo_group_features
    a_arg = ptr(a, _F32, "a", numel=1) if on_dev else float(a)
    prof = None if profile is None else ptr(profile, (torch.int64, torch.uint64), "profile",
                                            numel=6 * nbins)
    check(fn(
        *[ptr(t, _F32, "pos", numel=n) for t in pos],
        ptr(tree.child, torch.int32, "children", numel=2 * M),
        ptr(tree.parent, torch.int32, "parents", numel=M),
        ptr(tree.mass, torch.float64, "n_mass", numel=M),
        ptr(tree.com, torch.float64, "n_com", numel=3 * M),
        ptr(tree.aabb_min, _F32, "n_min", numel=3 * M),
        ptr(tree.aabb_max, _F32, "n_max", numel=3 * M), n, a_arg, float(n_grid), prof,
        int(nbins), ptr(out, _F32, "feat", numel=G * GROUP_FEATURES), stream_handle()),
        "group_features")
    return out


def straddle_profiles(tree, straddle_masses, profiles=None, scratch=None, nbins=PROFILE_BINS):
    """Straddle profiles (uint64 [J, 2, 3, nbins]) of the tree for J straddle
    limits at once (one pass over the nodes); block j equals
    straddle_profile(tree, straddle_masses[j]).  See cosmo_straddle_profiles."""
    import ctypes
    n = tree.n
    M = 2 * n - 1
    J = len(straddle_masses)
    dev = tree.child.device
    profiles = (torch.empty(J * 6 * nbins, dtype=torch.int64, device=dev) if profiles is None
                else profiles)
    scratch = (torch.empty(J * 6 * nbins, dtype=torch.int64, device=dev) if scratch is None
               else scratch)
    lim = (ctypes.c_double * J)(*[float(m) for m in straddle_masses])
    check(load().cosmo_straddle_profiles(
        ptr(tree.parent, torch.int32, "parents", numel=M),
        ptr(tree.mass, torch.float64, "n_mass", numel=M),
        ptr(tree.aabb_min, _F32, "n_min", numel=3 * M),
        ptr(tree.aabb_max, _F32, "n_max", numel=3 * M), n, ctypes.addressof(lim), J,
        ptr(profiles, (torch.int64, torch.uint64), "profiles", numel=J * 6 * nbins),
        ptr(scratch, (torch.int64, torch.uint64), "scratch", numel=J * 6 * nbins), int(nbins),
        stream_handle()), "straddle_profiles")
    return profiles


def group_features_multi(pos, tree, a, n_grid, profiles, count, out=None, nbins=PROFILE_BINS):
    """group_features for `count` profiles (straddle_profiles) at once:
    float32 [count, groups, 9], block j as group_features with profile j.
    a: the scale factor, a float or a device float tensor."""
    n = pos[0].shape[0]
    G = (n + 63) // 64
    out = (torch.empty((count, G, GROUP_FEATURES), dtype=_F32, device=pos[0].device)
           if out is None else out)
    M = 2 * n - 1
    on_dev = isinstance(a, torch.Tensor)
    check(load().cosmo_group_features_multi(
        *[ptr(t, _F32, "pos", numel=n) for t in pos],
        ptr(tree.child, torch.int32, "children", numel=2 * M),
        ptr(tree.parent, torch.int32, "parents", numel=M),
        ptr(tree.mass, torch.float64, "n_mass", numel=M),
        ptr(tree.com, torch.float64, "n_com", numel=3 * M),
        ptr(tree.aabb_min, _F32, "n_min", numel=3 * M),
        ptr(tree.aabb_max, _F32, "n_max", numel=3 * M), n, 0.0 if on_dev else float(a),
        ptr(a, _F32, "a", numel=1) if on_dev else None, float(n_grid),
        ptr(profiles, (torch.int64, torch.uint64), "profiles", numel=count * 6 * nbins),
        int(count), int(nbins), ptr(out, _F32, "feat", numel=count * G * GROUP_FEATURES),
        stream_handle()), "group_features_multi")
    return out


def assemble_rows(feat, thetas, out=None, cand2=None, fidx=None):
    """MLP input rows [theta_k, features, cand2_k] for every group and
    candidate k (cand2: the candidates' straddle-mass codes, or None: 0).
    feat: float32 [groups, 9], or [blocks, groups, 9] with fidx (int32 [K]):
    candidate k takes its features from block fidx[k]."""
    G = feat.shape[-2]
    K = thetas.shape[0]
    out = torch.empty((G * K, IN_PAD), dtype=_I16, device=feat.device) if out is None else out
    c2 = None if cand2 is None else ptr(cand2, _F32, "cand2", numel=K)
    fi = None if fidx is None else ptr(fidx, torch.int32, "fidx", numel=K)
    if fidx is None and feat.dim() != 2:
        raise ValueError("assemble_rows: blocks of features need fidx")
    check(load().cosmo_assemble_rows(ptr(feat, _F32, "feat"), G, ptr(thetas, _F32, "thetas"), c2,
                                     fi, K, ptr(out, _I16, "X", numel=G * K * IN_PAD),
                                     stream_handle()), "assemble_rows")
    return out


def forward(X, wbf, params, out=None):
    rows = X.shape[0]
    out = torch.empty(rows, dtype=_F32, device=X.device) if out is None else out
    check(load().cosmo_mlp_forward(ptr(X, _I16, "X", numel=rows * IN_PAD), rows,
                                   ptr(wbf, _I16, "wbf", numel=BF16_WEIGHTS),
                                   ptr(params, _F32, "params", numel=PARAMS),
                                   ptr(out, _F32, "Y", numel=rows), stream_handle()),
          "mlp_forward")
    return out


def grad(X, T, wbf, params, partial=None, out=None, nblocks=GRAD_BLOCKS):
    """Gradient of mean((MLP(X) - T)^2); out[PARAMS] = sum of squared errors."""
    rows = X.shape[0]
    dev = X.device
    partial = (torch.empty(nblocks * (PARAMS + 1), dtype=_F32, device=dev)
               if partial is None else partial)
    out = torch.empty(PARAMS + 1, dtype=_F32, device=dev) if out is None else out
    check(load().cosmo_mlp_grad(ptr(X, _I16, "X", numel=rows * IN_PAD), ptr(T, _F32, "T"), rows,
                                ptr(wbf, _I16, "wbf", numel=BF16_WEIGHTS),
                                ptr(params, _F32, "params", numel=PARAMS),
                                ptr(partial, _F32, "partial", numel=nblocks * (PARAMS + 1)),
                                int(nblocks), ptr(out, _F32, "grad", numel=PARAMS + 1),
                                stream_handle()), "mlp_grad")
    return out


def adam(params, m, v, g, lr, step, wbf, beta1=0.9, beta2=0.999, eps=1e-8):
    check(load().cosmo_mlp_adam(ptr(params, _F32, "params"), ptr(m, _F32, "m"), ptr(v, _F32, "v"),
                                ptr(g, _F32, "grad"), float(lr), float(beta1), float(beta2),
                                float(eps), int(step), ptr(wbf, _I16, "wbf"), stream_handle()),
          "mlp_adam")


def to_bf16(params, wbf):
    check(load().cosmo_mlp_to_bf16(ptr(params, _F32, "params", numel=PARAMS),
                                   ptr(wbf, _I16, "wbf", numel=BF16_WEIGHTS), stream_handle()),
          "mlp_to_bf16")


def select_theta(pred, groups, thetas, cost, lam, fallback, out, n_fallback=None, mus=None,
                 fallback_mu=0.0, mu_out=None, k_out=None):
    """Per-group theta (and, with mus and mu_out, straddle mass) of the
    candidate minimising cost + lam * 2^pred; k_out (int32 per group) gets
    the chosen candidate's index, -1 for a fallback."""
    K = thetas.shape[0]
    if (mus is None) != (mu_out is None):
        raise ValueError("select_theta: mus and mu_out go together")
    check(load().cosmo_select_theta(ptr(pred, _F32, "pred", numel=groups * K), int(groups),
                                    ptr(thetas, _F32, "thetas", numel=K),
                                    None if mus is None else ptr(mus, _F32, "mus", numel=K),
                                    ptr(cost, _F32, "cost", numel=K), K, float(lam),
                                    float(fallback), float(fallback_mu),
                                    ptr(out, _F32, "theta_group", numel=groups),
                                    None if mu_out is None else
                                    ptr(mu_out, _F32, "mu_group", numel=groups),
                                    None if k_out is None else
                                    ptr(k_out, torch.int32, "k_group", numel=groups),
                                    None if n_fallback is None else
                                    ptr(n_fallback, (torch.int64, torch.uint64), "n_fallback"),
                                    stream_handle()), "select_theta")
    return out
