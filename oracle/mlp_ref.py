"""CPU mirror of the learned-theta MLP and its group features.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker for the MFMA
kernels in cosmo-nbody_amd/csrc/mlp.hip.  The network has no reference
counterpart.  Arithmetic is FP64; with ``rounding=True`` the BF16 roundings
the GPU applies are inserted at the same points (inputs, weights, the
layer-1 activations, and the layer gradients fed to the MFMA products), so
the GPU and this mirror differ only by FP32 accumulation order.  With
``rounding=False`` it is a plain FP64 network whose gradient is pinned by
central finite differences (tests/test_mlp.py).

Parameter vector layout (kParams = 5313 floats): W1[16][64] (rows 10..15
zero), b1[64], W2[64][64], b2[64], W3[64], b3.
"""
from __future__ import annotations

import numpy as np

IN, IN_PAD, HID = 11, 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN_PAD * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
N_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round to BF16 (nearest even) and return as float64 (via FP32)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(nan, (u.astype(np.uint32) | 0x400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def bf16_bits(x):
    """BF16 bit patterns (uint16) of x, nearest even."""
    return (bf16(x).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def unpack(p):
    p = np.asarray(p, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(IN_PAD, HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(HID, HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, W3, b3


def init_params(seed=0):
    """He-style initialisation; W1 rows >= 10 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    W1 = rng.standard_normal((IN_PAD, HID)) * np.sqrt(2.0 / IN)
    W1[IN:] = 0
    p[OFF_W1:OFF_B1] = W1.reshape(-1)
    p[OFF_W2:OFF_B2] = (rng.standard_normal((HID, HID)) * np.sqrt(2.0 / HID)).reshape(-1)
    p[OFF_W3:OFF_B3] = rng.standard_normal(HID) * np.sqrt(1.0 / HID)
    return p


def forward(X, p, rounding=True):
    """X: [B, 16] (values already BF16-representable when rounding)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    X = np.asarray(X, np.float64)
    z1 = X @ r(W1) + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2) + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ r(W3) + b3
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(X, T, p, rounding=True):
    """mean((y - T)^2) and its gradient (vector of N_PARAMS), plus the sum of
    squared errors."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    y, (X, z1, h1, z2, h2) = forward(X, p, rounding)
    B = len(y)
    err = y - np.asarray(T, np.float64)
    dy = 2.0 * err / B
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, r(W3)) * (z2 > 0)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2r = r(dz2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2r).reshape(-1)
    dh1 = dz2r @ r(W2).T
    dz1 = dh1 * (z1 > 0)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    gW1 = X.T @ r(dz1)
    gW1[IN:] = 0
    g[OFF_W1:OFF_B1] = gW1.reshape(-1)
    return float(np.mean(err ** 2)), g, float(np.sum(err ** 2))


def straddle_profile(parents, n_mass, n_min, n_max, straddle_mass, nbins=1024):
    """Mirror of cosmo_straddle_profile: uint64 [2, 3, nbins] sums of
    (M/limit)^2 and M/limit (units of 2^-32) over the straddle-level nodes
    (lighter than the limit, parent not) spanning each bin."""
    n = (len(n_mass) + 1) // 2
    prof = np.zeros((2, 3, nbins), np.uint64)
    k = np.arange(n, 2 * n - 1)
    m = np.asarray(n_mass, np.float64)[k]
    par = np.asarray(parents)[k]
    pm = np.where(par >= 0, np.asarray(n_mass, np.float64)[np.maximum(par, 0)], np.inf)
    sel = (m < straddle_mass) & (pm >= straddle_mass)
    k, m = k[sel], m[sel]
    r = m / straddle_mass
    w = [(r * r * 4294967296.0).astype(np.uint64), (r * 4294967296.0).astype(np.uint64)]
    for ax in range(3):
        b0 = (np.asarray(n_min, np.float32)[k, ax] * np.float32(nbins)).astype(np.int64)
        b1 = (np.asarray(n_max, np.float32)[k, ax] * np.float32(nbins)).astype(np.int64)
        b0 = np.clip(b0, 0, nbins - 1)
        b1 = np.clip(b1, 0, nbins - 1)
        for q in range(2):
            for j in range(len(k)):
                prof[q, ax, b0[j]:b1[j] + 1] += w[q][j]
    return prof


def group_features(pos, children, parents, n_mass, n_com, n_min, n_max, a, n_grid,
                   profile=None):
    """Mirror of k_group_features (FP32 arithmetic where the kernel uses it);
    profile: straddle_profile's output, or None."""
    px, py, pz = (np.asarray(v, np.float32) for v in pos)
    n = len(px)
    G = (n + 63) // 64
    out = np.zeros((G, 9), np.float32)
    f32 = np.float32
    floor_len = f32(1.0) / (f32(64.0) * f32(n_grid))
    for g in range(G):
        b0, b1 = 64 * g, min(64 * g + 64, n)
        cnt = b1 - b0
        P = np.stack([px[b0:b1], py[b0:b1], pz[b0:b1]], 1)
        lo, hi = P.min(0), P.max(0)
        e = (hi - lo).astype(np.float32)
        ext = f32(max(f32(0), *e))
        vol = f32(1.0)
        for k in range(3):
            vol = f32(vol * (e[k] if e[k] > floor_len else floor_len))
        cen = (f32(0.5) * (hi + lo)).astype(np.float32)
        out[g, 0] = np.log2(max(f32(ext * f32(n_grid)), f32(1e-6)))
        out[g, 1] = np.log2(f32(f32(cnt) / f32(vol * f32(n))))
        out[g, 2] = a
        need = (cnt - 0.5) / n
        node = b0
        while n_mass[node] < need:
            p = parents[node]
            if p < 0:
                break
            node = p
        for lvl in range(2):
            p = parents[node]
            r = mlog = 0.0
            if p >= 0:
                c0, c1 = children[p]
                s = c1 if c0 == node else c0
                d = np.array([n_com[s, k] for k in range(3)], np.float32) - cen
                d = (d - np.rint(d)).astype(np.float32)
                dist = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
                es = (n_max[s] - n_min[s]).astype(np.float32)
                r = f32(es.max()) / max(dist, f32(1e-6))
                mlog = np.log2(max(f32(f32(n_mass[s]) * f32(n) / f32(cnt)), f32(1e-6)))
                node = p
            out[g, 3 + 2 * lvl] = r
            out[g, 4 + 2 * lvl] = mlog
        s2 = s1 = f32(0)
        if profile is not None:
            nb = profile.shape[2]
            for ax in range(3):
                x = f32(cen[ax] + f32(0.5))
                x = f32(x - np.floor(x))
                b = min(max(int(f32(x * f32(nb))), 0), nb - 1)
                s2 = f32(s2 + f32(profile[0, ax, b]))
                s1 = f32(s1 + f32(profile[1, ax, b]))
        out[g, 7] = np.log2(f32(1) + f32(s2 * f32(2.3283064365386963e-10)))
        out[g, 8] = np.log2(f32(1) + f32(s1 * f32(2.3283064365386963e-10)))
    return out
