"""Learned per-group opening angle (no reference counterpart; the reference
uses one theta, src/utils/config.py:31).

A group is 64 Morton-consecutive particles, i.e. one wavefront of the
traversal.  For each group and each candidate theta_k the MLP predicts
y = log2(mean over the group of |a_BH - a_direct|^2 / mean over all sampled
targets of |a_direct|^2), the group's share of the squared relative RMS
error.  The step then picks, per group, the theta_k that minimises
cost_k + lambda * 2^y, where cost_k is the measured number of interactions
per particle at theta_k; lambda sets the error budget and is calibrated
against the FP64 direct sum.  Non-finite predictions fall back to the fixed
theta, so with poisoned weights the step is the plain method.

Everything runs on the GPU with the project's own kernels: features
(cosmo_group_features), rows (cosmo_assemble_rows), the MFMA MLP forward
and backward (cosmo_mlp_forward / cosmo_mlp_grad), Adam (cosmo_mlp_adam),
selection (cosmo_select_theta), the traversal and the FP64 direct sum.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from src.kernels import mlp_kernels as K
from src.kernels.physics_kernels import direct_sum, traverse

DEFAULT_THETAS = (0.15, 0.2, 0.25, 0.3, 0.35, 0.4, 0.45, 0.5, 0.55, 0.6, 0.65, 0.7, 0.8, 0.9,
                  1.0, 1.2)


def init_params(seed=0, device="cuda"):
    """He initialisation of the 10-64-64-1 network (W1 rows >= 10 zero)."""
    g = torch.Generator().manual_seed(seed)
    p = torch.zeros(K.PARAMS, dtype=torch.float32)
    w1 = torch.randn(K.IN_PAD, K.HIDDEN, generator=g) * math.sqrt(2.0 / K.IN)
    w1[K.IN:] = 0
    o = 0
    p[o:o + K.IN_PAD * K.HIDDEN] = w1.reshape(-1)
    o += K.IN_PAD * K.HIDDEN + K.HIDDEN
    p[o:o + K.HIDDEN * K.HIDDEN] = (torch.randn(K.HIDDEN, K.HIDDEN, generator=g) *
                                    math.sqrt(2.0 / K.HIDDEN)).reshape(-1)
    o += K.HIDDEN * K.HIDDEN + K.HIDDEN
    p[o:o + K.HIDDEN] = torch.randn(K.HIDDEN, generator=g) * math.sqrt(1.0 / K.HIDDEN)
    return p.to(device)


class ThetaModel:
    """MLP parameters (FP32 master + BF16 copy) and Adam state on the GPU."""

    def __init__(self, params=None, seed=0, device="cuda"):
        self.params = init_params(seed, device) if params is None else params.to(device).float()
        self.wbf = torch.empty(K.BF16_WEIGHTS, dtype=torch.int16, device=device)
        K.to_bf16(self.params, self.wbf)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._partial = torch.empty(K.GRAD_BLOCKS * (K.PARAMS + 1), dtype=torch.float32,
                                    device=device)
        self._grad = torch.empty(K.PARAMS + 1, dtype=torch.float32, device=device)

    def predict(self, X, out=None):
        return K.forward(X, self.wbf, self.params, out)

    def loss(self, X, T):
        g = K.grad(X, T, self.wbf, self.params, self._partial, self._grad)
        return float(g[-1].item()) / X.shape[0]

    def fit(self, X, T, steps=2000, batch=8192, lr=2e-3, seed=0, log=None):
        """Minibatch Adam on (X, T); returns the list of logged losses."""
        n = X.shape[0]
        gen = torch.Generator(device=X.device).manual_seed(seed)
        hist = []
        for s in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=gen)
            xb, tb = X[idx].contiguous(), T[idx].contiguous()
            g = K.grad(xb, tb, self.wbf, self.params, self._partial, self._grad)
            self.step += 1
            lr_s = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
            K.adam(self.params, self.m, self.v, g, lr_s, self.step, self.wbf)
            if log is not None and (s % log == 0 or s == steps - 1):
                hist.append((s, float(g[-1].item()) / xb.shape[0]))
        return hist


@dataclasses.dataclass
class Samples:
    X: torch.Tensor        # int16 [groups*K, 16] BF16 rows
    T: torch.Tensor        # float32 [groups*K] labels log2(relative squared error)
    err2: torch.Tensor     # float64 [groups, K] group squared-error sums
    cost: np.ndarray       # interactions per particle at each theta
    norm: float            # sum over targets of |a_direct|^2
    groups: np.ndarray     # sampled group ids


def _targets_of(groups, n):
    idx = (np.asarray(groups)[:, None] * 64 + np.arange(64)[None, :]).reshape(-1)
    valid = idx < n
    return idx, valid


def gather_groups(pos, groups):
    """Positions of whole sampled groups, 64-aligned (padding repeats the last
    particle of a partial group; padded lanes are ignored by the caller)."""
    n = pos[0].shape[0]
    idx, valid = _targets_of(groups, n)
    idx = np.where(valid, idx, np.minimum(idx, n - 1))
    it = torch.from_numpy(idx).to(pos[0].device)
    return [p.index_select(0, it).contiguous() for p in pos], idx, valid


def group_errors(engine, groups, thetas, theta_group=None):
    """BH forces of the sampled groups at each theta (or one per-group theta
    vector) against the FP64 direct sum.  Returns (err2[groups, K],
    inter[K] per particle, norm = sum |a_direct|^2 over valid targets)."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    n = st.n
    tpos, idx, valid = gather_groups(st.pos, groups)
    T = tpos[0].shape[0]
    tidx = torch.from_numpy(idx[valid]).to(st.pos[0].device)
    a_ref = direct_sum(*st.pos, st.mass, tidx, c.G, c.SOFTENING)
    vmask = torch.from_numpy(valid).to(a_ref.device)
    norm = float((a_ref ** 2).sum().item())
    force = torch.empty((T, 3), dtype=torch.float32, device=a_ref.device)
    stats = torch.zeros(2, dtype=torch.int64, device=a_ref.device)
    runs = [(th, None) for th in thetas] if theta_group is None else [(0.5, theta_group)]
    err2, inter = [], []
    for th, tg in runs:
        stats.zero_()
        traverse(*tpos, tm.nodes, tm.root_idx, force, th, c.G, c.SOFTENING, c.BOX_SIZE, mac=engine.mac,
                 theta_group=tg, stats=stats, n_nodes=tm.n_nodes)
        d = torch.zeros((T, 3), dtype=torch.float64, device=a_ref.device)
        d[vmask] = force[vmask].double() - a_ref
        e = (d ** 2).sum(1).reshape(len(groups), 64).sum(1)
        err2.append(e)
        inter.append(float(stats[0].item()) / T)
    return torch.stack(err2, 1), np.array(inter), norm


def sample(engine, n_groups, thetas=DEFAULT_THETAS, seed=0):
    """Training samples from the engine's current tree (built)."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    G = (st.n + 63) // 64
    rng = np.random.default_rng(seed)
    groups = np.sort(rng.choice(G, size=min(n_groups, G), replace=False))
    feat = K.group_features(st.pos, tm, engine.a, c.N_GRID)
    th = torch.tensor(thetas, dtype=torch.float32, device=feat.device)
    X = K.assemble_rows(feat[torch.from_numpy(groups).to(feat.device)].contiguous(), th)
    err2, cost, norm = group_errors(engine, groups, thetas)
    counts = np.minimum(64, st.n - groups * 64).astype(np.float64)
    mean_norm = norm / counts.sum()
    cnt = torch.from_numpy(counts).to(err2.device)[:, None]
    T = torch.log2(err2 / cnt / mean_norm + 1e-30).float().reshape(-1).contiguous()
    return Samples(X, T, err2, cost, norm, groups)


class LearnedTheta:
    """Selects engine.theta_group each step from the model's predictions."""

    def __init__(self, model, thetas, cost, lam, fallback):
        self.model = model
        self.thetas_list = list(thetas)
        self.cost_np = np.asarray(cost, np.float32)
        self.lam = float(lam)
        self.fallback = float(fallback)
        self._dev = None

    def _ensure(self, device, G):
        if self._dev is None or self._dev[0] != G:
            th = torch.tensor(self.thetas_list, dtype=torch.float32, device=device)
            cost = torch.from_numpy(self.cost_np).to(device)
            Kc = len(self.thetas_list)
            X = torch.empty((G * Kc, K.IN_PAD), dtype=torch.int16, device=device)
            Y = torch.empty(G * Kc, dtype=torch.float32, device=device)
            tg = torch.empty(G, dtype=torch.float32, device=device)
            nf = torch.zeros(1, dtype=torch.int64, device=device)
            self._dev = (G, th, cost, X, Y, tg, nf)
        return self._dev

    def theta_groups(self, engine, lam=None):
        st, tm, c = engine.state, engine.tree_manager, engine.config
        G = (st.n + 63) // 64
        _, th, cost, X, Y, tg, nf = self._ensure(st.pos[0].device, G)
        feat = K.group_features(st.pos, tm, engine.a, c.N_GRID)
        K.assemble_rows(feat, th, X)
        self.model.predict(X, Y)
        K.select_theta(Y, G, th, cost, self.lam if lam is None else lam, self.fallback, tg, nf)
        return tg

    @property
    def n_fallback(self):
        return 0 if self._dev is None else int(self._dev[6].item())

    def save(self, path):
        np.savez(path, params=self.model.params.cpu().numpy(), thetas=np.array(self.thetas_list),
                 cost=self.cost_np, lam=self.lam, fallback=self.fallback)

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path, allow_pickle=False)
        model = ThetaModel(torch.from_numpy(d["params"]), device=device)
        return cls(model, d["thetas"].tolist(), d["cost"], float(d["lam"]), float(d["fallback"]))
