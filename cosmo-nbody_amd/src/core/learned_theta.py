"""Learned per-group opening angle (no reference counterpart; the reference
uses one theta, src/utils/config.py:31).

A group is 64 Morton-consecutive particles, i.e. one wavefront of the
traversal.  For each group and each candidate theta_k the MLP predicts
y = log2(mean over the group of |a_BH - a_direct|^2 / mean over all sampled
targets of |a_direct|^2), the group's share of the squared relative RMS
error.  The step then picks, per group, the theta_k that minimises
cost_k + lambda * 2^y, where cost_k is the measured cost of theta_k: the
traversal time of one group at theta_k (the whole-state traversal at fixed
theta_k divided by the number of groups), or the interactions per particle;
lambda sets the error budget and is calibrated against the FP64 direct sum.
Non-finite predictions fall back to the fixed theta, so with poisoned
weights the step is the plain method.

Everything runs on the GPU with the project's own kernels: features
(cosmo_group_features), rows (cosmo_assemble_rows), the MFMA MLP forward
and backward (cosmo_mlp_forward / cosmo_mlp_grad), Adam (cosmo_mlp_adam),
selection (cosmo_select_theta), the traversal and the FP64 direct sum.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np
import torch

from src.kernels import mlp_kernels as K
from src.kernels.physics_kernels import direct_sum, order_groups, traverse

N_MODELS = 4           # networks averaged by the selection
WAVES_PER_SIMD = 8     # resident wavefronts per SIMD of the traversal kernel (its VGPR use)
WAVE_BUDGET = 0.5      # largest single-wavefront time allowed, as a fraction of the fixed traversal
DEFAULT_THETAS = (0.04, 0.05, 0.06, 0.07, 0.08, 0.1, 0.12, 0.15, 0.2, 0.25, 0.3, 0.4, 0.5, 0.6,
                  0.8, 1.0, 1.4, 2.0)
LABEL_FLOOR = -40.0   # log2 of the smallest squared-error share a label keeps
# Candidates of the joint selection: opening angles x straddle masses (the
# configuration's straddle mass times these factors: 1/4 to 4 in steps of
# sqrt(2) above 1/2).
PAIR_THETAS = (0.1, 0.12, 0.15, 0.2, 0.25, 0.3, 0.4, 0.5, 0.6, 0.8, 1.0, 1.4)
MU_FACTORS = (0.25, 0.5, 0.7071, 1.0, 1.4142, 2.0, 2.8284, 4.0)
FIT_STEPS = 160000     # Adam steps per network of the ensemble


def candidates(c, thetas=None, pairs=True, mu_factors=None):
    """(thetas, mus) of the candidates for configuration c: every theta with
    every straddle mass c.STRADDLE_MASS * mu_factors (default MU_FACTORS)
    when pairs is set and the AABB test has a straddle mass, else each theta
    with c.STRADDLE_MASS."""
    mu0 = float(getattr(c, "STRADDLE_MASS", 0.0))
    if pairs and mu0 > 0:
        th = PAIR_THETAS if thetas is None else tuple(thetas)
        fs = MU_FACTORS if mu_factors is None else tuple(mu_factors)
        return ([t for _ in fs for t in th], [mu0 * f for f in fs for _ in th])
    th = DEFAULT_THETAS if thetas is None else tuple(thetas)
    return list(th), [mu0] * len(th)


def feature_blocks(engine, mus_l, feat=None, prof=None, a=None):
    """Group features computed with the straddle profile of each distinct
    straddle mass among the candidates' (mus_l): (feat float[J, G, 9], fidx
    list of the block of each candidate).  feat and prof (int64 [2, J,
    6 * PROFILE_BINS]: the profiles, then their scratch) are reused when
    given.  With every distinct straddle mass positive, the J profiles come
    from one pass over the nodes and the J feature blocks from one pass over
    the groups."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    G = (st.n + 63) // 64
    distinct = sorted(set(float(m) for m in mus_l))
    J = len(distinct)
    dev = st.pos[0].device
    if feat is None:
        feat = torch.empty((J, G, K.GROUP_FEATURES), dtype=torch.float32, device=dev)
    if prof is None:
        prof = torch.empty((2, J, 6 * K.PROFILE_BINS), dtype=torch.int64, device=dev)
    a_val = engine.a if a is None else a
    if all(m > 0 for m in distinct) and J <= K.MAX_PROFILES:
        K.straddle_profiles(tm, distinct, prof[0], prof[1])
        K.group_features_multi(st.pos, tm, a_val, c.N_GRID, prof[0], J, out=feat)
    else:
        for j, m in enumerate(distinct):
            profile = K.straddle_profile(tm, m, prof[0, j], prof[1, j]) if m > 0 else None
            K.group_features(st.pos, tm, a_val, c.N_GRID, feat[j], profile=profile)
    return feat, [distinct.index(float(m)) for m in mus_l]


def mu_code(mu, n):
    """MLP input of a straddle mass: log2(1 + mass in particle masses)."""
    return math.log2(1.0 + float(mu) * float(n))


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


class Ensemble:
    """Several networks trained on the same samples from different initial
    weights and minibatch orders; predicts the mean of their log2 errors
    (a non-finite prediction of any member makes the mean non-finite, so
    the fixed-theta fallback still applies).  Allocates only on the first
    call for a given number of rows."""

    def __init__(self, models):
        self.models = list(models)
        self._buf = None

    @property
    def params(self):
        return torch.stack([m.params for m in self.models])

    def predict(self, X, out=None):
        rows = X.shape[0]
        out = self.models[0].predict(X, out)
        if len(self.models) > 1:
            if self._buf is None or self._buf.shape[0] != rows:
                self._buf = torch.empty(rows, dtype=torch.float32, device=X.device)
            for m in self.models[1:]:
                m.predict(X, self._buf)
                out.add_(self._buf)
            out.mul_(1.0 / len(self.models))
        return out


@dataclasses.dataclass
class Samples:
    X: torch.Tensor        # int16 [groups*K, 16] BF16 rows
    T: torch.Tensor        # float32 [groups*K] labels log2(relative squared error)
    err2: torch.Tensor     # float64 [groups, K] group squared-error sums
    cost: np.ndarray       # cost of each theta (ms per group, or interactions per particle)
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


class GroupProbe:
    """Sampled groups of the engine's current tree with their FP64 direct-sum
    accelerations, for repeated error measurements."""

    def __init__(self, engine, groups):
        st, c = engine.state, engine.config
        self.engine = engine
        self.groups = np.asarray(groups)
        self.tpos, idx, valid = gather_groups(st.pos, self.groups)
        dev = st.pos[0].device
        self.T = self.tpos[0].shape[0]
        self.vmask = torch.from_numpy(valid).to(dev)
        tidx = torch.from_numpy(idx[valid]).to(dev)
        self.a_ref = direct_sum(*st.pos, st.mass, tidx, c.G, c.SOFTENING)
        self.norm = float((self.a_ref ** 2).sum().item())
        self.counts = np.minimum(64, st.n - self.groups * 64).astype(np.float64)
        self._force = torch.empty((self.T, 3), dtype=torch.float32, device=dev)
        self._stats = torch.zeros(4, dtype=torch.int64, device=dev)

    def errors(self, theta=0.5, theta_group=None, straddle=None, straddle_group=None):
        """(squared-error sum per group, interactions per target); straddle:
        the straddle mass (default the engine's), straddle_group per group."""
        e, tm, c = self.engine, self.engine.tree_manager, self.engine.config
        self._stats.zero_()
        traverse(*self.tpos, tm.nodes, tm.root_idx, self._force, theta, c.G, c.SOFTENING,
                 c.BOX_SIZE, mac=e.mac, theta_group=theta_group, stats=self._stats,
                 n_nodes=tm.n_nodes, compensated=bool(getattr(c, "COMPENSATED", False)),
                 straddle_mass=e.straddle_mass if straddle is None else straddle,
                 straddle_group=straddle_group)
        d = torch.zeros((self.T, 3), dtype=torch.float64, device=self.a_ref.device)
        d[self.vmask] = self._force[self.vmask].double() - self.a_ref
        err2 = (d ** 2).sum(1).reshape(len(self.groups), 64).sum(1)
        return err2, float(self._stats[0].item()) / float(self.counts.sum())

    def rel_rms(self, theta=0.5, theta_group=None, straddle=None, straddle_group=None):
        err2, inter = self.errors(theta, theta_group, straddle, straddle_group)
        return math.sqrt(float(err2.sum().item()) / self.norm), inter


def sample(engine, n_groups, thetas=DEFAULT_THETAS, seed=0, cost_kind="time", cost=None,
           mus=None):
    """Training samples from the engine's current tree (built) for the
    candidates (thetas[k], mus[k]) (mus None: the engine's straddle mass).
    cost_kind: "time" (ms of the whole-state traversal at each candidate per
    group) or "interactions" (per particle, on the sampled groups); a given
    `cost` array is used as is."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    G = (st.n + 63) // 64
    mus = [engine.straddle_mass] * len(thetas) if mus is None else list(mus)
    rng = np.random.default_rng(seed)
    groups = np.sort(rng.choice(G, size=min(n_groups, G), replace=False))
    feat, fidx = feature_blocks(engine, mus)
    th = torch.tensor(thetas, dtype=torch.float32, device=feat.device)
    codes = torch.tensor([mu_code(m, st.n) for m in mus], dtype=torch.float32,
                         device=feat.device)
    fi = torch.tensor(fidx, dtype=torch.int32, device=feat.device)
    X = K.assemble_rows(feat[:, torch.from_numpy(groups).to(feat.device)].contiguous(), th,
                        cand2=codes, fidx=fi)
    probe = GroupProbe(engine, groups)
    cols, measured = [], []
    for t, m in zip(thetas, mus):
        e2, inter = probe.errors(t, straddle=m)
        cols.append(e2)
        if cost is None:
            measured.append(inter if cost_kind == "interactions"
                            else _time_traversal(engine, t, None, reps=2, straddle=m) / G)
    cost = np.asarray(measured if cost is None else cost, np.float64)
    err2 = torch.stack(cols, 1)
    mean_norm = probe.norm / probe.counts.sum()
    cnt = torch.from_numpy(probe.counts).to(err2.device)[:, None]
    # log2 of the group's share of the squared error, floored at LABEL_FLOOR:
    # shares that far below the budget of any useful target never decide a
    # selection, and unfloored they would dominate the squared loss
    T = torch.log2(err2 / cnt / mean_norm + 1e-30).clamp_min(LABEL_FLOOR).float()
    T = T.reshape(-1).contiguous()
    return Samples(X, T, err2, cost, probe.norm, groups)


class LearnedTheta:
    """Selects engine.theta_group (and, with straddle-mass candidates,
    mu_group) each step from the model's predictions.  Candidate k is the
    opening angle thetas[k] with the straddle mass mu_factors[k] times the
    engine's (all 1 when mu_factors is None: opening angles only), so a
    network keeps its meaning at another N when the configuration scales
    its straddle mass with the particle mass.  keep (or theta_min) drops
    candidates: one group at a very small theta is a single wavefront whose
    duration bounds the whole traversal."""

    def __init__(self, model, thetas, cost, lam, fallback, theta_min=0.0, mu_factors=None,
                 keep=None, fallback_mu_factor=1.0):
        factors = [1.0] * len(thetas) if mu_factors is None else list(mu_factors)
        keep = ([k for k, t in enumerate(thetas) if t >= theta_min] if keep is None
                else list(keep))
        self.model = model
        self.thetas_list = [float(thetas[k]) for k in keep]
        self.mu_factors = [float(factors[k]) for k in keep]
        self.cost_np = np.asarray(cost, np.float32)[keep]
        self.lam = float(lam)
        # a group whose predictions are not finite gets the fixed pair: theta
        # `fallback` with fallback_mu_factor times the engine's straddle mass
        self.fallback = float(fallback)
        self.fallback_mu_factor = float(fallback_mu_factor)
        # straddle masses are chosen per group only when the candidates differ
        self.joint = len(set(self.mu_factors)) > 1
        self.mu_group = None
        self._dev = None

    def _ensure(self, device, G, n, mu_engine):
        mus_l = [mu_engine * f for f in self.mu_factors]
        if self._dev is None or self._dev[0] != G or self._dev[13] != mus_l:
            th = torch.tensor(self.thetas_list, dtype=torch.float32, device=device)
            mus = torch.tensor(mus_l, dtype=torch.float32, device=device)
            codes = torch.tensor([mu_code(m, n) for m in mus_l], dtype=torch.float32,
                                 device=device)
            cost = torch.from_numpy(self.cost_np).to(device)
            Kc = len(self.thetas_list)
            X = torch.empty((G * Kc, K.IN_PAD), dtype=torch.int16, device=device)
            Y = torch.empty(G * Kc, dtype=torch.float32, device=device)
            tg = torch.empty(G, dtype=torch.float32, device=device)
            mg = torch.empty(G, dtype=torch.float32, device=device)
            nf = torch.zeros(1, dtype=torch.int64, device=device)
            J = len(set(mus_l))
            feat = torch.empty((J, G, K.GROUP_FEATURES), dtype=torch.float32, device=device)
            order = torch.empty(G, dtype=torch.int32, device=device)
            prof = torch.empty((2, J, 6 * K.PROFILE_BINS), dtype=torch.int64, device=device)
            distinct = sorted(set(mus_l))
            fidx = torch.tensor([distinct.index(m) for m in mus_l], dtype=torch.int32,
                                device=device)
            self._dev = (G, th, cost, X, Y, tg, nf, feat, order, prof, mus, codes, mg, mus_l,
                         fidx)
            # the chosen candidates of this and the previous selection (-1:
            # none), swapped at every selection
            self.k_group = torch.full((G,), -1, dtype=torch.int32, device=device)
            self.k_prev = torch.full((G,), -1, dtype=torch.int32, device=device)
        return self._dev

    @property
    def cost_dev(self):
        """The candidates' cost per group (device float32[K])."""
        return self._dev[2]

    def theta_groups(self, engine, lam=None, a=None):
        """Per-group theta for the engine's current tree (device float[G]);
        the per-group straddle masses, when chosen, are left in mu_group.
        a: the scale factor (default engine.a; a device tensor inside a
        captured step).  Allocates nothing after the first call."""
        st, tm, c = engine.state, engine.tree_manager, engine.config
        G = (st.n + 63) // 64
        _, th, cost, X, Y, tg, nf, feat, _, prof, mus, codes, mg, mus_l, fidx = self._ensure(
            st.pos[0].device, G, st.n, float(engine.straddle_mass))
        feature_blocks(engine, mus_l, feat, prof, a)
        K.assemble_rows(feat, th, X, cand2=codes, fidx=fidx)
        self.model.predict(X, Y)
        self.k_group, self.k_prev = self.k_prev, self.k_group
        if self.joint:
            K.select_theta(Y, G, th, cost, self.lam if lam is None else lam, self.fallback, tg,
                           nf, mus=mus,
                           fallback_mu=float(engine.straddle_mass) * self.fallback_mu_factor,
                           mu_out=mg, k_out=self.k_group)
            self.mu_group = mg
        else:
            K.select_theta(Y, G, th, cost, self.lam if lam is None else lam, self.fallback, tg,
                           nf, k_out=self.k_group)
            self.mu_group = None
        return tg

    def launch_order(self, theta_group):
        """Groups by increasing theta: the traversal's launch order."""
        order = self._dev[8]
        return order_groups(theta_group, order)

    @property
    def n_fallback(self):
        return 0 if self._dev is None else int(self._dev[6].item())

    def save(self, path):
        np.savez(path, params=self.model.params.cpu().numpy(), thetas=np.array(self.thetas_list),
                 mu_factors=np.array(self.mu_factors), cost=self.cost_np, lam=self.lam,
                 fallback=self.fallback, fallback_mu_factor=self.fallback_mu_factor)

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path, allow_pickle=False)
        p = torch.from_numpy(d["params"])
        if p.dim() == 2:
            model = Ensemble([ThetaModel(q, device=device) for q in p])
        else:
            model = ThetaModel(p, device=device)
        factors = d["mu_factors"].tolist() if "mu_factors" in d.files else None
        fmu = float(d["fallback_mu_factor"]) if "fallback_mu_factor" in d.files else 1.0
        return cls(model, d["thetas"].tolist(), d["cost"], float(d["lam"]), float(d["fallback"]),
                   mu_factors=factors, fallback_mu_factor=fmu)


def _time_traversal(engine, theta, theta_group, reps=5, group_order=None, straddle=None,
                    straddle_group=None):
    """Average HIP-event time (ms) of the full traversal (all particles).
    With the engine's LONGEST_FIRST schedule every launch (the untimed first
    one included) stores its groups' pops and the next one is launched in
    the order they give, the order kernel inside the timed region, whatever
    group_order says; fixed and per-group theta are timed alike."""
    st, tm, c = engine.state, engine.tree_manager, engine.config
    comp = bool(getattr(c, "COMPENSATED", False))
    mu = engine.straddle_mass if straddle is None else straddle
    sched = engine.longest_first

    def launch():
        # (every launch here walks the same selection: plain pops)
        order = engine.launch_order(scale=False) if sched else group_order
        traverse(*st.pos, tm.nodes, tm.root_idx, st.force, theta, c.G, c.SOFTENING, c.BOX_SIZE,
                 mac=engine.mac, theta_group=theta_group, group_order=order,
                 compensated=comp, straddle_mass=mu, straddle_group=straddle_group,
                 group_pops=engine.group_pops if sched else None)

    launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _time_steps(c, state, selector, steps=3, graph=False):
    """Average time (ms) of full engine steps (sort, tree, per-group theta
    selection when a selector is given, traversal, update) starting from
    `state`; one untimed warm-up step.  graph: the steps (with the density
    splat) are replayed from captured HIP graphs and timed by the host
    clock between device synchronisations."""
    eng = _engine_at(c, *state)
    eng.theta_selector = selector
    eng.step()
    if graph:
        eng.capture_graphs(steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.step_graph()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        eng.step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _bisect(f, lo, hi, target, iters=20, log=False, increasing=True):
    """x in [lo, hi] with f(x) ~ target for a monotone f (returns the side
    that meets the target: f(x) <= target)."""
    best = None
    for _ in range(iters):
        mid = math.sqrt(lo * hi) if log else 0.5 * (lo + hi)
        v = f(mid)
        ok = v <= target
        if ok:
            best = mid
        if ok == increasing:
            lo = mid
        else:
            hi = mid
    return best


def _engine_at(c, seed, steps):
    from src.core.engine import SimulationEngine
    from src.core.state import SimulationState
    from src.utils.cosmology import generate_ics
    pos, vel = generate_ics(c, seed)
    # the states that train and judge a network run the plain method
    eng = SimulationEngine(c.replace(LEARNED_THETA=False), SimulationState(len(pos[0]), pos, vel))
    for _ in range(steps):
        eng.step()
    eng.build_tree()
    return eng


def default_states(c):
    """(training states, held-out state) as (seed, steps) pairs: later steps of
    one realisation for the Zel'dovich ICs, other realisations for the halo
    state."""
    if getattr(c, "IC_KIND", "zeldovich") == "halos":
        return [(c.SEED + 1, 0), (c.SEED + 2, 0), (c.SEED + 3, 0)], (c.SEED, 0)
    return [(c.SEED, 0), (c.SEED, 10), (c.SEED, 20)], (c.SEED, 30)


def train(c, states, n_groups, thetas, fit_steps, seed=0, log=print, n_models=N_MODELS,
          mus=None):
    """Samples from the training states for the candidates (thetas[k],
    mus[k]) (cost of each candidate timed on the first state), trains
    n_models networks (an Ensemble when more than one); returns (model,
    cost, info)."""
    Xs, Ts, costs = [], [], []
    for k, (sd, steps) in enumerate(states):
        eng = _engine_at(c, sd, steps)
        smp = sample(eng, n_groups, thetas, seed=seed + k, cost=costs[0] if costs else None,
                     mus=mus)
        Xs.append(smp.X)
        Ts.append(smp.T)
        costs.append(smp.cost)
        log(f"samples from (seed {sd}, step {steps}): {smp.X.shape[0]} rows, "
            f"cost {np.round(smp.cost, 5)}")
        del eng
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    models, hists = [], []
    t0 = time.perf_counter()
    for i in range(n_models):
        m = ThetaModel(seed=seed + i)
        hists.append(m.fit(X, T, steps=fit_steps, seed=seed + i, log=max(1, fit_steps // 10)))
        models.append(m)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    first = float(np.mean([h[0][1] for h in hists]))
    last = float(np.mean([h[-1][1] for h in hists]))
    log(f"fit {n_models} x {fit_steps} steps in {t_fit:.1f}s: loss {first:.3f} -> {last:.3f}")
    info = {"train_states": list(states), "train_rows": int(X.shape[0]), "fit_s": t_fit,
            "n_models": n_models, "loss_first": first, "loss_last": last}
    model = models[0] if n_models == 1 else Ensemble(models)
    return model, np.mean(np.stack(costs), 0), info


def measure_costs(c, state, thetas, log=print, mus=None):
    """ms per group of a whole traversal of `state` at each candidate."""
    eng = _engine_at(c, *state)
    G = (eng.state.n + 63) // 64
    mus = [eng.straddle_mass] * len(thetas) if mus is None else list(mus)
    out = []
    for t, m in zip(thetas, mus):
        out.append(_time_traversal(eng, t, None, reps=2, straddle=m) / G)
        log(f"cost at theta {t}, straddle {m:.3g}: {out[-1] * G:.1f} ms")
    return np.asarray(out)


def hindsight_choice(E, cost, norm, target, iters=30):
    """The oracle choice behind the hindsight bound: per group the candidate
    of least cost + lam * E / norm (E: squared errors, groups x candidates;
    cost: per candidate), with lam the setting of least mean cost whose
    relative RMS error sqrt(sum E / norm) meets the target.  lam is swept on
    a log grid, then bisected in log lam between the best passing grid
    setting and the next smaller one when that one misses the target, as
    theta* and the selection's lambda are: on the grid alone the mean cost
    jumps by up to the grid step where the error crosses the target.
    Returns (lam, mean cost per group) or None."""
    def at(lam):
        k = np.argmin(cost[None, :] + lam * E / norm, axis=1)
        return math.sqrt(E[np.arange(len(k)), k].sum() / norm), float(cost[k].mean())

    grid = np.logspace(-6, 16, 221)
    best, i_best = None, -1
    for i, lam in enumerate(grid):
        e, mc = at(lam)
        if e <= target and (best is None or mc < best[1]):
            best, i_best = (float(lam), mc), i
    if best is None or i_best == 0 or at(grid[i_best - 1])[0] <= target:
        return best
    lo, hi = math.log(grid[i_best - 1]), math.log(grid[i_best])
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        e, mc = at(math.exp(mid))
        if e <= target:
            hi = mid
            if mc < best[1]:
                best = (math.exp(mid), mc)
        else:
            lo = mid
    return best


def evaluate(c, model, cost, thetas, val, target=1e-3, n_val_groups=1024, seed=0,
             target_theta=None, lambdas=None, log=print, graph=False, mus=None):
    """Fixed (theta, straddle mass) against the learned per-group selection
    on the held-out state `val` at equal relative RMS error (FP64 direct sum
    on sampled groups).  The candidates are (thetas[k], mus[k]) (mus None:
    the configuration's straddle mass for all).  The fixed method is the best
    fixed pair: for every straddle mass of the candidates, the largest theta
    within the target on the candidate thetas, refined by bisection, is
    timed, and the fastest pair is the baseline.  lambda is swept on a log
    grid, the cheapest settings meeting the target by the cost model are
    timed and the fastest is kept.  Reports traversal and full-step times,
    interactions and the hindsight bound (per group the candidate of least
    cost + lambda * measured error).  theta*, the straddle mass and lambda
    are tuned on one sample of groups; both errors are then measured again
    on a disjoint sample of as many groups (err_*_confirm).  Returns
    (result, LearnedTheta), or (result, None) when the target cannot be
    met."""
    eng = _engine_at(c, *val)
    G = (eng.state.n + 63) // 64
    mu0 = eng.straddle_mass
    mus = [mu0] * len(thetas) if mus is None else list(mus)
    rng = np.random.default_rng(seed + 12345)
    # the tuning sample (theta*, lambda) and a disjoint confirmation sample on
    # which both errors are measured again
    drawn = rng.permutation(G)
    nv = min(n_val_groups, G // 2)
    vgroups = np.sort(drawn[:nv])
    cgroups = np.sort(drawn[nv:2 * nv])
    probe = GroupProbe(eng, vgroups)
    err2_k = []
    for t, m in zip(thetas, mus):
        e2, _ = probe.errors(t, straddle=m)
        err2_k.append(e2.cpu().numpy())
    th_fixed = sorted(set(float(t) for t in thetas))
    if target_theta is not None:
        target = probe.rel_rms(target_theta)[0]
        log(f"target = error of fixed theta {target_theta}: {target:.4e}")

    def fixed_at(mu):
        # the largest (cheapest) theta whose error is within the target, found
        # on the candidate curve and refined by bisection up to the next
        # candidate (the curve need not be monotone), and its traversal time
        curve = []
        for t in th_fixed:
            e, inter = probe.rel_rms(t, straddle=mu)
            curve.append((t, e, inter))
        log(f"fixed theta curve at straddle {mu:.3g} (theta, rel_rms, interactions): " +
            ", ".join(f"({t:.2f}, {e:.2e}, {i:.0f})" for t, e, i in curve))
        ok_c = [k for k, (_, e, _) in enumerate(curve) if e <= target]
        if not ok_c:
            return curve, None
        k_ok = max(ok_c)
        th = float(th_fixed[k_ok])
        if k_ok + 1 < len(th_fixed):
            t_b = _bisect(lambda t: probe.rel_rms(t, straddle=mu)[0], th_fixed[k_ok],
                          th_fixed[k_ok + 1], target, iters=14)
            th = t_b if t_b is not None else th
        e, inter = probe.rel_rms(th, straddle=mu)
        t_ms = _time_traversal(eng, th, None, straddle=mu)
        log(f"  fixed pair meeting {target:.1e}: theta {th:.4f}, straddle {mu:.3g} "
            f"(error {e:.4e}): traversal {t_ms:.1f} ms")
        return curve, (th, e, inter, t_ms)

    fixed_by_mu = {}
    fixed_curve = None
    for mu in sorted(set(float(m) for m in mus)):
        curve, best_mu = fixed_at(mu)
        if mu == float(mu0):
            fixed_curve = curve
        if best_mu is not None:
            fixed_by_mu[mu] = best_mu
    if not fixed_by_mu:
        return {"reachable": False, "val_state": val, "target": target,
                "fixed_curve": fixed_curve}, None
    mu_star = min(fixed_by_mu, key=lambda m: fixed_by_mu[m][3])
    theta_star, e_fixed, i_fixed, t_fixed = fixed_by_mu[mu_star]
    log(f"best fixed pair: theta {theta_star:.4f}, straddle {mu_star:.3g} "
        f"(error {e_fixed:.4e}), traversal {t_fixed:.1f} ms")
    E = np.stack(err2_k, 1)
    cost_arr = np.asarray(cost, np.float64)
    bound = hindsight_choice(E, cost_arr, probe.norm, target)
    # Candidates whose average single-wavefront time would exceed half the
    # fixed-theta traversal are dropped: the traversal cannot end before its
    # slowest wavefront.  All groups at candidate k take cost_k * G ms with
    # `concurrent` wavefronts resident at a time.
    props = torch.cuda.get_device_properties(eng.state.pos[0].device)
    concurrent = min(props.multi_processor_count * 4 * WAVES_PER_SIMD, G)
    wave_ms = cost_arr * concurrent
    keep = [k for k in range(len(thetas)) if wave_ms[k] <= WAVE_BUDGET * t_fixed]
    if not keep:
        keep = [int(np.argmin(wave_ms))]
    theta_min = float(min(thetas[k] for k in keep))
    # the same bound over the candidates the selection keeps
    bound_kept = hindsight_choice(E[:, keep], cost_arr[keep], probe.norm, target)
    lt = LearnedTheta(model, thetas, cost, 1.0, theta_star, keep=keep,
                      mu_factors=[m / mu0 for m in mus] if mu0 > 0 else None,
                      fallback_mu_factor=(mu_star / mu0) if mu0 > 0 else 1.0)
    vg = torch.from_numpy(vgroups).to(eng.state.pos[0].device)
    cost_sel = np.asarray(lt.cost_np, np.float64)
    pairs_sel = list(zip(lt.thetas_list, [mu0 * f for f in lt.mu_factors]))

    def sub(t):
        return None if t is None else t.index_select(0, vg).contiguous()

    def model_cost(tg, mg):
        tg = tg.cpu().numpy()
        mg = np.full_like(tg, mu0) if mg is None else mg.cpu().numpy()
        idx = {(np.float32(a), np.float32(b)): k for k, (a, b) in enumerate(pairs_sel)}
        return float(np.mean([cost_sel[idx.get((np.float32(a), np.float32(b)), 0)]
                              for a, b in zip(tg, mg)]))

    lambdas = np.logspace(-2, 14, 129) if lambdas is None else lambdas
    sweep = []
    for j, lam in enumerate(lambdas):
        tg = lt.theta_groups(eng, float(lam))
        mg = lt.mu_group
        e, inter = probe.rel_rms(0.5, sub(tg), straddle=mu0, straddle_group=sub(mg))
        sweep.append((float(lam), e, inter, model_cost(tg, mg)))
        if j % 32 == 31:
            log(f"lambda sweep {j + 1}/{len(lambdas)}")
    ok = sorted((sw for sw in sweep if sw[1] <= target), key=lambda sw: sw[3])
    if not ok:
        return {"reachable": False, "val_state": val, "target": target,
                "theta_fixed": theta_star, "straddle_fixed": mu_star, "theta_min": theta_min,
                "sweep": sweep}, None
    # The grid steps lambda by 1.33x, so its cheapest setting can sit well
    # below the target; as theta* is bisected to the target, lambda is
    # bisected (in log lambda) between that setting and the next smaller
    # grid lambda that misses the target, and the result joins the timed
    # candidates.
    lam_ok = ok[0][0]
    below = [sw for sw in sweep if sw[0] < lam_ok and sw[1] > target]
    if below:
        lo_l, hi_l = math.log(max(sw[0] for sw in below)), math.log(lam_ok)
        for _ in range(10):
            mid = 0.5 * (lo_l + hi_l)
            tg = lt.theta_groups(eng, math.exp(mid))
            mg = lt.mu_group
            e, inter = probe.rel_rms(0.5, sub(tg), straddle=mu0, straddle_group=sub(mg))
            if e <= target:
                hi_l = mid
                refined = (math.exp(mid), e, inter, model_cost(tg, mg))
            else:
                lo_l = mid
        if hi_l < math.log(lam_ok):
            ok = [refined] + ok
            log(f"lambda refined to {refined[0]:.4g}: error {refined[1]:.4e}")
    best = None
    for lam, _, _, _ in ok[:5]:  # the cheapest settings by the cost model, timed
        tg = lt.theta_groups(eng, lam).clone()
        mg = None if lt.mu_group is None else lt.mu_group.clone()
        t = _time_traversal(eng, 0.5, tg, group_order=lt.launch_order(tg).clone(),
                            straddle_group=mg)
        log(f"lambda {lam:.3g}: learned traversal {t:.1f} ms")
        if best is None or t < best[1]:
            best = (lam, t)
    lt.lam = best[0]
    tg = lt.theta_groups(eng).clone()
    mg = None if lt.mu_group is None else lt.mu_group.clone()
    e_learned, i_learned = probe.rel_rms(0.5, sub(tg), straddle=mu0, straddle_group=sub(mg))
    cprobe = GroupProbe(eng, cgroups)
    cg = torch.from_numpy(cgroups).to(eng.state.pos[0].device)
    e_fixed_c, _ = cprobe.rel_rms(theta_star, straddle=mu_star)
    e_learned_c, _ = cprobe.rel_rms(
        0.5, tg.index_select(0, cg).contiguous(), straddle=mu0,
        straddle_group=None if mg is None else mg.index_select(0, cg).contiguous())
    log(f"confirmation sample ({len(cgroups)} other groups): fixed {e_fixed_c:.4e}, "
        f"learned {e_learned_c:.4e}")
    del cprobe
    t_learned_unordered = _time_traversal(eng, 0.5, tg, straddle_group=mg)
    t_learned = _time_traversal(eng, 0.5, tg, group_order=lt.launch_order(tg).clone(),
                                straddle_group=mg)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        lt.theta_groups(eng)
    e1.record()
    torch.cuda.synchronize()
    t_select = e0.elapsed_time(e1) / 5
    step_fixed = _time_steps(c.replace(THETA=theta_star, STRADDLE_MASS=mu_star), val, None,
                             graph=graph)
    log(f"fixed-theta step: {step_fixed:.1f} ms")
    step_learned = _time_steps(c, val, lt, graph=graph)
    log(f"learned-theta step: {step_learned:.1f} ms")
    tg_np = tg.cpu().numpy()
    hist_t = {float(t): int(np.sum(np.isclose(tg_np, t))) for t in sorted(set(lt.thetas_list))}
    hist_mu = None
    if mg is not None:
        mg_np = mg.cpu().numpy()
        hist_mu = {float(mu0 * f): int(np.sum(np.isclose(mg_np, mu0 * f)))
                   for f in sorted(set(lt.mu_factors))}
    res = {"reachable": True, "preset": c.NAME, "mac": c.MAC, "target": target,
           "bucket": int(getattr(c, "BUCKET", 0)),
           "compensated": bool(getattr(c, "COMPENSATED", False)),
           "straddle_mass": float(getattr(c, "STRADDLE_MASS", 0.0)),
           "joint_straddle": bool(lt.joint), "candidates": len(lt.thetas_list),
           "n": eng.state.n, "val_state": val, "val_groups": int(len(vgroups)),
           "hip_graph_steps": bool(graph),
           "theta_fixed": theta_star, "straddle_fixed": mu_star,
           "fixed_by_straddle": {f"{m:.4g}": {"theta": v[0], "err": v[1], "traversal_ms": v[3]}
                                 for m, v in sorted(fixed_by_mu.items())},
           "err_fixed": e_fixed,
           "inter_fixed": i_fixed, "lambda": lt.lam, "err_learned": e_learned,
           "inter_learned": i_learned, "confirm_groups": int(len(cgroups)),
           "err_fixed_confirm": e_fixed_c, "err_learned_confirm": e_learned_c,
           "traversal_ms_fixed": t_fixed,
           "traversal_ms_learned": t_learned,
           "traversal_ms_learned_unordered": t_learned_unordered, "select_ms": t_select,
           "speedup_traversal": t_fixed / t_learned,
           "speedup_with_select": t_fixed / (t_learned + t_select),
           "step_ms_fixed": step_fixed, "step_ms_learned": step_learned,
           "speedup_step": step_fixed / step_learned,
           "hindsight_bound_speedup": ((t_fixed / G) / bound[1] if bound else None),
           "hindsight_bound_kept": ((t_fixed / G) / bound_kept[1] if bound_kept else None),
           "interaction_ratio": i_fixed / i_learned,
           "theta_group_mean": float(tg.mean().item()), "theta_min": theta_min,
           "theta_group_histogram": hist_t, "straddle_group_histogram": hist_mu,
           "fallbacks": lt.n_fallback}
    return res, lt


def experiment(preset="repo_128", target=1e-3, mac="aabb", states=None, n_groups=2048,
               n_val_groups=1024, fit_steps=FIT_STEPS, seed=0, thetas=None,
               target_theta=None, lambdas=None, log=print, overrides=None, val_states=None,
               pairs=True, mu_factors=None):
    """Trains the network on the GPU from samples the method generates on the
    training states, then evaluates it (see evaluate) on each held-out
    state.  The candidates are opening angles, and with pairs (and a
    straddle mass) opening angles x straddle masses (see candidates()).
    Returns [(result, LearnedTheta or None)] per held-out state and the
    training info."""
    from src.utils import config as cfg
    c = cfg.preset(preset, MAC=mac, **(overrides or {}))
    th, mus = candidates(c, thetas, pairs, mu_factors)
    train_states, val = default_states(c) if states is None else states
    val_states = [val] if val_states is None else list(val_states)
    model, cost, info = train(c, train_states, n_groups, th, fit_steps, seed, log, mus=mus)
    out = []
    for v in val_states:
        res, lt = evaluate(c, model, cost, th, v, target, n_val_groups, seed, target_theta,
                           lambdas, log, mus=mus)
        res.update(info)
        out.append((res, lt))
    return out


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--preset", default="repo_128")
    ap.add_argument("--mac", default="aabb")
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--fit-steps", type=int, default=FIT_STEPS)
    ap.add_argument("--mu-factors", default="",
                    help="comma-separated straddle-mass factors of the candidates (default "
                         + ",".join(str(f) for f in MU_FACTORS) + ")")
    ap.add_argument("--groups", type=int, default=10240)
    ap.add_argument("--target-theta", type=float, default=None)
    ap.add_argument("--val-groups", type=int, default=1024)
    ap.add_argument("--val-seeds", default="",
                    help="comma-separated seeds of extra held-out states (step 0)")
    ap.add_argument("--train-seeds", default="",
                    help="comma-separated seeds of the training states (step 0) in place of "
                         "the preset's")
    ap.add_argument("--bucket", type=int, default=None)
    ap.add_argument("--compensated", type=int, default=None)
    ap.add_argument("--straddle", type=float, default=None)
    ap.add_argument("--eval-preset", default="",
                    help="evaluate on this preset (its own theta costs, same network); "
                         "e.g. a larger N than the training preset")
    ap.add_argument("--eval-straddle", type=float, default=None,
                    help="straddle mass of --eval-preset (default: that of the training run)")
    ap.add_argument("--eval-thetas", default="",
                    help="comma-separated candidate thetas for --eval-preset")
    ap.add_argument("--graph", action="store_true",
                    help="time full steps replayed from HIP graphs (with the density splat)")
    ap.add_argument("--theta-only", action="store_true",
                    help="candidates are opening angles only (no per-group straddle mass)")
    ap.add_argument("--save", default="")
    args = ap.parse_args()
    over = {}
    if args.bucket is not None:
        over["BUCKET"] = args.bucket
    if args.compensated is not None:
        over["COMPENSATED"] = bool(args.compensated)
    if args.straddle is not None:
        over["STRADDLE_MASS"] = args.straddle
    from src.utils import config as cfg
    c = cfg.preset(args.preset, MAC=args.mac, **over)
    train_states, val = default_states(c)
    if args.train_seeds:
        train_states = [(int(x), 0) for x in args.train_seeds.split(",") if x]
    vals = [val] + [(int(x), 0) for x in args.val_seeds.split(",") if x]
    mu_f = tuple(float(x) for x in args.mu_factors.split(",") if x) or None
    if args.eval_preset:
        log = lambda m: print(m, flush=True)  # noqa: E731
        pairs = not args.theta_only
        th_t, mus_t = candidates(c, None, pairs, mu_f)
        model, _, info = train(c, train_states, args.groups, th_t, args.fit_steps, log=log,
                               mus=mus_t)
        over_e = dict(over)
        if args.eval_straddle is not None:
            over_e["STRADDLE_MASS"] = args.eval_straddle
        ce = cfg.preset(args.eval_preset, MAC=args.mac, **over_e)
        th = (tuple(float(x) for x in args.eval_thetas.split(",")) if args.eval_thetas
              else None)
        th, mus = candidates(ce, th, pairs, mu_f)
        cost = measure_costs(ce, vals[0], th, log, mus=mus)
        for v in vals:
            res, _ = evaluate(ce, model, cost, th, v, args.target, args.val_groups, log=log,
                              graph=args.graph, mus=mus)
            res.update(info)
            res["train_preset"] = args.preset
            print(json.dumps(res, default=float), flush=True)
        return
    out = experiment(args.preset, args.target, args.mac, fit_steps=args.fit_steps,
                     n_groups=args.groups, n_val_groups=args.val_groups,
                     target_theta=args.target_theta, log=lambda m: print(m, flush=True),
                     overrides=over, val_states=vals, states=(train_states, val),
                     pairs=not args.theta_only, mu_factors=mu_f)
    for res, lt in out:
        print(json.dumps(res, default=float), flush=True)
    if args.save and out[0][1] is not None:
        out[0][1].save(args.save)


if __name__ == "__main__":
    main()
