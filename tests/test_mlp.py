"""Learned-theta MLP: the CPU mirror is pinned by finite differences; the
MFMA forward/backward kernels, Adam, the group features and the theta
selection are checked against the mirror on the GPU."""
import numpy as np
import pytest

from conftest import gpu_available

from oracle import mlp_ref as M


def _rand_rows(rng, B):
    X = np.zeros((B, M.IN_PAD))
    X[:, :M.IN] = rng.standard_normal((B, M.IN))
    return M.bf16(X)


def test_mirror_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1).astype(np.float64)
    p[M.OFF_B1:M.OFF_W2] = rng.standard_normal(M.HID) * 0.1
    p[M.OFF_B2:M.OFF_W3] = rng.standard_normal(M.HID) * 0.1
    X = _rand_rows(rng, 37)
    T = rng.standard_normal(37)
    loss, g, _ = M.loss_and_grad(X, T, p, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_W1 + M.IN * M.HID), 40),
                          rng.choice(np.arange(M.OFF_B1, M.N_PARAMS), 80), [M.OFF_B3]])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = M.loss_and_grad(X, T, q, rounding=False)[0]
        q[i] -= 2 * h
        lm = M.loss_and_grad(X, T, q, rounding=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded input rows of W1 never receive gradient
    assert np.all(g[M.OFF_W1 + M.IN * M.HID:M.OFF_B1] == 0)


def test_bf16_rounding_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert np.isfinite(r[4])


gpu = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a).astype(dtype)).to("cuda")


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
@pytest.mark.parametrize("B", [1, 16, 1000, 4096])
def test_forward_and_grad_match_mirror(B):
    import torch
    from src.kernels import mlp_kernels as K
    rng = np.random.default_rng(B)
    p = M.init_params(2)
    p[M.OFF_B1:M.OFF_W2] = rng.standard_normal(M.HID) * 0.1
    p[M.OFF_B2:M.OFF_W3] = rng.standard_normal(M.HID) * 0.1
    p[M.OFF_B3] = 0.3
    X = _rand_rows(rng, B)
    T = rng.standard_normal(B)
    params = _dev(torch, p, np.float32)
    wbf = torch.empty(K.BF16_WEIGHTS, dtype=torch.int16, device="cuda")
    K.to_bf16(params, wbf)
    Xg = _dev(torch, M.bf16_bits(X).view(np.int16), np.int16)
    y = K.forward(Xg, wbf, params).cpu().numpy()
    y_ref, _ = M.forward(X, p)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    g = K.grad(Xg, _dev(torch, T, np.float32), wbf, params).cpu().numpy()
    _, g_ref, sse = M.loss_and_grad(X, T, p)
    assert abs(g[-1] - sse) <= 1e-4 * sse
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                   (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        a, b = g[lo:hi].astype(np.float64), g_ref[lo:hi]
        assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b) + 1e-7, (lo, hi)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_adam_step_and_training_reduces_loss():
    import torch
    from src.kernels import mlp_kernels as K
    rng = np.random.default_rng(5)
    B = 8192
    X = _rand_rows(rng, B)
    T = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]
    p = M.init_params(3)
    params = _dev(torch, p, np.float32)
    wbf = torch.empty(K.BF16_WEIGHTS, dtype=torch.int16, device="cuda")
    K.to_bf16(params, wbf)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    Xg = _dev(torch, M.bf16_bits(X).view(np.int16), np.int16)
    Tg = _dev(torch, T, np.float32)
    g = K.grad(Xg, Tg, wbf, params)
    # one Adam step against numpy
    gn = g[:-1].cpu().numpy().astype(np.float64)
    K.adam(params, m, v, g, 1e-3, 1, wbf)
    mh = 0.1 * gn / (1 - 0.9)
    vh = 0.001 * gn * gn / (1 - 0.999)
    expect = p - 1e-3 * mh / (np.sqrt(vh) + 1e-8)
    expect[M.OFF_W1 + M.IN * M.HID:M.OFF_B1] = 0
    assert np.allclose(params.cpu().numpy(), expect, rtol=1e-5, atol=1e-6)
    loss0 = float(g[-1].item()) / B
    for step in range(2, 400):
        g = K.grad(Xg, Tg, wbf, params)
        K.adam(params, m, v, g, 3e-3, step, wbf)
    g = K.grad(Xg, Tg, wbf, params)
    loss1 = float(g[-1].item()) / B
    assert loss1 < 0.1 * loss0, (loss0, loss1)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_group_features_match_mirror():
    import torch
    from conftest import zeldovich
    from src.core.octree import OctreeManager
    from src.core.state import SimulationState
    from src.kernels import mlp_kernels as K
    from src.utils import config as cfg
    pos, vel = zeldovich(16)
    n = len(pos[0])
    st = SimulationState(n, pos, vel)
    st.sort_buffer_by_morton(cfg.preset("small_32"))
    tree = OctreeManager(n)
    tree.build(st.pos, st.mass, st.codes)
    h = lambda t: t.cpu().numpy()
    for mu in (0.0, 3e-3, 2e-2):
        prof = K.straddle_profile(tree, mu) if mu > 0 else None
        f = K.group_features(st.pos, tree, 0.02, 16, profile=prof).cpu().numpy()
        ref_prof = (M.straddle_profile(h(tree.parent), h(tree.mass), h(tree.aabb_min),
                                       h(tree.aabb_max), mu) if mu > 0 else None)
        if mu > 0:
            assert np.array_equal(h(prof).view(np.uint64).reshape(2, 3, -1), ref_prof)
        ref = M.group_features([h(t) for t in st.pos], h(tree.child), h(tree.parent),
                               h(tree.mass), h(tree.com), h(tree.aabb_min), h(tree.aabb_max),
                               0.02, 16, profile=ref_prof)
        assert np.allclose(f, ref, rtol=1e-5, atol=1e-5)
        assert (mu == 0) == (not np.any(f[:, 7:]))


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_multi_profile_features_equal_single():
    """The one-pass profiles of several straddle masses and the
    one-wavefront-per-group features of all of them equal, bit for bit, the
    per-mass profiles and features (a group count that is not a multiple of
    64 particles; the scale factor given on the host and on the device)."""
    import torch
    from conftest import zeldovich
    from src.core.octree import OctreeManager
    from src.core.state import SimulationState
    from src.kernels import mlp_kernels as K
    from src.utils import config as cfg
    pos, vel = zeldovich(16)
    n = 4000
    pos = [p[:n].copy() for p in pos]
    vel = [v[:n].copy() for v in vel]
    st = SimulationState(n, pos, vel)
    st.sort_buffer_by_morton(cfg.preset("small_32"))
    tree = OctreeManager(n)
    tree.build(st.pos, st.mass, st.codes)
    mus = [1e-3, 3e-3, 2e-2, 5e-2, 0.2]
    profs = K.straddle_profiles(tree, mus).reshape(len(mus), -1)
    a_dev = torch.tensor([0.02], dtype=torch.float32, device="cuda")
    for a in (0.02, a_dev):
        f = K.group_features_multi(st.pos, tree, a, 16, profs, len(mus))
        for j, mu in enumerate(mus):
            one = K.straddle_profile(tree, mu)
            assert torch.equal(profs[j], one)
            fj = K.group_features(st.pos, tree, 0.02, 16, profile=one)
            assert torch.equal(f[j], fj)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_select_theta_and_nan_fallback():
    import torch
    from src.kernels import mlp_kernels as K
    G, Kc = 1000, 4
    thetas = torch.tensor([0.3, 0.5, 0.7, 0.9], dtype=torch.float32, device="cuda")
    cost = torch.tensor([4.0, 2.0, 1.0, 0.5], dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(1)
    pred = rng.uniform(-8, 2, (G, Kc)).astype(np.float32)
    out = torch.empty(G, dtype=torch.float32, device="cuda")
    nf = torch.zeros(1, dtype=torch.int64, device="cuda")
    K.select_theta(_dev(torch, pred.reshape(-1), np.float32), G, thetas, cost, 3.0, 0.5, out, nf)
    score = cost.cpu().numpy()[None, :] + 3.0 * np.exp2(pred)
    expect = thetas.cpu().numpy()[np.argmin(score, 1)]
    assert np.array_equal(out.cpu().numpy(), expect)
    assert int(nf.item()) == 0
    pred[7, 2] = np.nan
    K.select_theta(_dev(torch, pred.reshape(-1), np.float32), G, thetas, cost, 3.0, 0.5, out, nf)
    assert out[7].item() == pytest.approx(0.5) and int(nf.item()) == 1


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_select_pairs_and_rows():
    """Joint candidates (theta, straddle mass): the selection returns the
    theta and the straddle mass of the same best candidate (and both
    fallbacks for a non-finite prediction); the MLP rows carry the
    candidate's straddle code in input 10."""
    import torch
    from src.kernels import mlp_kernels as K
    G = 500
    th = np.array([0.3, 0.5, 0.3, 0.5, 0.7], np.float32)
    mus = np.array([1e-5, 1e-5, 4e-5, 4e-5, 2e-5], np.float32)
    cost = np.array([4.0, 2.0, 3.0, 1.5, 0.5], np.float32)
    rng = np.random.default_rng(2)
    pred = rng.uniform(-8, 2, (G, len(th))).astype(np.float32)
    pred[11, 3] = np.inf
    dv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    out = torch.empty(G, dtype=torch.float32, device="cuda")
    mout = torch.empty(G, dtype=torch.float32, device="cuda")
    nf = torch.zeros(1, dtype=torch.int64, device="cuda")
    kout = torch.empty(G, dtype=torch.int32, device="cuda")
    K.select_theta(dv(pred.reshape(-1)), G, dv(th), dv(cost), 3.0, 0.45, out, nf, mus=dv(mus),
                   fallback_mu=3e-5, mu_out=mout, k_out=kout)
    k = np.argmin(cost[None, :] + 3.0 * np.exp2(pred), 1)
    o, mo = out.cpu().numpy(), mout.cpu().numpy()
    ok = np.ones(G, bool)
    ok[11] = False
    assert np.array_equal(o[ok], th[k][ok]) and np.array_equal(mo[ok], mus[k][ok])
    assert o[11] == np.float32(0.45) and mo[11] == np.float32(3e-5) and int(nf.item()) == 1
    # the chosen candidate's index, -1 for the fallback
    ko = kout.cpu().numpy()
    assert np.array_equal(ko[ok], k[ok]) and ko[11] == -1
    feat = dv(rng.standard_normal((7, K.GROUP_FEATURES)).astype(np.float32))
    codes = dv(np.log2(1.0 + mus * 2 ** 21).astype(np.float32))
    X = K.assemble_rows(feat, dv(th), cand2=codes).cpu().numpy().view(np.uint16)
    bf = lambda a: (np.asarray(a, np.float32).view(np.uint32) + 0x7FFF +  # noqa: E731
                    ((np.asarray(a, np.float32).view(np.uint32) >> 16) & 1)) >> 16
    assert np.array_equal(X[:, 10], np.tile(bf(codes.cpu().numpy()), 7).astype(np.uint16))
    assert np.array_equal(X[:, 0], np.tile(bf(th), 7).astype(np.uint16))
    assert not X[:, 11:].any()
