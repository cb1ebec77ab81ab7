"""Golden fixtures (tests/golden/za8.npz, za16.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them, and
the HIP path must match them bit for bit on the stored inputs."""
import hashlib
import os

import numpy as np
import pytest

from conftest import gpu_available, rel_rms

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_stages(pos, vel, n_grid):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.stages(list(pos), list(vel), n_grid)


def test_oracle_reproduces_za8():
    g = load("za8.npz")
    out = oracle_stages(g["pos"], g["vel"], 8)
    for k, v in out.items():
        assert np.array_equal(v, g[k]), k


def test_oracle_reproduces_za16_digests():
    g = load("za16.npz")
    out = oracle_stages(g["pos"], g["vel"], 16)
    for k, d in zip(g["digest_names"], g["digests"]):
        assert digest(out[str(k)]) == str(d), k


def gpu_stages(pos, vel, n_grid, force_in, mvp, res, theta):
    """The same stages on the HIP path; the update takes the stored FP32
    oracle force so that it is compared bit for bit."""
    import torch

    from src.kernels import _hip
    from src.kernels.morton import compute_codes
    from src.kernels.physics_kernels import integrate, traverse
    from src.kernels.render_kernels import render_density
    from src.kernels.tree_kernels import (compute_multipoles, normalize_com, pack_nodes,
                                          build_tree, reset_tree_buffers)

    def d(a):
        a = np.ascontiguousarray(a)
        if a.dtype == np.uint64:
            a = a.view(np.int64)
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        return torch.from_numpy(a).to("cuda")

    def h(t, dt=None):
        a = t.cpu().numpy()
        return a.view(dt) if dt else a

    n = pos.shape[1]
    M = 2 * n - 1
    soft = 0.02 / n_grid
    out = {}
    P = [d(p) for p in pos]
    ctx = _hip.Context(n)
    for ms, tag in ((0, "codes_full21"), (1, "codes_ref10")):
        codes = torch.zeros(n, dtype=torch.int64, device="cuda")
        compute_codes(*P, codes, None, n, mask_set=ms)
        out[tag] = h(codes, np.uint64)
        sc = torch.empty_like(codes)
        perm = torch.empty(n, dtype=torch.int32, device="cuda")
        _hip.check(_hip.load().cosmo_sort_by_code(ctx.handle, _hip.ptr(codes, _hip.U64, "c"),
                                                  _hip.ptr(sc, _hip.U64, "s"),
                                                  _hip.ptr(perm, _hip.U32, "p"), n,
                                                  63 if ms == 0 else 30,
                                                  _hip.stream_handle()), "sort")
        out["perm" if ms == 0 else "perm_ref10"] = h(perm, np.uint32)
        if ms == 0:
            sorted_codes = sc
    perm = out["perm"]
    ps = [d(p[perm]) for p in pos]
    ch = torch.empty((M, 2), dtype=torch.int32, device="cuda")
    par = torch.empty(M, dtype=torch.int32, device="cuda")
    cnt = torch.empty(M, dtype=torch.int32, device="cuda")
    root = torch.empty(1, dtype=torch.int32, device="cuda")
    reset_tree_buffers(ch, par, cnt, root)
    build_tree(sorted_codes, ch, par)
    out["children"], out["parents"] = h(ch), h(par)
    mass = d(np.full(n, 1.0 / n, np.float32))
    nm = torch.zeros(M, dtype=torch.float64, device="cuda")
    nc = torch.zeros((M, 3), dtype=torch.float64, device="cuda")
    lo = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    hi = torch.zeros((M, 3), dtype=torch.float32, device="cuda")
    compute_multipoles(*ps, mass, ch, par, nm, nc, lo, hi, cnt)
    out["n_mass"], out["n_com_sum"], out["n_min"], out["n_max"] = h(nm), h(nc), h(lo), h(hi)
    normalize_com(nm, nc, M)
    out["n_com"] = h(nc)
    root.fill_(n)
    for mac, tag in ((0, "ref"), (1, "aabb")):
        nodes = pack_nodes(ch, nm, nc, lo, hi, 1.0, mac)
        f = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        stats = torch.zeros(4, dtype=torch.int64, device="cuda")
        traverse(*ps, nodes, root, f, theta, 1.0, soft, 1.0, mac=mac, stats=stats)
        out[f"force32_{tag}"] = h(f)
        out[f"stats_{tag}"] = h(stats)
    vs = [d(v[perm]) for v in vel]
    a0, H, dt = (float(x) for x in out_args(n_grid))
    integrate(*ps, *vs, d(force_in), a0, H, dt)
    out["update"] = np.stack([h(t) for t in ps + vs])
    grid = torch.zeros((res, res), dtype=torch.int64, device="cuda")
    render_density(*ps, grid, mvp, res)
    out["grid"] = h(grid, np.uint64)
    return out


def out_args(n_grid):
    from src.utils import config as cfg
    from src.utils.cosmology import get_hubble_param
    c = cfg.preset("small_32", N_GRID=n_grid)
    H = get_hubble_param(c.A_START, c)
    return c.A_START, H, c.DA / (c.A_START * H) * 0.8


BIT_EXACT = ["codes_full21", "codes_ref10", "perm", "perm_ref10", "children", "parents",
             "n_mass", "n_com_sum", "n_min", "n_max", "n_com", "update", "grid"]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_hip_matches_za8():
    g = load("za8.npz")
    out = gpu_stages(g["pos"], g["vel"], 8, g["force32_ref"], g["mvp"], int(g["res"]),
                     float(g["theta"]))
    for k in BIT_EXACT:
        assert np.array_equal(out[k], g[k]), k
    for tag in ("ref", "aabb"):
        assert int(out[f"stats_{tag}"][0]) == int(g[f"n_inter_{tag}"].sum())
        assert int(out[f"stats_{tag}"][1]) == int(g[f"n_visit_{tag}"].sum())
        assert rel_rms(out[f"force32_{tag}"], g[f"force32_{tag}"]) <= 2e-6
    assert rel_rms(out["force32_ref"], g["force64_ref"]) <= 1e-5


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_hip_matches_za16_digests():
    g = load("za16.npz")
    # the update and grid digests were made with the oracle's FP32 force
    ref = oracle_stages(g["pos"], g["vel"], 16)
    out = gpu_stages(g["pos"], g["vel"], 16, ref["force32_ref"], ref["mvp"], int(g["res"]),
                     float(g["theta"]))
    names = [str(k) for k in g["digest_names"]]
    for k, dg in zip(names, g["digests"]):
        if k in BIT_EXACT:
            assert digest(out[k]) == str(dg), k
