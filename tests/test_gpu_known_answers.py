"""Hand-derived known answers on the GPU path: the reference acceptance test
(x-extent only, box/2 clamp) on hand-built trees through the traversal, the
split walk and the compute_forces drop-in, and the density splat's row and
column orientation (derivations in kat_cases.py)."""
import numpy as np
import pytest

import kat_cases as K
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("name", ["x_extent", "clamp", "no_clamp"])
def test_reference_mac_known_answers_gpu(orc, name):
    import torch
    from src.kernels.physics_kernels import compute_forces, traverse, traverse_split
    from src.kernels.tree_kernels import pack_nodes
    case = K.FORCE_CASES[name]
    pos, mass, ch, par = K.tree_arrays(case["points"], case["masses"])
    nm, nc, lo, hi = orc.compute_multipoles(*pos, mass, ch, par)
    com = orc.normalize_com(nm, nc)
    expect = K.expected_force(case)
    P = [_dev(p) for p in pos]
    nodes = pack_nodes(_dev(ch), _dev(nm), _dev(com), _dev(lo), _dev(hi), 1.0, 0)
    root = torch.tensor([3], dtype=torch.int32, device="cuda")
    f = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    st = torch.zeros(4, dtype=torch.int64, device="cuda")
    traverse(*P, nodes, root, f, case["theta"], K.G, K.SOFT, 1.0, off=0, count=1, stats=st)
    s = st.cpu().numpy()
    assert int(s[0]) == case["interactions"] and int(s[1]) == case["visits"]
    got = f.cpu().numpy()[0].astype(np.float64)
    assert np.linalg.norm(got - expect) <= 2e-6 * np.linalg.norm(expect)
    assert not f.cpu().numpy()[1:].any()
    # the split walk and the reference-signature drop-in (FP32 node arrays,
    # as octree.py:36-39 allocates them) give the same answer
    f2 = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    st2 = torch.zeros(2, dtype=torch.int64, device="cuda")
    traverse_split(*P, nodes, root, f2, case["theta"], K.G, K.SOFT, 1.0, off=0, count=1,
                   stats=st2)
    assert st2.cpu().tolist() == [case["interactions"], case["visits"]]
    assert np.linalg.norm(f2.cpu().numpy()[0] - expect) <= 2e-6 * np.linalg.norm(expect)
    f3 = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    compute_forces[1, 1](*P, _dev(ch), _dev(nm.astype(np.float32)),
                         _dev(com.astype(np.float32)), _dev(lo), _dev(hi), f3, case["theta"],
                         K.G, K.SOFT, 1.0, root, None, None, 0)
    assert np.linalg.norm(f3.cpu().numpy()[0] - expect) <= 2e-6 * np.linalg.norm(expect)


def test_render_orientation_known_answers_gpu(orc):
    import torch
    from src.kernels.render_kernels import render_density
    from src.utils import config as cfg
    from src.utils.visualizer import get_mvp_matrix
    c = cfg.preset("repo_128")
    for res in (256, 1024):
        mvp = get_mvp_matrix(0, (0.5, 0.5, 0.5), 1.0, c)
        for name, p in K.SPLAT_POINTS.items():
            row, col, fr, fc = K.pinhole_pixel(p, res)
            grid = torch.zeros((res, res), dtype=torch.int64, device="cuda")
            render_density[1, 1](*[_dev(np.array([v], np.float32)) for v in p], grid, mvp, res)
            g = grid.cpu().numpy()
            ys, xs = np.nonzero(g)
            assert len(ys) == 1
            ref = orc.render_density(*[np.array([v], np.float32) for v in p], mvp, res)
            assert np.array_equal(g.view(np.uint64), ref)
            if 1e-3 < fr < 1 - 1e-3:
                assert int(ys[0]) == row, (name, res)
            if 1e-3 < fc < 1 - 1e-3:
                assert int(xs[0]) == col, (name, res)
