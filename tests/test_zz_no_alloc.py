"""cosmo_sort_by_code_noalloc (the sort without allocation, workspace in a
static device array) against cosmo_sort_by_code (workspace in a context),
through the C ABI: the same sorted codes and permutation bit for bit on two
calls in a row, the second smaller, and sizes beyond the array refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

COSMO_OK, COSMO_E_TOO_LARGE = 0, 3


def _sorts(L, ctx, codes_np, bits, in_place):
    import torch
    from src.kernels import _hip
    n = len(codes_np)
    res = []
    for which in ("ctx", "noalloc"):
        c = torch.from_numpy(codes_np.view(np.int64)).cuda()
        out = c if in_place else torch.empty_like(c)
        perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        args = (_hip.ptr(c, _hip.U64, "c"), _hip.ptr(out, _hip.U64, "o"),
                _hip.ptr(perm, _hip.U32, "p"), n, bits, _hip.stream_handle())
        rc = (L.cosmo_sort_by_code(ctx.handle, *args) if which == "ctx"
              else L.cosmo_sort_by_code_noalloc(*args))
        assert rc == COSMO_OK, which
        torch.cuda.synchronize()
        res.append((out.cpu().numpy().view(np.uint64), perm.cpu().numpy().view(np.uint32)))
    return res


def test_noalloc_sort_matches_context_sort():
    import torch
    from src.kernels import _hip
    L = _hip.load()
    rng = np.random.default_rng(77)
    big = rng.integers(0, 2 ** 63, size=1_500_001, dtype=np.int64).astype(np.uint64)
    small = (rng.integers(0, 5, size=4097).astype(np.uint64) << np.uint64(33))  # ties
    ctx = _hip.Context(len(big))
    for codes, bits, in_place in ((big, 63, False), (small, 40, True)):
        (a_out, a_perm), (b_out, b_perm) = _sorts(L, ctx, codes, bits, in_place)
        assert np.array_equal(a_perm, b_perm) and np.array_equal(a_out, b_out)
        assert np.array_equal(b_perm, np.argsort(codes, kind="stable"))
        assert np.array_equal(b_out, codes[b_perm])
    ctx.close()
    # beyond the static array: refused before any launch
    t = torch.zeros(1, dtype=torch.int64, device="cuda")
    p = _hip.ptr(t, _hip.U64, "t")
    rc = L.cosmo_sort_by_code_noalloc(p, p, p, (1 << 21) + 1, 63, _hip.stream_handle())
    assert rc == COSMO_E_TOO_LARGE
    assert "capacity" in L.cosmo_status_string(rc).decode()
