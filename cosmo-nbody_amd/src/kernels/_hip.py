"""ctypes binding of libcosmo_hip.so (the gfx950 kernels) and the launch
syntax of the reference's Numba kernels.

The reference launches ``kern[blocks, threads](*args)`` (Numba CUDA).  The
drop-ins keep that syntax: the launch configuration only fixes the index
range a kernel covers (``[0, blocks*threads)``, or ``[off, off +
blocks*threads)`` for compute_forces); the library chooses its own
geometry.  Arguments are torch tensors on the HIP device in the reference's
layouts; every call is enqueued on torch's current stream.  There is no
fallback: if the library is missing or a call fails, a RuntimeError says so.
"""
from __future__ import annotations

import ctypes as C
import functools
import os
import threading

import torch

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libcosmo_hip.so")

MASK_FULL21 = 0
MASK_REF10 = 1
MAC_REF = 0
MAC_AABB = 1
TRAV_COMPENSATED = 1
TRAV_STATS = 2
TRAV_PAIRS = 4
TRAV_GROUP_STATS = 8
RENDER_FRAC_BITS = 20

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_f64 = C.c_double
_f32 = C.c_float
_int = C.c_int

# name -> (argtypes, restype), mirroring include/cosmo_abi.h and include/cosmo_mlp.h
SIGNATURES = {
    "cosmo_status_string": ([_int], C.c_char_p),
    "cosmo_ctx_create": ([_i64, C.POINTER(_p)], _int),
    "cosmo_ctx_destroy": ([_p], _int),
    "cosmo_ctx_nodes": ([_p], _p),
    "cosmo_tree_bytes": ([_i64, _int], _i64),
    "cosmo_compute_codes": ([_p, _p, _p, _p, _p, _i64, _int, _p], _int),
    "cosmo_sort_by_code": ([_p, _p, _p, _p, _i64, _int, _p], _int),
    "cosmo_sort_by_code_noalloc": ([_p, _p, _p, _i64, _int, _p], _int),
    "cosmo_permute": ([_p, _int, C.POINTER(_p), C.POINTER(_p), _i64, _p], _int),
    "cosmo_reset_tree_buffers": ([_p, _p, _p, _p, _i64, _p], _int),
    "cosmo_build_tree": ([_p, _p, _p, _i64, _p], _int),
    "cosmo_build_tree_spans": ([_p, _p, _p, _p, _i64, _p], _int),
    "cosmo_compute_multipoles": ([_p] * 11 + [_i64, _p], _int),
    "cosmo_normalize_com": ([_p, _p, _i64, _p], _int),
    "cosmo_compute_multipoles_f32": ([_p] * 12 + [_i64, _p], _int),
    "cosmo_normalize_com_f32": ([_p, _p, _i64, _p], _int),
    "cosmo_pack_nodes_f32": ([_p] * 5 + [_i64, _f64, _int, _p, _p], _int),
    "cosmo_compute_forces_f32": ([_p] * 10 + [_f64] * 4 + [_p, _p, _p, _i64, _i64, _i64, _p],
                                 _int),
    "cosmo_compute_forces_packed": ([_p] * 5 + [_f64] * 4 + [_p, _i64, _i64, _i64, C.c_uint64,
                                                              _p], _int),
    "cosmo_grid_add_fixed": ([_p, _p, _i64, _int, _p], _int),
    "cosmo_count_nonfinite": ([_p, _i64, _p, _p], _int),
    "cosmo_zeldovich_work_bytes": ([_i64], _i64),
    "cosmo_zeldovich_ics": ([_i64, _p, _f64, _f64, _p] + [_p] * 6 + [_p], _int),
    "cosmo_find_root": ([_p, _p, _i64, _p], _int),
    "cosmo_build_top_tree_cache": ([_p, _p, _p, _p, _i64, _p], _int),
    "cosmo_pack_nodes": ([_p] * 5 + [_i64, _f64, _int, _p, _p], _int),
    "cosmo_normalize_pack": ([_p] * 5 + [_i64, _f64, _int, _p, _p], _int),
    "cosmo_normalize_pack_buckets": ([_p] * 6 + [_i64, _f64, _int, _int, _p, _p], _int),
    "cosmo_traverse": ([_p] * 4 + [_i64, _p, _p, _f64, _p, _f64, _f64, _f64, _i64, _i64, _i64,
                                    _int, _int, _p, _p], _int),
    "cosmo_traverse_split": ([_p] * 5 + [_i64, _p, _p, _f64, _f64, _f64, _f64, _i64, _i64, _i64,
                              _p, _p], _int),
    "cosmo_traverse_ordered": ([_p] * 4 + [_i64, _p, _p, _f64, _p, _f64, _f64, _f64, _i64, _i64,
                                            _i64, _int, _int, _p, _p, _p], _int),
    "cosmo_traverse_ex": ([_p] * 4 + [_i64, _p, _p, _f64, _p, _f64, _f64, _f64, _i64, _i64,
                                       _i64, _int, _int, _p, _p, _f64, _p, _p, _p], _int),
    "cosmo_order_groups": ([_p, _i64, _p, _p], _int),
    "cosmo_order_groups_by_pops": ([_p, _i64, _p, _p, _p, _int, _p, _p], _int),
    "cosmo_compute_forces": ([_p] * 10 + [_f64] * 4 + [_p, _p, _p, _i64, _i64, _i64, _p], _int),
    "cosmo_integrate": ([_p] * 7 + [_f64, _f64, _f64, _i64, _p], _int),
    "cosmo_kick_drift": ([_p] * 7 + [_f32, _f32, _f32, _i64, _p], _int),
    "cosmo_render_density": ([_p, _p, _p, _p, _p, _i32, _i64, _int, _p], _int),
    "cosmo_kick_drift_dev": ([_p] * 8 + [_i64, _p], _int),
    "cosmo_render_density_dev": ([_p, _p, _p, _p, _p, _i32, _i64, _int, _p], _int),
    "cosmo_colorize_frame": ([_p, _i32, _int, _p, _p, _p, _p], _int),
    "cosmo_direct_sum": ([_p, _p, _p, _p, _i64, _p, _i64, _f64, _f64, _p, _p], _int),
    "cosmo_validate_tree": ([_p, _p, _p, _i64, _p, _p], _int),
    # include/cosmo_mlp.h
    "cosmo_straddle_profile": ([_p] * 4 + [_i64, _f64, _p, _p, _int, _p], _int),
    "cosmo_group_features": ([_p] * 9 + [_i64, _f64, _f64, _p, _int, _p, _p], _int),
    "cosmo_group_features_dev": ([_p] * 9 + [_i64, _p, _f64, _p, _int, _p, _p], _int),
    "cosmo_straddle_profiles": ([_p] * 4 + [_i64, _p, _int, _p, _p, _int, _p], _int),
    "cosmo_group_features_multi": ([_p] * 9 + [_i64, _f64, _p, _f64, _p, _int, _int, _p, _p],
                                   _int),
    "cosmo_assemble_rows": ([_p, _i64, _p, _p, _p, _int, _p, _p], _int),
    "cosmo_mlp_forward": ([_p, _i64, _p, _p, _p, _p], _int),
    "cosmo_mlp_grad": ([_p, _p, _i64, _p, _p, _p, _int, _p, _p], _int),
    "cosmo_mlp_adam": ([_p, _p, _p, _p, _f32, _f32, _f32, _f32, _int, _p, _p], _int),
    "cosmo_mlp_to_bf16": ([_p, _p, _p], _int),
    "cosmo_select_theta": ([_p, _i64, _p, _p, _p, _int, _f64, _f64, _f64, _p, _p, _p, _p, _p],
                           _int),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load libcosmo_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(make -C cosmo-nbody_amd/csrc)")
                L = C.CDLL(LIB_PATH)
                for name, (args, res) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = res
                _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = load().cosmo_status_string(status).decode()
        raise RuntimeError(f"{what}: {msg} (status {status})")


def stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t, dtypes, name, numel=None, shape=None):
    """Device pointer of tensor t after checking device, dtype, layout."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor on the HIP device, got {type(t)}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: tensor must live on the HIP device (got {t.device})")
    if not isinstance(dtypes, tuple):
        dtypes = (dtypes,)
    if t.dtype not in dtypes:
        raise TypeError(f"{name}: dtype {t.dtype} not in {dtypes}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if numel is not None and t.numel() < numel:
        raise ValueError(f"{name}: needs at least {numel} elements, has {t.numel()}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != {tuple(shape)}")
    return C.c_void_p(t.data_ptr())


U64 = (torch.int64, torch.uint64)
U32 = (torch.int32, torch.uint32)


class Context:
    """Owns cosmo_ctx scratch (sort workspace, packed node records)."""

    def __init__(self, n_max: int):
        self.n_max = int(n_max)
        h = C.c_void_p()
        check(load().cosmo_ctx_create(self.n_max, C.byref(h)), "cosmo_ctx_create")
        self.handle = h
        self.device = torch.cuda.current_device()
        self.packed_key = None  # what the records in cosmo_ctx_nodes were packed from
        self.packed_epoch = 0   # bumped at every pack: keys the library's planned frontiers

    @property
    def nodes_ptr(self):
        return C.c_void_p(load().cosmo_ctx_nodes(self.handle))

    def close(self):
        if self.handle:
            load().cosmo_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_shared_ctx = {}

# Tree generation: bumped by every drop-in that writes tree arrays
# (reset_tree_buffers, build_tree, compute_multipoles, normalize_com,
# find_root and the packing variants).  compute_forces repacks its node records only when the
# generation, the arrays or their torch versions changed since its last pack.
_tree_generation = [0]


def bump_tree_generation() -> None:
    _tree_generation[0] += 1


def tree_generation() -> int:
    return _tree_generation[0]


def shared_context(n: int) -> Context:
    """Per-device context with capacity >= n, grown on demand."""
    dev = torch.cuda.current_device()
    ctx = _shared_ctx.get(dev)
    if ctx is None or ctx.n_max < n:
        if ctx is not None:
            torch.cuda.synchronize()
            ctx.close()
        ctx = Context(max(n, 1))
        _shared_ctx[dev] = ctx
    return ctx


class Kernel:
    """Reference-style launcher: ``kern[blocks, threads](*args)``."""

    def __init__(self, fn):
        self._fn = fn
        functools.update_wrapper(self, fn)

    def __getitem__(self, cfg):
        if not isinstance(cfg, tuple) or len(cfg) < 2:
            raise TypeError("launch configuration must be [blocks, threads]")
        blocks, threads = int(cfg[0]), int(cfg[1])
        if blocks < 0 or threads < 1:
            raise ValueError("invalid launch configuration")
        return functools.partial(self._fn, _launch=(blocks, threads))

    def __call__(self, *args, **kw):
        return self._fn(*args, _launch=None, **kw)


def kernel(fn):
    return Kernel(fn)


def launch_extent(launch, n):
    """Number of indices covered: min(n, blocks*threads), or n if unset."""
    if launch is None:
        return n
    return min(n, launch[0] * launch[1])
