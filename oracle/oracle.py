"""numpy front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Every function
takes and returns numpy arrays in the reference's layouts and restates one
stage of the reference (file:line given per function).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MASK_FULL21 = 0
MASK_REF10 = 1
RENDER_FRAC_BITS = 20
MAC_REF = 0
MAC_AABB = 1

_i64 = C.c_int64
_i32 = C.c_int32
_f64 = C.c_double
_f32 = C.c_float
_p = C.c_void_p


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_compute_codes": [_p, _p, _p, _p, _p, _i64, C.c_int],
            "orc_argsort_stable": [_p, _p, _i64],
            "orc_reset_tree_buffers": [_p, _p, _p, _p, _i64],
            "orc_build_tree": [_p, _p, _p, _i64],
            "orc_compute_multipoles": [_p] * 11 + [_i64],
            "orc_normalize_com": [_p, _p, _i64],
            "orc_find_root": [_p, _p, _i64],
            "orc_build_top_tree_cache": [_p, _p, _p, _p, _i64],
            "orc_validate_tree": [_p, _p, _i64],
            "orc_compute_forces": [_p] * 9 + [_f64] * 4 + [_i32, _i64, _i64, _i64,
                                                           C.c_int, C.c_int, _p, C.c_int, _f64,
                                                           _p, _p, _p, _p, _p, _p],
            "orc_direct_sum": [_p, _p, _p, _p, _i64, _p, _i64, _f64, _f64, _p],
            "orc_kick_drift": [_p] * 7 + [_f32] * 3 + [_i64],
            "orc_integrate_coeffs": [_f64, _f64, _f64, _p, _p, _p],
            "orc_render_density": [_p, _p, _p, _p, _p, _i32, _i64, C.c_int],
            "orc_num_threads": [],
            "orc_set_num_threads": [C.c_int],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int if name in ("orc_validate_tree", "orc_num_threads") else None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(t: int) -> None:
    lib().orc_set_num_threads(int(t))


def compute_codes(px, py, pz, mask_set=MASK_FULL21):
    """src/kernels/morton.py:40-91 -> (codes uint64[n], indices uint32[n])."""
    px, py, pz = (_c(a, np.float32) for a in (px, py, pz))
    n = px.shape[0]
    codes = np.zeros(n, np.uint64)
    idx = np.zeros(n, np.uint32)
    lib().orc_compute_codes(_ptr(px), _ptr(py), _ptr(pz), _ptr(codes), _ptr(idx), n, mask_set)
    return codes, idx


def argsort_stable(codes):
    """src/core/state.py:35 (argsort on (code, index))."""
    codes = _c(codes, np.uint64)
    perm = np.zeros(codes.shape[0], np.uint32)
    lib().orc_argsort_stable(_ptr(codes), _ptr(perm), codes.shape[0])
    return perm


def build_tree(codes):
    """src/kernels/tree_kernels.py:236-262 then :60-129 -> (children[M,2], parents[M])."""
    codes = _c(codes, np.uint64)
    n = codes.shape[0]
    M = 2 * n - 1
    children = np.full((M, 2), -1, np.int32)
    parents = np.full(M, -1, np.int32)
    lib().orc_build_tree(_ptr(codes), _ptr(children), _ptr(parents), n)
    return children, parents


def compute_multipoles(px, py, pz, mass, children, parents):
    """src/kernels/tree_kernels.py:131-191 in FP64 -> (n_mass, n_com_sum, n_min, n_max)."""
    px, py, pz, mass = (_c(a, np.float32) for a in (px, py, pz, mass))
    children = _c(children, np.int32)
    parents = _c(parents, np.int32)
    M = parents.shape[0]
    n = px.shape[0]
    n_mass = np.zeros(M, np.float64)
    n_com = np.zeros((M, 3), np.float64)
    n_min = np.zeros((M, 3), np.float32)
    n_max = np.zeros((M, 3), np.float32)
    counts = np.zeros(M, np.int32)
    lib().orc_compute_multipoles(_ptr(px), _ptr(py), _ptr(pz), _ptr(mass), _ptr(children),
                                 _ptr(parents), _ptr(n_mass), _ptr(n_com), _ptr(n_min),
                                 _ptr(n_max), _ptr(counts), n)
    return n_mass, n_com, n_min, n_max


def normalize_com(n_mass, n_com):
    """src/kernels/tree_kernels.py:193-214 (returns a normalised copy)."""
    n_mass = _c(n_mass, np.float64)
    out = np.array(n_com, dtype=np.float64, order="C", copy=True)
    lib().orc_normalize_com(_ptr(n_mass), _ptr(out), n_mass.shape[0])
    return out


def find_root(parents):
    """src/kernels/tree_kernels.py:216-234."""
    parents = _c(parents, np.int32)
    r = np.zeros(1, np.int32)
    lib().orc_find_root(_ptr(parents), _ptr(r), parents.shape[0])
    return int(r[0])


def build_top_tree_cache(children, root, cache_size=256):
    """src/kernels/tree_kernels.py:264-314 -> (top_nodes, node_to_cache)."""
    children = _c(children, np.int32)
    M = children.shape[0]
    top = np.full(cache_size, -1, np.int32)
    n2c = np.full(M, -1, np.int32)
    r = np.array([root], np.int32)
    lib().orc_build_top_tree_cache(_ptr(children), _ptr(r), _ptr(top), _ptr(n2c), cache_size)
    return top, n2c


def validate_tree(children, parents, n):
    return int(lib().orc_validate_tree(_ptr(_c(children, np.int32)), _ptr(_c(parents, np.int32)), n))


def compute_forces(px, py, pz, children, n_mass, n_com, n_min, n_max, theta, G, soft,
                   box, root, off=0, count=None, precision=64, mac=MAC_REF, theta_g=None,
                   counts=False, want_f64=False, bucket=0, straddle_mass=0.0, targets=None,
                   straddle_g=None, drops=False):
    """src/kernels/physics_kernels.py:14-153.

    Node arrays may be FP64 (as produced by the FP64 upward pass); they are
    rounded to FP32, the storage type of the reference's node arrays
    (src/core/octree.py:37-40), before the traversal.
    Returns force float32[n,3] (rows outside [off, off+count) are zero) and,
    if counts, (n_inter, n_visit) int64[n]; if drops, the per-particle
    overflow drops int64[n]; if want_f64 also the FP64 sums.
    bucket (AABB test only): rejected internal nodes of at most `bucket`
    leaves are summed leaf by leaf in leaf order instead of being opened.
    straddle_mass (AABB test only): nodes lighter than straddle_mass pass the
    minimum-image guard.
    straddle_g: optional per-particle straddle mass (float32[n]).
    targets: optional particle indices; then every output has one row per
    target (row t belongs to particle targets[t]) and off/count are ignored.
    """
    px, py, pz = (_c(a, np.float32) for a in (px, py, pz))
    n = px.shape[0]
    if count is None:
        count = n - off
    children = _c(children, np.int32)
    nm = _c(np.asarray(n_mass).astype(np.float32), np.float32)
    nc = _c(np.asarray(n_com).astype(np.float32), np.float32)
    nmin = _c(n_min, np.float32)
    nmax = _c(n_max, np.float32)
    rows = n
    if targets is not None:
        targets = _c(targets, np.int64)
        if targets.size and (targets.min() < 0 or targets.max() >= n):
            raise ValueError("target index out of range")
        rows = count = targets.shape[0]
        off = 0
    force = np.zeros((rows, 3), np.float32)
    tg = None if theta_g is None else _c(theta_g, np.float32)
    sg = None if straddle_g is None else _c(straddle_g, np.float32)
    ni = np.zeros(rows, np.int64) if counts else None
    nv = np.zeros(rows, np.int64) if counts else None
    f64 = np.zeros((rows, 3), np.float64) if want_f64 else None
    nd = np.zeros(rows, np.int64) if drops else None
    lib().orc_compute_forces(_ptr(px), _ptr(py), _ptr(pz), _ptr(children), _ptr(nm), _ptr(nc),
                             _ptr(nmin), _ptr(nmax), _ptr(force), float(theta), float(G),
                             float(soft), float(box), int(root), int(off), int(count), n,
                             int(precision), int(mac), _ptr(tg), int(bucket),
                             float(straddle_mass), _ptr(sg), _ptr(targets), _ptr(ni), _ptr(nv),
                             _ptr(nd), _ptr(f64))
    out = [force]
    if counts:
        out += [ni, nv]
    if drops:
        out.append(nd)
    if want_f64:
        out.append(f64)
    return out[0] if len(out) == 1 else tuple(out)


def direct_sum(px, py, pz, mass, targets, G, soft):
    """FP64 direct sum for the target indices -> float64[t,3]."""
    px, py, pz, mass = (_c(a, np.float32) for a in (px, py, pz, mass))
    targets = _c(targets, np.int64)
    acc = np.zeros((targets.shape[0], 3), np.float64)
    lib().orc_direct_sum(_ptr(px), _ptr(py), _ptr(pz), _ptr(mass), px.shape[0], _ptr(targets),
                         targets.shape[0], float(G), float(soft), _ptr(acc))
    return acc


def integrate_coeffs(a, H, dt):
    cd, ck, cf = _f32(), _f32(), _f32()
    lib().orc_integrate_coeffs(float(a), float(H), float(dt), C.byref(cd), C.byref(ck),
                               C.byref(cf))
    return cd.value, ck.value, cf.value


def integrate(px, py, pz, vx, vy, vz, force, a, H, dt):
    """src/kernels/physics_kernels.py:155-203 (FP32 kick-drift); returns new copies."""
    arrs = [np.array(x, np.float32, order="C", copy=True) for x in (px, py, pz, vx, vy, vz)]
    force = _c(force, np.float32)
    cd, ck, cf = integrate_coeffs(a, H, dt)
    lib().orc_kick_drift(*(_ptr(x) for x in arrs), _ptr(force), cd, ck, cf, arrs[0].shape[0])
    return tuple(arrs)


def render_density(px, py, pz, mvp, res, frac_bits=RENDER_FRAC_BITS):
    """src/kernels/render_kernels.py:12-47 into fixed-point uint64 bins."""
    px, py, pz = (_c(a, np.float32) for a in (px, py, pz))
    mvp = _c(mvp, np.float32).reshape(16)
    grid = np.zeros((res, res), np.uint64)
    lib().orc_render_density(_ptr(px), _ptr(py), _ptr(pz), _ptr(grid), _ptr(mvp), int(res),
                             px.shape[0], int(frac_bits))
    return grid


def colorize_frame(grid):
    """src/utils/visualizer.py:115-171 in numpy float32, as the reference's
    CuPy float32 arrays compute it (Python-float operands take the array's
    float32 type).  grid: float32 density [res, res].  Returns the float32
    frame [res, res, 3] and the bytes (frame * 255).astype(uint8) that
    engine.py:139 stores."""
    g = np.asarray(grid, np.float32)
    img = np.log1p(g * np.float32(120.0))
    img = img / (img.max() + np.float32(1e-9))
    val = np.power(img, np.float32(2.2))
    T1, T2, T3 = np.float32(0.6), np.float32(0.72), np.float32(0.82)
    r = np.zeros_like(val)
    gg = np.zeros_like(val)
    b = np.zeros_like(val)
    m1 = val <= T1
    b[m1] = (val[m1] / T1) * np.float32(0.8)
    m2 = (val > T1) & (val <= T2)
    t = (val[m2] - T1) / np.float32(0.72 - 0.6)
    b[m2] = (np.float32(1.0) - t) * np.float32(0.8)
    gg[m2] = t * np.float32(0.9)
    m3 = (val > T2) & (val <= T3)
    t = (val[m3] - T2) / np.float32(0.82 - 0.72)
    gg[m3] = np.float32(0.9) + t * np.float32(0.1)
    r[m3] = t
    m4 = val > T3
    t = (val[m4] - T3) / np.float32(1.0 - 0.82)
    r[m4] = np.float32(1.0)
    gg[m4] = np.float32(1.0) - t
    frame = np.clip(np.stack([r, gg, b], axis=-1) * np.float32(1.3), 0, 1).astype(np.float32)
    return frame, (frame * np.float32(255)).astype(np.uint8)

