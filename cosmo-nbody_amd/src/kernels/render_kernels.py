"""Density splat (src/kernels/render_kernels.py).

``render_density[blocks, threads](pos_x, pos_y, pos_z, grid, mvp, res)``
projects every particle with the MVP matrix and adds 2.5/w to its pixel
(render_kernels.py:35-47).  The grid holds fixed-point uint64 bins (int64
tensor of shape (res, res)): each visible particle adds
rint(2.5/w * 2^frac_bits), so the grid is independent of arrival order.
``to_float`` converts it back to the reference's float density.  A float32
grid (the reference's own, camera.py:26) is filled from such bins.
"""
from __future__ import annotations

import numpy as np
import torch

from ._hip import RENDER_FRAC_BITS, U64, check, kernel, load, ptr, stream_handle

__all__ = ["render_density", "to_float", "colorize", "RENDER_FRAC_BITS"]


def _mvp_host(mvp):
    if isinstance(mvp, torch.Tensor):
        mvp = mvp.detach().to("cpu")
        mvp = mvp.numpy()
    m = np.ascontiguousarray(np.asarray(mvp, dtype=np.float32).reshape(16))
    return m


_bins = {}


def _bins_for(device, res):
    """Fixed-point bins for a float32 grid, one buffer per device and size."""
    key = (device, res)
    b = _bins.get(key)
    if b is None:
        b = torch.zeros((res, res), dtype=torch.int64, device=device)
        _bins[key] = b
    return b


@kernel
def render_density(pos_x, pos_y, pos_z, grid, mvp, res, frac_bits=RENDER_FRAC_BITS,
                   _launch=None):
    """grid: int64/uint64 fixed-point bins (res, res), or the reference's
    float32 density grid (camera.py:26), which gains the splat rounded from
    fixed-point bins (the same value whatever the arrival order)."""
    n = pos_x.shape[0] if _launch is None else min(pos_x.shape[0], _launch[0] * _launch[1])
    res = int(res)
    m = _mvp_host(mvp)
    as_float = isinstance(grid, torch.Tensor) and grid.dtype == torch.float32
    bins = _bins_for(grid.device, res) if as_float else grid
    if as_float:
        ptr(grid, torch.float32, "grid", numel=res * res)
        bins.zero_()
    check(load().cosmo_render_density(
        ptr(pos_x, torch.float32, "pos_x", numel=n), ptr(pos_y, torch.float32, "pos_y", numel=n),
        ptr(pos_z, torch.float32, "pos_z", numel=n), ptr(bins, U64, "grid", numel=res * res),
        m.ctypes.data, res, n, int(frac_bits), stream_handle()), "render_density")
    if as_float:
        check(load().cosmo_grid_add_fixed(ptr(bins, U64, "bins"), ptr(grid, torch.float32, "grid"),
                                          res * res, int(frac_bits), stream_handle()),
              "render_density")


def to_float(grid, frac_bits=RENDER_FRAC_BITS):
    """Fixed-point bins -> float64 density."""
    return grid.to(torch.float64) * (2.0 ** -frac_bits)


def colorize(grid, frac_bits=RENDER_FRAC_BITS, frame=None, rgb=None, scratch=None):
    """Colour mapping of the fixed-point density grid on the GPU
    (src/utils/visualizer.py:115-171): fills frame (float32 [res, res, 3])
    and/or rgb (uint8 [res, res, 3], the (frame * 255).astype(uint8) bytes
    of engine.py:139).  With neither given, returns a new float frame."""
    res = grid.shape[0]
    if frame is None and rgb is None:
        frame = torch.empty((res, res, 3), dtype=torch.float32, device=grid.device)
    scratch = (torch.empty(1, dtype=torch.int64, device=grid.device) if scratch is None
               else scratch)
    check(load().cosmo_colorize_frame(
        ptr(grid, U64, "grid", numel=res * res), res, int(frac_bits),
        ptr(scratch, U64, "scratch", numel=1),
        None if frame is None else ptr(frame, torch.float32, "frame", numel=3 * res * res),
        None if rgb is None else ptr(rgb, torch.uint8, "rgb", numel=3 * res * res),
        stream_handle()), "colorize_frame")
    return frame if frame is not None else rgb

