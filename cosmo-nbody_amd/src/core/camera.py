"""Density capture (src/core/camera.py:10-48).

``Camera(n_steps, resolution, config).capture(step, px, py, pz)`` clears
the fixed-point grid, builds the orbiting MVP on the host and splats every
particle with render_density; it returns the (res, res) int64 grid.  The
reference then colours the frame and writes PNG/video; those presentation
steps are not part of this package.
"""
from __future__ import annotations

import torch

from src.kernels.render_kernels import RENDER_FRAC_BITS, render_density, to_float
from src.utils.visualizer import get_mvp_matrix


class Camera:
    def __init__(self, n_steps, resolution, config, device="cuda"):
        self.config = config
        self.n = n_steps
        self.res = int(resolution)
        self.frac_bits = int(getattr(config, "RENDER_FRAC_BITS", RENDER_FRAC_BITS))
        self.grid_d = torch.zeros((self.res, self.res), dtype=torch.int64, device=device)
        b = config.BOX_SIZE
        self.box_center = (b / 2, b / 2, b / 2)

    def capture(self, step, px_d, py_d, pz_d):
        self.grid_d.zero_()
        mvp = get_mvp_matrix(step, self.box_center, self.config.BOX_SIZE, self.config)
        render_density(px_d, py_d, pz_d, self.grid_d, mvp, self.res, frac_bits=self.frac_bits)
        return self.grid_d

    def density(self):
        return to_float(self.grid_d, self.frac_bits)
