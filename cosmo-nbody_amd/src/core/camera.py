"""Density capture (src/core/camera.py:10-48).

``Camera(n_steps, resolution, config).capture(step, px, py, pz)`` clears
the fixed-point grid, builds the orbiting MVP on the host and splats every
particle with render_density; it returns the (res, res) int64 grid.
``frame_bytes()`` colours it on the GPU (visualizer.py:115-171) into the
uint8 [res, res, 3] frame the reference's run loop stores
(engine.py:139); PNG/video writing is not part of this package.
"""
from __future__ import annotations

import torch

from src.kernels.render_kernels import RENDER_FRAC_BITS, colorize, render_density, to_float
from src.utils.visualizer import get_mvp_matrix


class Camera:
    def __init__(self, n_steps, resolution, config, device="cuda"):
        self.config = config
        self.n = n_steps
        self.res = int(resolution)
        self.frac_bits = int(getattr(config, "RENDER_FRAC_BITS", RENDER_FRAC_BITS))
        self.grid_d = torch.zeros((self.res, self.res), dtype=torch.int64, device=device)
        self.rgb_d = torch.zeros((self.res, self.res, 3), dtype=torch.uint8, device=device)
        self._max_d = torch.zeros(1, dtype=torch.int64, device=device)
        b = config.BOX_SIZE
        self.box_center = (b / 2, b / 2, b / 2)

    def capture(self, step, px_d, py_d, pz_d):
        self.grid_d.zero_()
        mvp = get_mvp_matrix(step, self.box_center, self.config.BOX_SIZE, self.config)
        render_density(px_d, py_d, pz_d, self.grid_d, mvp, self.res, frac_bits=self.frac_bits)
        return self.grid_d

    def density(self):
        return to_float(self.grid_d, self.frac_bits)

    def frame_bytes(self):
        """The coloured frame of the last capture as uint8 [res, res, 3]."""
        return colorize(self.grid_d, self.frac_bits, rgb=self.rgb_d, scratch=self._max_d)
