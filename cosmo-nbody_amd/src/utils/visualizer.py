"""Camera matrices for the density splat (src/utils/visualizer.py:9-112).

The MVP matrix is built on the host in numpy float32 exactly as the
reference builds it (look_at, perspective, proj @ view), and returned as a
float32[4, 4] numpy array; render_density takes its 16 values as launch
arguments.  colorize_frame (visualizer.py:115-171) runs as one fused gfx950
pass over the fixed-point grid (render_kernels.colorize).
"""
from __future__ import annotations

import numpy as np


def normalize(v):
    """visualizer.py:9-20."""
    norm = np.linalg.norm(v)
    return v if norm == 0 else v / norm


def look_at(eye, target, up):
    """visualizer.py:22-48."""
    z = normalize(eye - target)
    x = normalize(np.cross(up, z))
    y = normalize(np.cross(z, x))
    view = np.identity(4, dtype=np.float32)
    view[0, :3] = x
    view[1, :3] = y
    view[2, :3] = z
    view[0, 3] = -np.dot(x, eye)
    view[1, 3] = -np.dot(y, eye)
    view[2, 3] = -np.dot(z, eye)
    return view


def perspective(fov_deg, aspect, near, far):
    """visualizer.py:50-78."""
    fov_rad = np.radians(fov_deg)
    f = 1.0 / np.tan(fov_rad / 2.0)
    proj = np.zeros((4, 4), dtype=np.float32)
    proj[0, 0] = f / aspect
    proj[1, 1] = f
    proj[2, 2] = (far + near) / (near - far)
    proj[2, 3] = (2 * far * near) / (near - far)
    proj[3, 2] = -1.0
    return proj


def get_mvp_matrix(step, box_center, box_size, config):
    """visualizer.py:81-112: orbiting camera, returns float32[4, 4] (host)."""
    angle = step * config.ROTATION_SPEED
    radius = box_size * config.CAM_DIST_MULT
    eye_x = box_center[0] + radius * np.cos(angle)
    eye_y = box_center[1] + (box_size * 0.2)
    eye_z = box_center[2] + radius * np.sin(angle)
    eye = np.array([eye_x, eye_y, eye_z], dtype=np.float32)
    target = np.array(box_center, dtype=np.float32)
    up = np.array([0, 1, 0], dtype=np.float32)
    view = look_at(eye, target, up)
    proj = perspective(config.FOV, 1.0, 0.1, radius * 3.0)
    return np.ascontiguousarray(proj @ view, dtype=np.float32)


def colorize_frame(grid_d, frac_bits=None):
    """src/utils/visualizer.py:115-171: the float32 frame [res, res, 3] in
    [0, 1] of the density grid (the fixed-point grid of render_density)."""
    from src.kernels.render_kernels import RENDER_FRAC_BITS, colorize
    return colorize(grid_d, RENDER_FRAC_BITS if frac_bits is None else frac_bits)

