"""Zel'dovich initial conditions and the Hubble rate.

``generate_zeldovich_ics`` restates src/utils/cosmology.py:12-73 on the host
with numpy: P(k) = k^-2.5, uniform phases in [0, 2pi), displacement
d = Re ifftn(i k / (k^2 + 1e-10) * sqrt(P) e^{i phi}), rescaled to
std(dx) = 0.3 / n_grid, added to the ij-ordered lattice and wrapped into the
unit box; velocities d * H(a_start) * Om(a_start)^0.55 * a_start / 20.  The
phases come from numpy's PCG64 seeded with config.SEED, so the realisation
differs from CuPy's generator; the statistics (std, zero mean, v/d ratio)
are those of the reference.  Positions and velocities are returned as FP32.
"""
from __future__ import annotations

import numpy as np


def get_hubble_param(a, config):
    """src/utils/cosmology.py:76-85: H0 * sqrt(Om a^-3 + OL), FP64."""
    return config.H0 * np.sqrt(config.OMEGA_M * a ** -3 + config.OMEGA_L)


def velocity_coupling(config) -> float:
    """src/utils/cosmology.py:64-67."""
    a0 = config.A_START
    h_start = config.H0 * np.sqrt(config.OMEGA_M * a0 ** -3 + config.OMEGA_L)
    om_a = config.OMEGA_M * a0 ** -3 / (config.OMEGA_M * a0 ** -3 + config.OMEGA_L)
    return float(h_start * om_a ** 0.55 * a0 / 20)


def draw_phases(n_grid: int, seed: int):
    """Phases U[0, 2 pi) of the n^3 modes (cosmology.py:35), PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(0, 2 * np.pi, (n_grid, n_grid, n_grid))


def displacement_field(n_grid: int, seed: int):
    """Unscaled Zel'dovich displacement (cosmology.py:27-46), FP64 [3, n^3]."""
    k = np.fft.fftfreq(n_grid) * n_grid
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k_mag = np.sqrt(k2)
    k_mag[0, 0, 0] = 1e-10
    P = k_mag ** -2.5
    P[0, 0, 0] = 0
    phase = draw_phases(n_grid, seed)
    delta_k = np.sqrt(P) * (np.cos(phase) + 1j * np.sin(phase))
    del phase, P
    k_mag2 = k_mag ** 2 + 1e-10
    out = np.empty((3, n_grid ** 3), np.float64)
    for c, kc in enumerate((kx, ky, kz)):
        term = 1j * kc / k_mag2
        term[0, 0, 0] = 0
        out[c] = np.fft.ifftn(term * delta_k).real.reshape(-1)
    return out


def generate_zeldovich_ics(config, seed=None):
    """Returns ([pos_x, pos_y, pos_z], [vel_x, vel_y, vel_z]) as FP32 numpy."""
    n_grid = config.N_GRID
    seed = config.SEED if seed is None else seed
    d = displacement_field(n_grid, seed)
    scale = (1.0 / n_grid) * getattr(config, "ZA_AMPLITUDE", 0.3) / np.std(d[0])
    d *= scale
    lin = np.linspace(0, 1, n_grid, endpoint=False)
    gx, gy, gz = np.meshgrid(lin, lin, lin, indexing="ij")
    grid = (gx.reshape(-1), gy.reshape(-1), gz.reshape(-1))
    coupling = velocity_coupling(config)
    pos = [((g + d[c]) % 1.0).astype(np.float32) for c, g in enumerate(grid)]
    vel = [(d[c] * coupling).astype(np.float32) for c in range(3)]
    return pos, vel


def generate_zeldovich_ics_device(config, seed=None, device="cuda"):
    """generate_zeldovich_ics with everything after the phase draw on the GPU
    (cosmo_zeldovich_ics: spectrum, three hipFFT inverse transforms, std,
    lattice + displacement): the same phases, and FP32 outputs equal to the
    host version's except where the FFT libraries' different rounding
    crosses an FP32 rounding boundary.  Returns ([pos_x, pos_y, pos_z],
    [vel_x, vel_y, vel_z]) as float32 tensors on `device`."""
    import torch
    from src.kernels import _hip
    n_grid = int(config.N_GRID)
    seed = config.SEED if seed is None else seed
    dev = torch.device(device)
    phase = torch.from_numpy(draw_phases(n_grid, seed).reshape(-1)).to(dev)
    L = _hip.load()
    work = torch.empty(int(L.cosmo_zeldovich_work_bytes(n_grid)), dtype=torch.uint8, device=dev)
    n3 = n_grid ** 3
    pos = [torch.empty(n3, dtype=torch.float32, device=dev) for _ in range(3)]
    vel = [torch.empty(n3, dtype=torch.float32, device=dev) for _ in range(3)]
    _hip.check(L.cosmo_zeldovich_ics(
        n_grid, _hip.ptr(phase, torch.float64, "phase"),
        float(getattr(config, "ZA_AMPLITUDE", 0.3)), velocity_coupling(config),
        _hip.ptr(work, torch.uint8, "work"),
        *[_hip.ptr(t, torch.float32, "pos") for t in pos],
        *[_hip.ptr(t, torch.float32, "vel") for t in vel], _hip.stream_handle()),
        "zeldovich_ics")
    return pos, vel


def generate_halo_ics(config, seed=None):
    """A heterogeneous test state with no reference counterpart: a fraction
    HALO_FRACTION of the particles in N_HALOS Plummer spheres (scale radii
    log-uniform in HALO_RADII, truncated at 10 scale radii, equal particle
    counts) at uniform random centres, the rest uniform in the box; zero
    velocities.  Used where the error/cost trade-off must differ between
    groups (the reference's Zel'dovich states stay close to uniform)."""
    n = config.N_GRID ** 3
    seed = config.SEED if seed is None else seed
    rng = np.random.Generator(np.random.PCG64(seed))
    n_halo = int(round(n * config.HALO_FRACTION))
    k = max(1, int(config.N_HALOS))
    lo, hi = config.HALO_RADII
    radii = np.exp(rng.uniform(np.log(lo), np.log(hi), k))
    centres = rng.uniform(0.0, 1.0, (k, 3))
    which = np.arange(n_halo) % k
    u = rng.uniform(1e-6, 1.0, n_halo)
    r = radii[which] / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, 10.0 * radii[which])
    v = rng.normal(size=(n_halo, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    halo = centres[which] + v * r[:, None]
    bg = rng.uniform(0.0, 1.0, (n - n_halo, 3))
    allp = np.concatenate([halo, bg]) % 1.0
    allp = allp[rng.permutation(n)]
    pos = [allp[:, c].astype(np.float32) for c in range(3)]
    pos = [np.where(p >= 1.0, np.float32(0.0), p) for p in pos]
    vel = [np.zeros(n, np.float32) for _ in range(3)]
    return pos, vel


def generate_ics(config, seed=None):
    """Initial conditions of the configured kind (IC_KIND)."""
    if getattr(config, "IC_KIND", "zeldovich") == "halos":
        return generate_halo_ics(config, seed)
    return generate_zeldovich_ics(config, seed)
