import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cosmo-nbody_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


def zeldovich(n_grid, seed=42):
    from src.utils import config as cfg
    from src.utils.cosmology import generate_zeldovich_ics
    c = cfg.preset("small_32", N_GRID=n_grid, SEED=seed)
    return generate_zeldovich_ics(c)


@pytest.fixture(scope="session")
def ics16():
    """Positions of 16^3 Zel'dovich ICs."""
    return zeldovich(16)[0]


@pytest.fixture(scope="session")
def ics32():
    """Positions of 32^3 Zel'dovich ICs."""
    return zeldovich(32)[0]


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
