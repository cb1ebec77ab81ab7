"""The per-kernel harness keeps the reference harness's layout
(benchmark/Performance_Benchmarking.py:44-74 sizes and constants, :407-417
CSV header), so the two CSV files line up column by column.  CPU only: the
constants and the header, not the timings."""
import os
import sys

import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cosmo-nbody_amd")


@pytest.fixture(scope="module")
def harness():
    if PKG not in sys.path:
        sys.path.insert(0, PKG)
    from benchmark import Performance_Benchmarking as H
    return H


def test_constants_match_reference(harness):
    H = harness
    # benchmark/Performance_Benchmarking.py:44-72
    assert H.N_VALUES == [10_000, 15_000, 25_000, 37_500, 50_000, 75_000, 100_000, 175_000,
                          250_000, 375_000, 500_000, 750_000, 1_000_000, 1_500_000, 2_000_000,
                          3_000_000]
    assert (H.WARMUP, H.ITERATIONS, H.TPB, H.CACHE_SIZE, H.RES) == (5, 10, 256, 256, 1024)
    assert (H.BOX_SIZE, H.SOFTENING, H.G) == (1.0, 0.001, 1.0)
    assert H.THETA_LIST == [0.3, 0.5, 0.7, 0.9, 1.1] and H.THETA_PIE == 0.7
    assert H.bpg(5000) == 20 and H.bpg(256) == 1


def test_header_starts_with_reference_columns(harness):
    H = harness
    # :407-409: N, the nine stage kernels in dict order, compute_forces, one
    # column per theta
    ref = (["N", "compute_codes", "reset_tree", "build_tree", "compute_multipoles",
            "normalize_com", "find_root", "build_cache", "integrate", "render_density",
            "compute_forces"] + [f"compute_forces_theta_{t}" for t in H.THETA_LIST])
    header = (["N"] + H.KERNEL_COLUMNS + ["compute_forces"]
              + [f"compute_forces_theta_{t}" for t in H.THETA_LIST])
    assert header == ref


def test_fixed_camera_matches_reference_construction(harness):
    """:303-321: look-at from (1.5, 0.2, 1.5) to the box centre, 60 degrees,
    near 0.1, far 100; the box centre projects to the middle of the frame."""
    import numpy as np
    m = harness._bench_mvp()
    c = m @ np.array([0.5, 0.5, 0.5, 1.0], np.float32)
    assert m.shape == (4, 4) and m.dtype == np.float32
    assert abs(c[0] / c[3]) < 1e-6 and abs(c[1] / c[3]) < 1e-6
    assert c[3] > 0.01
