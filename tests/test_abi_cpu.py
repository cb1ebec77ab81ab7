"""CPU-side checks of the C ABI: the library loads without a GPU and exports
every function include/*.h declares; the ctypes table matches."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if h.endswith(".h"):
            text = open(os.path.join(INCLUDE, h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cosmo_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("cosmo_compute_codes", "cosmo_sort_by_code", "cosmo_build_tree",
                 "cosmo_compute_multipoles", "cosmo_compute_forces", "cosmo_integrate",
                 "cosmo_render_density", "cosmo_traverse"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from src.kernels import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.skip("libcosmo_hip.so not built")
    lib = ctypes.CDLL(_hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from src.kernels import _hip
    assert sorted(_hip.SIGNATURES) == declared_functions()


def test_status_strings():
    from src.kernels import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.skip("libcosmo_hip.so not built")
    L = _hip.load()
    assert L.cosmo_status_string(0) == b"ok"
    assert b"invalid" in L.cosmo_status_string(1)
    assert b"malformed" in L.cosmo_status_string(4)


def test_invalid_arguments_rejected_without_gpu():
    """Argument checks run on the host before any launch."""
    from src.kernels import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.skip("libcosmo_hip.so not built")
    L = _hip.load()
    assert L.cosmo_compute_codes(None, None, None, None, None, 10, 0, None) == 1
    assert L.cosmo_compute_codes(None, None, None, None, None, 0, 0, None) == 0
    assert L.cosmo_sort_by_code(None, None, None, None, 10, 63, None) == 1
    assert L.cosmo_traverse(None, None, None, None, 10, None, None, 0.5, None, 1.0, 0.0, 1.0, 0,
                            10, 10, 7, 0, None, None) == 1
    assert L.cosmo_tree_bytes(10, 0) == 20 * 32  # one record per node id and the sentinel
    assert L.cosmo_tree_bytes(10, 1) == 10 * 16 + 10 * 64
    assert L.cosmo_render_density(None, None, None, None, None, 0, 10, 20, None) == 1


def test_size_limits_refused_without_gpu():
    """The walk addresses its records with 32-bit byte offsets: a tree whose
    records pass 4 GiB is refused with COSMO_E_TOO_LARGE (3) before any
    launch, at exactly the boundary (2^26 leaves fill 4 GiB of reference-test
    records; AABB records take 80 bytes per leaf); a tree at the boundary
    passes the size check and stops at the null pointers (1).  Particle
    counts above 2^29 are invalid at every stage."""
    from src.kernels import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.skip("libcosmo_hip.so not built")
    L = _hip.load()
    ok, invalid, too_large = 0, 1, 3

    def trav(n_tree, mac):
        return L.cosmo_traverse(None, None, None, None, n_tree, None, None, 0.5, None, 1.0, 0.0,
                                1.0, 0, 64, 64, mac, 0, None, None)

    assert L.cosmo_tree_bytes(1 << 26, 0) == 1 << 32
    assert trav(1 << 26, 0) == invalid
    assert trav((1 << 26) + 1, 0) == too_large
    aabb_max = (1 << 32) // 80
    assert L.cosmo_tree_bytes(aabb_max, 1) <= 1 << 32 < L.cosmo_tree_bytes(aabb_max + 1, 1)
    assert trav(aabb_max, 1) == invalid
    assert trav(aabb_max + 1, 1) == too_large
    big = (1 << 29) + 1
    assert L.cosmo_compute_codes(None, None, None, None, None, big, 0, None) == invalid
    assert L.cosmo_build_tree(None, None, None, big, None) == invalid
    assert L.cosmo_build_tree(None, None, None, 1, None) == ok  # one leaf: nothing to build
    assert L.cosmo_compute_multipoles(None, None, None, None, None, None, None, None, None,
                                      None, None, big, None) == invalid
    assert trav(big, 0) == invalid


def test_zeldovich_ics_arguments_checked_on_host():
    """Sizes outside [1, 1024] and null pointers are refused before any HIP
    call (so this runs without a GPU)."""
    from src.kernels import _hip
    if not os.path.exists(_hip.LIB_PATH):
        pytest.skip("libcosmo_hip.so not built")
    L = _hip.load()
    assert L.cosmo_zeldovich_work_bytes(0) == 0
    assert L.cosmo_zeldovich_work_bytes(2048) == 0
    assert L.cosmo_zeldovich_work_bytes(4) == 64 * 16 + 3 * 64 * 8 + 1026 * 8
    assert L.cosmo_zeldovich_ics(0, None, 0.3, 1.0, None, None, None, None, None, None, None,
                                 None) == 1
    assert L.cosmo_zeldovich_ics(4, None, 0.3, 1.0, None, None, None, None, None, None, None,
                                 None) == 1
    assert b"hipFFT" in L.cosmo_status_string(5)
