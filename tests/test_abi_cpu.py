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
