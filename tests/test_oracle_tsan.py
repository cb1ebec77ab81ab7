"""The oracle's OpenMP loops under ThreadSanitizer (host code only): the
recipe in oracle/Makefile builds oracle/tsan_check.c and cosmo_oracle.c with
LLVM's clang, -fsanitize=thread and libomp with its libarcher annotations,
runs every parallel entry point, and fails on any reported race or on any
bit that differs between 1 and 8 threads."""
import os
import shutil
import subprocess

import pytest

LLVM = os.environ.get("LLVM", "/opt/rocm/llvm")
ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(not (os.path.exists(os.path.join(LLVM, "bin", "clang")) and
                         os.path.exists(os.path.join(LLVM, "lib", "libarcher.so")) and
                         shutil.which("make")),
                    reason="needs LLVM clang with libomp and libarcher")
def test_oracle_openmp_loops_race_free():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "tsan", f"LLVM={LLVM}"],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ThreadSanitizer" not in out, out[-4000:]
    assert "bit-identical" in out
