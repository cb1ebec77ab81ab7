"""bench.py's command line without a GPU: more than one GPU is refused with
a message (one MI355X per run), and the roofline constants are MI355X's."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr
    assert r.stdout.strip() == ""


def test_peaks():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.PEAK_FP32_TFLOPS == 157.3 and bench.PEAK_HBM_GBS == 8000.0
    assert bench.FLOP_PER_INTERACTION == 20
