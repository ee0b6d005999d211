"""bench.py's command-line contract (no GPU needed for these paths)."""
import os
import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_default_parts_and_extra_parts():
    sys.path.insert(0, ROOT)
    import bench

    assert bench.PARTS == ("solve", "config4", "shard", "batch", "unet", "warm", "cpu")
    assert set(bench.EXTRA_PARTS) == {"batch3b", "warm5", "prec4", "prec4sweep"}
    assert bench.HBM_PEAK_GBS == 8000.0 and bench.BF16_PEAK_TFLOPS == 2500.0
