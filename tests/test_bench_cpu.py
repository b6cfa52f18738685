"""bench.py's contract, on the CPU: one GPU only (--gpus above 1 exits with a
message and status 2 before any device work), and the per-launch byte counts
behind the roofline, against figures worked out by hand from the 4096^2 FP32
layouts (1367 kept kx columns, 2731 kept ky rows, 8-byte complex values)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_kernel_bytes_at_4096_fp32():
    sys.path.insert(0, ROOT)
    import bench
    own, model = bench.kernel_bytes(4096, True)
    mixed = 1367 * 4096 * 8            # 44,793,856 B
    compact = 1367 * 2731 * 8          # 29,866,216 B
    half = 4096 * 2049 * 8             # 67,141,632 B
    assert own["k_row"] == 5 * mixed == 223_969_280
    assert own["k_col"] == pytest.approx(5 * mixed + 2.75 * compact)
    assert model["k_row"] == 5 * half == 335_708_160
    assert model["k_col"] == 11 * half == 738_557_952
    # FP64 doubles every figure
    own64, model64 = bench.kernel_bytes(4096, False)
    assert own64["k_row"] == 2 * own["k_row"] and model64["k_col"] == 2 * model["k_col"]
