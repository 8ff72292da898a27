"""bench.py's contract on the CPU side: more than one GPU is refused with a message
(the build measures one MI355X), before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bjontegaard_rate_closed_forms():
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    r = np.array([1e6, 2e6, 4e6, 8e6])
    p = np.array([30.0, 33.0, 36.0, 39.0])
    assert abs(bench.bjontegaard_rate(r, p, r, p)) < 1e-9
    # the same PSNR for 10 % more bits everywhere: +10 %; 5 % fewer: -5 %
    assert abs(bench.bjontegaard_rate(r, p, 1.1 * r, p) - 10.0) < 1e-6
    assert abs(bench.bjontegaard_rate(r, p, 0.95 * r, p) + 5.0) < 1e-6
    # a curve shifted up in PSNR needs fewer bits for the same quality
    assert bench.bjontegaard_rate(r, p, r, p + 0.5) < 0
