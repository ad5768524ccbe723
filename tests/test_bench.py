"""bench.py's CPU-side contract: a multi-GPU request exits with a message
(this build runs on one GPU), and the roofline helpers compute the stated
byte models."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_poisson_roofline_bytes():
    class D:
        nx = ny = nz = 257
        mg_precision = 1

    r = bench.poisson_roofline(D, 4.0, 4.5)
    n = 255 ** 3
    assert r["fine_unknowns"] == n
    assert r["algorithmic_bytes_per_solve"] == n * (4.5 * 123 + 56)
    assert abs(r["achieved"] - r["algorithmic_bytes_per_solve"] / 4.0e-3 / 1e9) < 1e-6
    assert r["frac"] == r["achieved"] / bench.HBM_PEAK_GBS
    D.mg_precision = 0
    assert bench.poisson_roofline(D, 4.0, 1.0)["algorithmic_bytes_per_solve"] == n * (166 + 40 + 56)
    # pairs of cycles sharing one FP64 update, with a residual pass between the two
    D.mg_precision = 1
    r = bench.poisson_roofline(D, 3.0, 4.0, updates=2.0, pair_passes=2.0)
    assert r["algorithmic_bytes_per_solve"] == n * (4 * 83 + 2 * 40 + 2 * 8 + 56)
    assert r["updates_per_solve"] == 2.0 and r["pair_passes_per_solve"] == 2.0
