"""bench.py's contract that needs no GPU: one GPU only, and the HBM traffic figure of
the leaf kernel read from the committed rocprofv3 summary for the headline workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_leaf_traffic_comes_from_the_profile_summary():
    t = bench.leaf_traffic(8 << 30, 128, "murmur3", 0.02)
    # the kernel reads every input byte once plus the previous digests (N + 16 n)
    n = (8 << 30) // 128
    assert isinstance(t, int) and (8 << 30) + 16 * n <= t <= 1.05 * ((8 << 30) + 17 * n + 20 * n)
    assert bench.leaf_traffic(8 << 30, 128, "murmur3", 0.2) is None   # no summary for this workload
    assert bench.leaf_traffic(1 << 30, 128, "murmur3", 0.02) is None
