"""Host-side checks of bench.py (no GPU): the one-GPU guard and the reading
of the committed per-kernel HBM traffic."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_measured_traffic_reads_the_committed_profile():
    import bench

    t = bench.measured_traffic("k_expand_relax", 24)
    assert isinstance(t, int) and t > 0
    assert bench.measured_traffic("no_such_kernel", 24) is None
    assert bench.measured_traffic("k_expand_relax", 99) is None
