"""bench.py's command line on a machine without a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_figure_is_read_from_the_profile():
    sys.path.insert(0, ROOT)
    import bench
    for kernel in ("k_sweep", "k_sweep2"):
        t = bench.traffic_from_profiles(kernel)
        assert t is not None and 7.0e9 < t < 7.6e9
    assert bench.traffic_from_profiles("no_such_kernel") is None
