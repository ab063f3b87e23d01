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


def test_committed_tuning_covers_the_benchmark_shapes():
    """The measured launch-shape picks in casim_amd/tuning.json are keyed as
    the autotuner looks them up for bench.py's FP64, FP32 and coupled legs."""
    from casim_amd import abi, autotune, workloads
    for spec in (workloads.batched_dyads(n_sites=4096, precision=abi.CASIM_FP64),
                 workloads.batched_dyads(n_sites=4096, precision=abi.CASIM_FP32),
                 workloads.coupled_dyads(512)):
        t = autotune.lookup(spec)
        assert t is not None, autotune.key(spec)
        assert t["threads"] % 64 == 0 and t["planes"] >= 0 and t.get("prefetch", 1) >= 1
