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


class _TimedDevice:
    """Stand-in for DyadHIP in the autotuner: each launch shape has a sweep
    launch time and a fused step time, and the step calls sleep for it."""

    def __init__(self, sweep_ms, step_ms, per_pass=2):
        self.sweep_ms, self.step_ms, self.per_pass = sweep_ms, step_ms, per_pass
        self.shape, self.calls = None, []

    def StepsPerPass(self):
        return self.per_pass

    def SetTuning(self, planes, threads, prefetch):
        self.shape = (planes, threads)

    def StepProfiled(self, dt, k):
        return self.sweep_ms.get(self.shape, 1.0) * k, 0.0

    def Step(self, dt, k):
        import time
        self.calls.append((self.shape, k))
        time.sleep(self.step_ms.get(self.shape, 1.0) * k * 1e-3)

    def Synchronize(self):
        pass


def test_autotuner_picks_by_the_fused_step_and_records_both_times():
    """The pick is the shape with the shortest fused step even where another
    has the shorter sweep launch; each candidate's graph is captured by one
    call before three timed ones."""
    from casim_amd import autotune, workloads
    spec = workloads.coupled_dyads(512)
    cands = autotune.candidates(spec)
    assert (67, 256, 1) in cands and (100, 256, 1) in cands  # thirds and halves of the 199 planes
    sweep = {(100, 256): 0.10, (67, 256): 0.15}
    step = {(100, 256): 0.30, (67, 256): 0.20}
    dev = _TimedDevice(sweep, step)
    best, res = autotune.tune(dev, spec, 1e-6, steps=5, warmup=1, save=False)
    assert (best["planes"], best["threads"]) == (67, 256)
    r100 = next(r for r in res if (r["planes"], r["threads"]) == (100, 256))
    assert r100["sweep_ms"] < best["sweep_ms"] and r100["step_ms"] > best["step_ms"]
    assert [k for s, k in dev.calls if s == (67, 256)] == [10, 10, 10, 10]
