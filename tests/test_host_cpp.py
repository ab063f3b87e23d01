"""The C++ host side (casim_amd/host): the DyadSeries-shaped adapter over the
C ABI and host_check, which drives the library through it in the step order
of ParallelHybridSolver and compares the protocol path with the fused path
bit for bit on the GPU."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "casim_amd", "host")


def test_host_layer_builds_against_the_header():
    """The adapter header compiles with -Wall -Wextra and host_check links
    against libcasim_dyad.so (no GPU needed)."""
    if not os.path.exists(os.path.join(ROOT, "casim_amd", "libcasim_dyad.so")):
        pytest.skip("libcasim_dyad.so is not built")
    r = subprocess.run(["make", "-s", "-C", HOST], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "warning" not in r.stderr
    assert os.access(os.path.join(HOST, "host_check"), os.X_OK)


@pytest.mark.gpu
def test_host_check_protocol_equals_fused(gpu):
    exe = os.path.join(HOST, "host_check")
    if not os.path.exists(exe):  # a tree without build products: the host part builds in seconds
        subprocess.run(["make", "-s", "-C", HOST], check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and out["ok"], out
    assert out["channels_changed"] > 0
    assert out["dims_ok"] and out["names_ok"], out
