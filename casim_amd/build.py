"""Builds libcasim_dyad.so in-tree with hipcc for gfx950.

The state-sweep instantiations live in sweep_inst.hip, compiled once per
(precision, dimension) pair; the objects compile in parallel and link into
one shared library.  Packed-FP32 instructions are disabled for all device
code (-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcasim_dyad.so")
OBJ = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CASIM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]

UNITS = [
    ("casim_dyad.hip", []),
    ("mlp.hip", []),
    ("qs.hip", []),
    ("cytosol_box.hip", []),
    ("sweep_inst.hip", ["-DCASIM_SWEEP_FP64=1", "-DCASIM_SWEEP_DIM=3"]),
    ("sweep_inst.hip", ["-DCASIM_SWEEP_FP64=1", "-DCASIM_SWEEP_DIM=2"]),
    ("sweep_inst.hip", ["-DCASIM_SWEEP_FP64=0", "-DCASIM_SWEEP_DIM=3"]),
    ("sweep_inst.hip", ["-DCASIM_SWEEP_FP64=0", "-DCASIM_SWEEP_DIM=2"]),
]


def _compile(unit, verbose):
    src, defs = unit
    tag = "_".join(d.split("=")[1] for d in defs) or "main"
    obj = os.path.join(OBJ, src.replace(".hip", "") + "_" + tag + ".o")
    cmd = [HIPCC] + FLAGS + defs + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on %s %s" % (src, " ".join(defs)))
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs or min(len(UNITS), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose), UNITS))
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libcasim_dyad.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
