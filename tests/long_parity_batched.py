"""The batched workload's sites over a long run, GPU against the oracle.

The first 16 sites of the heterogeneous batch (81 x 81 x 5 FP64 clefts, 14..97 RyRs in random
clusters on the 31 nm lattice, so that several sites need two 64-channel chunks), with the launch
shape the benchmark uses for the whole batch (one 448-thread workgroup per site, two steps per pass,
two half-batch streams), stochastic gating, dt = 10 ns, 10^5 steps (1 ms), in calls of 10^4 steps.
After every block: channel states and RNG states bit-exact, every node, mouth value, site current
and efflux within 1e-9 relative (the bars of tests/test_gpu_parity.py).  With --fp32: the FP32
state with its own launch shape (two 40-plane chunks per site, 256 threads), clamped channels (rates
zero: each site's first RyR stays open, the draws still happen, so FP32 rounding cannot flip one),
against the FP64 oracle: the FP32 state's drift per field after every block, every field within the
workload's 2 % accuracy bar.  Too long for the GPU test suite (about three minutes on 16 host threads, nearly all of it the
oracle), so it runs on its own:

    python3 tests/long_parity_batched.py [--fp32] [--sites 16] [--steps 100000] [--block 10000] [--threads 16]
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sites", type=int, default=16)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--block", type=int, default=10000)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--fp32", action="store_true")
    a = ap.parse_args()

    from casim_amd import abi, workloads
    from oracle import oracle
    from tests.helpers import rel_err
    from tests.test_gpu_parity import FP64_TOL, compare, setup_pair

    oracle.build(ref=False)
    precision, tol = (abi.CASIM_FP32, 2e-2) if a.fp32 else (abi.CASIM_FP64, FP64_TOL)
    sp = workloads.batched_dyads(n_sites=a.sites, clamped=a.fp32, precision=precision)
    print("batched long parity%s: %d sites (RyRs %s), %d steps in blocks of %d"
          % (" FP32" if a.fp32 else "", a.sites, [len(s) for s in sp.sites], a.steps, a.block), flush=True)
    ora, dev, _ = setup_pair(oracle, sp)
    dev.SetTuning(40, 256) if a.fp32 else dev.SetTuning(0, 448)
    assert dev.StepsPerPass() == 2
    dt = sp.meta["dt_ms"]
    t0 = time.time()
    done, worst, toggles = 0, 0.0, 0
    prev = dev.GetChannelsStates().copy()
    while done < a.steps:
        k = min(a.block, a.steps - done)
        dev.Step(dt, k)
        ora.step(dt, k, threads=a.threads)
        done += k
        st = dev.GetChannelsStates()
        assert np.array_equal(st, ora.channel_states()), ("channel states differ", done)
        assert np.array_equal(dev.GetRngStates(), ora.rng_states()), ("RNG states differ", done)
        toggles += int((st != prev).sum())
        prev = st.copy()
        if a.fp32:
            errs = [rel_err(dev.GetField(f), ora.field(f)) for f in range(sp.n_ions + sp.n_bound)]
            print("  FP32 drift per field: " + ", ".join("%s %.3g" % (n, e) for n, e in zip(sp.field_names(), errs)),
                  flush=True)
            assert max(errs) < tol, ("FP32 drift beyond the bar", done, errs)
            worst = max(worst, max(errs))
        else:
            worst = max(worst, compare(ora, dev, sp, tol))
        print("steps %d: states and streams equal, worst field rel err %.3g, open RyRs %d, %.0f s"
              % (done, worst, int((st == 1).sum()), time.time() - t0), flush=True)
    print("ok: %d sites x %d steps, worst rel err %.3g, state changes seen at block ends %d"
          % (a.sites, a.steps, worst, toggles), flush=True)


if __name__ == "__main__":
    main()
