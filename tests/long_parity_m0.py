"""The single-dyad workload at its stated length, GPU against the oracle.

61 x 61 x 5 FP64 clefts with 25 RyRs each (the centre one open at t = 0), 32 sites with their own
MRG32k3a streams, stochastic gating, dt = 10 ns, 10^5 steps (1 ms).  After every block of 10^4
steps: channel states and RNG states bit-exact, every node, mouth value, site current and efflux
within 1e-9 relative (the bars of tests/test_gpu_parity.py).  Too long for the GPU test suite (about
six minutes on 16 host threads, nearly all of it the oracle), so it runs on its own:

    python3 tests/long_parity_m0.py [--sites 32] [--steps 100000] [--block 10000] [--threads 16]
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
    ap.add_argument("--sites", type=int, default=32)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--block", type=int, default=10000)
    ap.add_argument("--threads", type=int, default=16)
    a = ap.parse_args()

    from casim_amd import workloads
    from casim_amd.dyad import DyadHIP
    from oracle import oracle
    from tests.test_gpu_parity import FP64_TOL, compare, setup_pair

    oracle.build(ref=False)
    sp = workloads.single_dyad(a.sites, clamped=False)
    ora, dev, _ = setup_pair(oracle, sp)
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
        worst = max(worst, compare(ora, dev, sp, FP64_TOL))
        print("steps %d: states and streams equal, worst field rel err %.3g, open RyRs %d, %.0f s"
              % (done, worst, int((st == 1).sum()), time.time() - t0), flush=True)
    print("ok: %d sites x %d steps, worst rel err %.3g, state changes seen at block ends %d"
          % (a.sites, a.steps, worst, toggles), flush=True)


if __name__ == "__main__":
    main()
