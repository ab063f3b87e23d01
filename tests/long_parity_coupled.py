"""The coupled workload's per-site shape over a long run, GPU against the oracle.

2D clefts of 201 x 201 nodes (h = 2 nm) with 64 RyRs, the centre one open at t = 0, each coupled at
its rim to a 51^3 cytosol box (h = 20 nm, one box step per 50 dyad steps) and to a depleting jSR;
the workload's own kinetics, 4 sites with their own MRG32k3a streams, dt = 2 ns, 10^5 steps
(0.2 ms, 2000 box steps).  After every block of 10^4 steps: channel states and RNG states
bit-exact; every dyad node, every box node and the jSR within 1e-9 relative (the bars of
tests/test_cytosol_box.py).  Too long for the GPU test suite (about three minutes, nearly all of it
the oracle, one host thread per site), so it runs on its own:

    python3 tests/long_parity_coupled.py [--sites 4] [--steps 100000] [--block 10000]
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
    ap.add_argument("--sites", type=int, default=4)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--block", type=int, default=10000)
    a = ap.parse_args()

    from casim_amd import workloads
    from oracle import oracle
    from tests.helpers import rel_err
    from tests.test_cytosol_box import FP64_TOL, compare_all, gpu_pair

    oracle.build(ref=False)
    sp = workloads.coupled_dyads(a.sites)
    box = workloads.coupled_box(sp)
    assert a.block % box.substeps == 0
    print("coupled long parity: %d sites, %d steps in blocks of %d" % (a.sites, a.steps, a.block), flush=True)
    ora, dev = gpu_pair(oracle, sp, box, workloads.coupled_pools())
    dt = sp.meta["dt_ms"]
    t0 = time.time()
    done, toggles = 0, 0
    prev = dev.GetChannelsStates().copy()
    while done < a.steps:
        k = min(a.block, a.steps - done)
        dev.Step(dt, k)
        ora.step(dt, k, threads=a.sites)
        done += k
        st = dev.GetChannelsStates()
        compare_all(ora, dev, sp, box, FP64_TOL)
        assert np.array_equal(dev.GetRngStates(), ora.dyad.rng_states()), ("RNG states differ", done)
        jsr = rel_err(dev.GetPools()[0], ora.jsr())
        assert jsr < FP64_TOL, ("jSR differs", done, jsr)
        worst = max(rel_err(dev.GetField(f), ora.dyad.field(f)) for f in range(sp.n_ions + sp.n_bound))
        worst_box = max(rel_err(dev.GetCytosolBox(f), ora.box_field(f)) for f in range(len(box.field_names(sp))))
        toggles += int((st != prev).sum())
        prev = st.copy()
        print("steps %d: states and streams equal, dyad rel err %.3g, box %.3g, jSR %.3g, open RyRs %d, "
              "jSR %.1f uM, %.0f s" % (done, worst, worst_box, jsr, int((st == 1).sum()),
                                       float(ora.jsr().min()), time.time() - t0), flush=True)
    print("ok: %d sites x %d steps, state changes seen at block ends %d" % (a.sites, a.steps, toggles), flush=True)


if __name__ == "__main__":
    main()
