"""The surrogate-in-the-loop path against the plain path on the same sites:
1024 sites of the batched workload (8..100 RyRs, kopen = 0.2/ms (c/15 uM)^2,
kclose = 0.5/ms), a depleting jSR (1 mM, calsequestrin 10 mM with Kd 0.8 mM,
refill tau 100 ms, SRPool.cpp:70-89) in both paths, 1 ms each.  The release
statistics of IndChannelCREAnalyzer (IndChannelCREAnalyzer.cpp:53-138) -
amplitude, time of the last event, integral of N dt, open probability
(integral / (n_RyR t_max)) and N at t_max - must be KS-indistinguishable at
p > 0.01, and so must the lowest jSR per site once the sites held open over
the whole window, which share one value, are compared on their own.  Both
runs are deterministic (fixed streams and geometry), so the outcome does not
vary between runs."""
import os

import numpy as np
import pytest

from casim_amd import analysis, spark_compare, surrogate

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_surrogate_path_ks_indistinguishable_from_plain(gpu):
    n_sites, t_max = 1024, 1.0
    spec = spark_compare.workload(n_sites, kinetics="bench")
    n_ch = np.array([len(s) for s in spec.sites])
    net, lin = surrogate.load_model(os.path.join(ROOT, "casim_amd", "surrogate_weights.npz"))
    pools = spark_compare.depleting_pool()
    win, full, wq, fb = spark_compare.run_qs(spec, net, t_max, pools=pools, lin=lin)
    jq = spark_compare.run_qs.jsr_low
    print("surrogate path: %.2f s, %d fallback sites, %.2f events per site" % (wq, fb, full[:, 5].mean()), flush=True)
    plain, wp = spark_compare.run_plain(spec, t_max, pools=pools)
    jp = spark_compare.run_plain.jsr_low
    print("plain path: %.1f s" % wp, flush=True)
    res = spark_compare.compare(plain, win, n_ch, t_max)
    po_plain = spark_compare.on_grid(plain[:, 3], spec.meta["dt_ms"]) / (n_ch * t_max)
    po_qs = spark_compare.on_grid(win[:, 2], spec.meta["dt_ms"]) / (n_ch * t_max)
    d, p = analysis.ks_2samp(po_plain, po_qs)
    assert abs(res["open_probability"]["p"] - p) < 1e-12
    print({k: (v["p"] if isinstance(v, dict) and "p" in v else v) for k, v in res.items()}, flush=True)
    for name in ("amplitude", "last_event_ms", "open_integral", "n_open_end", "open_probability"):
        assert res[name]["p"] > 0.01, (name, res[name])
    # the lowest jSR per site: sites held open over the whole window share one
    # depletion path (an atom of the distribution); compare its weight and
    # its depletion, and KS the other sites
    jr = spark_compare.compare_jsr_lowest(plain, win, jp, jq)
    print(jr, flush=True)
    assert jp.min() < 999.0 and jq.min() < 999.0
    assert jr["held_open_plain"] > 100 and jr["held_open_qs"] > 100
    assert jr["held_open_spread_uM"] < 1e-9 * 1000.0
    assert jr["held_open_drop_rel_diff"] < 1e-3
    assert jr["held_open_weight_p"] > 0.01
    assert jr["others"]["p"] > 0.01, jr["others"]
    assert wq < wp
