"""Release statistics and throughput of the plain path (full reaction-diffusion
with per-step gating) against the surrogate-driven quasi-stationary path on
the same sites.

    python -m casim_amd.spark_compare --sites 1024 --t-max 2 --weights casim_amd/surrogate_weights.npz

Both runs start every site from rest with its init channel open, with a
constant jSR; statistics are those of IndChannelCREAnalyzer over [0, t_max]
(IndChannelCREAnalyzer.cpp:69-117): amplitude, time of the last event,
integral of N dt, and N at t_max.  The two populations are compared with
two-sample KS tests.  Progress and results are JSON lines.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, analysis, config, workloads

# SheepCannell parameters of the reference's own example and test inputs
# (test_input/Parameters.json, examples/CannellCPU/Parameters.json)
REFERENCE_RYR = {"kopen0": 457000.0, "ncaopen": 2.12, "kopenmax": 800.0, "kclose0": 245.0, "ncaclose": -0.27}


def depleting_pool(volume_nm3=1e7):
    """The depleting jSR of the coupled-dyad configuration: 1 mM start, 10 mM
    calsequestrin-like buffer with Kd 0.8 mM, refill tau 100 ms, NSR 1 mM;
    V_jSR as the reference's example input ("V jSR (nm3)": 1e7)."""
    return config.PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=volume_nm3, jsr_c0=[1000.0],
                           jsr_cq_sites=[1.0], jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])


def workload(n_sites, geometry_seed=2024, kinetics="reference", precision=abi.CASIM_FP64):
    spec = workloads.batched_dyads(n_sites=n_sites, geometry_seed=geometry_seed, precision=precision)
    if kinetics == "reference":
        spec.channel_types = [config.ChannelType("RyR", "SheepCannell", dict(REFERENCE_RYR),
                                                 [workloads.unitary_conductance()])]
    return spec


BIN_MS = 0.01  # IndChannelCREAnalyzer's time-histogram bin (IndChannelCREAnalyzer.cpp:6-8)


def run_plain(spec, t_max, jsr=1000.0, chunk=1000, log=None, pools=None):
    """Full RD + gating for t_max ms: release stats [S, 6], wall seconds;
    per-bin open counts in run_plain.counts, lowest jSR in run_plain.jsr_low
    (with a depleting pool)."""
    from .dyad import DyadHIP
    dt = workloads.DT_MS
    steps = int(round(t_max / dt))
    dev = DyadHIP(spec)
    dev.RecordOpenCounts(int(round(BIN_MS / dt)), int(round(t_max / BIN_MS)) + 1)
    ri, rb = spec.rim_arrays()

    def start():
        dev.InitOpeningAll()
        if pools is not None:
            dev.UpdatePools(ri, rb)
        else:
            dev.Update(ri, rb, spec.jsr_array([jsr]))

    if pools is not None:
        dev.AttachPools(pools)
    start()
    dev.Step(dt, min(chunk, steps))  # graph capture outside the timed region
    dev.Reset()
    start()
    dev.Synchronize()
    jlow = np.full(spec.n_sites, np.inf)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        dev.Step(dt, k)
        done += k
        if pools is not None:
            jlow = np.minimum(jlow, dev.GetPools()[0][0])
        if log and done % (20 * chunk) == 0:
            dev.Synchronize()
            log({"stage": "plain", "steps": done, "of": steps, "s": round(time.perf_counter() - t0, 1)})
    dev.Synchronize()
    wall = time.perf_counter() - t0
    stats = dev.GetReleaseStats()
    run_plain.counts = dev.GetOpenCounts()
    run_plain.jsr_low = jlow if pools is not None else None
    dev.close()
    return stats, wall


def run_qs(spec, net, t_max, jsr=1000.0, threshold=0.02, log=None, pools=None, lin=None):
    """Surrogate-driven quasi-stationary path to t_max: window stats [S, 4],
    event-record stats [S, 6], wall seconds, fallback solves."""
    from .surrogate import QuasiStationarySites
    qs = QuasiStationarySites(spec, np.full(spec.n_sites, jsr), t_max, net, spread_threshold=threshold,
                              pools=pools, lin=lin)
    qs.RecordOpenCounts(BIN_MS, int(round(t_max / BIN_MS)) + 1)
    t0 = time.perf_counter()
    qs.RunSimulation()
    wall = time.perf_counter() - t0
    win, full = qs.GetWindowStats(), qs.GetStats()
    fb = qs.fallback_solves
    info = {"rounds": qs.rounds, "fallback_s": qs.fallback_s}
    run_qs.counts = qs.GetOpenCounts()
    run_qs.jsr_low = qs.GetJsr()[1] if pools is not None else None
    qs.close()
    if log:
        log({"stage": "qs", "s": round(wall, 2), "events": float(full[:, 5].sum()), "fallback_sites": fb, **info})
    run_qs.last = info
    return win, full, wall, fb


def on_grid(t, dt):
    """Times and time integrals at the fixed-step path's resolution: the plain
    path sees events only at multiples of dt and sums N dt over steps, so a
    continuous-time value is rounded to the nearest multiple of dt (without
    this, the point mass of sites open through the whole window, exactly
    t_max in one path and a 10^5-term FP sum in the other, separates the
    samples)."""
    return np.round(np.asarray(t) / dt) * dt


def compare(plain, qs_win, n_channels, t_max, alpha=0.01, dt=workloads.DT_MS):
    """KS tests of amplitude, last event time, integral of N dt, N at t_max
    and open probability (integral / (n_RyR t_max))."""
    n_ch = np.asarray(n_channels, dtype=np.float64)
    cols = {"amplitude": (plain[:, 0], qs_win[:, 0]),
            "last_event_ms": (on_grid(plain[:, 2], dt), on_grid(qs_win[:, 1], dt)),
            "open_integral": (on_grid(plain[:, 3], dt), on_grid(qs_win[:, 2], dt)),
            "n_open_end": (plain[:, 5], qs_win[:, 3]),
            "open_probability": (on_grid(plain[:, 3], dt) / (n_ch * t_max), on_grid(qs_win[:, 2], dt) / (n_ch * t_max))}
    out = {}
    for name, (a, b) in cols.items():
        d, p = analysis.ks_2samp(a, b)
        out[name] = {"D": d, "p": p, "pass": p > alpha, "mean_plain": float(np.mean(a)), "mean_qs": float(np.mean(b))}
    pa = analysis.Population(np.column_stack([plain[:, 0], plain[:, 1], plain[:, 2], plain[:, 3], plain[:, 4],
                                              plain[:, 5]]), n_channels, t_max)
    out["classes_plain"] = pa.classes()
    qsfull = np.zeros((len(qs_win), 6))
    qsfull[:, 0] = qs_win[:, 0]
    out["classes_qs"] = analysis.Population(qsfull, n_channels, t_max).classes()
    out["all_pass"] = all(v["pass"] for k, v in out.items() if isinstance(v, dict) and "pass" in v)
    return out


def held_open(plain, qs_win):
    """Sites whose init RyR stayed open over the whole window with nothing
    else happening (amplitude 1, no event, one RyR open at t_max), from each
    path's record: (plain mask, surrogate mask)."""
    a = (plain[:, 0] == 1) & (plain[:, 2] == 0) & (plain[:, 5] == 1)
    b = (qs_win[:, 0] == 1) & (qs_win[:, 1] == 0) & (qs_win[:, 3] == 1)
    return a, b


def compare_jsr_lowest(plain, qs_win, low_plain, low_qs, jsr0=1000.0, alpha=0.01):
    """The lowest jSR per site, plain vs surrogate.  Sites held open over the
    whole window (held_open) share one deterministic depletion path, so each
    path puts them on a single value: an atom of the distribution.  Two-sample
    KS is not meaningful at an atom (any difference in its location, however
    small, shows as a jump of the atom's weight), so the atom is compared on
    its own: its weight (two-proportion test) and the relative difference of
    its depletion; KS covers the other sites."""
    from scipy import stats
    a, b = held_open(plain, qs_win)
    out = {"held_open_plain": int(a.sum()), "held_open_qs": int(b.sum())}
    if a.any() and b.any():
        drop_p = jsr0 - float(np.mean(low_plain[a]))
        drop_q = jsr0 - float(np.mean(low_qs[b]))
        out["held_open_drop_plain_uM"] = drop_p
        out["held_open_drop_qs_uM"] = drop_q
        out["held_open_drop_rel_diff"] = abs(drop_p - drop_q) / max(abs(drop_p), 1e-300)
        out["held_open_spread_uM"] = float(max(np.ptp(low_plain[a]), np.ptp(low_qs[b])))
    n1, n2 = len(a), len(b)
    tab = [[int(a.sum()), n1 - int(a.sum())], [int(b.sum()), n2 - int(b.sum())]]
    out["held_open_weight_p"] = float(stats.chi2_contingency(tab)[1]) if min(map(min, tab)) > 0 else 1.0
    d, pv = analysis.ks_2samp(low_plain[~a], low_qs[~b])
    out["others"] = {"D": d, "p": pv, "pass": pv > alpha, "mean_plain": float(np.mean(low_plain[~a])),
                     "mean_qs": float(np.mean(low_qs[~b]))}
    d, pv = analysis.ks_2samp(low_plain, low_qs)
    out["all_sites_ks"] = {"D": d, "p": pv}
    return out


def compare_histograms(counts_plain, counts_qs, amp_plain, amp_qs, n_channels, nadir=4):
    """Mean open count against time (the analyzer's time histograms reduced
    to their first moment), all sites and per class."""
    out = {}
    hp = analysis.time_histograms(counts_plain, amp_plain, n_channels, nadir)
    hq = analysis.time_histograms(counts_qs, amp_qs, n_channels, nadir)
    for name in ("all", "quarks", "blips", "sparks"):
        def mean_curve(h):
            n = h.sum(axis=1)
            return (h * np.arange(h.shape[1])).sum(axis=1) / np.maximum(n, 1)
        a, b = mean_curve(hp[name]), mean_curve(hq[name])
        out[name] = {"max_abs_diff": float(np.max(np.abs(a - b))), "plain_end": float(a[-1]), "qs_end": float(b[-1]),
                     "sites_plain": int(hp[name][0].sum()), "sites_qs": int(hq[name][0].sum())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sites", type=int, default=1024)
    ap.add_argument("--t-max", type=float, default=2.0)
    ap.add_argument("--kinetics", default="reference", choices=["reference", "bench"])
    ap.add_argument("--weights", default="casim_amd/surrogate_weights.npz")
    ap.add_argument("--threshold", type=float, default=0.02)
    ap.add_argument("--jsr", type=float, default=1000.0)
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--deplete", action="store_true", help="depleting jSR pool in both paths")
    a = ap.parse_args(argv)

    def log(d):
        print(json.dumps(d), flush=True)

    from . import surrogate
    spec = workload(a.sites, kinetics=a.kinetics)
    n_ch = np.array([len(s) for s in spec.sites])
    net, lin = surrogate.load_model(a.weights)
    pools = depleting_pool() if a.deplete else None
    win, full, wq, fb = run_qs(spec, net, a.t_max, a.jsr, a.threshold, log, pools, lin)
    res = {"stage": "result", "sites": a.sites, "t_max_ms": a.t_max, "kinetics": a.kinetics, "deplete": a.deplete,
           "qs_wall_s": wq, "qs_dyad_ms_per_s": a.sites * a.t_max / wq, "fallback_sites": fb,
           "events_per_site": float(full[:, 5].mean())}
    if not a.skip_plain:
        plain, wp = run_plain(spec, a.t_max, a.jsr, log=log, pools=pools)
        if pools is not None:
            q = [0.01, 0.1, 0.5, 0.9, 0.99]
            res["jsr_lowest"] = {"mean_plain": float(np.mean(run_plain.jsr_low)),
                                 "mean_qs": float(np.mean(run_qs.jsr_low)),
                                 "quantiles": q,
                                 "plain_q": [float(v) for v in np.quantile(run_plain.jsr_low, q)],
                                 "qs_q": [float(v) for v in np.quantile(run_qs.jsr_low, q)],
                                 **compare_jsr_lowest(plain, win, run_plain.jsr_low, run_qs.jsr_low)}
        res.update({"plain_wall_s": wp, "plain_dyad_ms_per_s": a.sites * a.t_max / wp,
                    "speedup": wp / wq, "ks": compare(plain, win, n_ch, a.t_max),
                    "mean_open_vs_time": compare_histograms(run_plain.counts, run_qs.counts, plain[:, 0], win[:, 0],
                                                            int(n_ch.max()))})
    log(res)
    return 0


if __name__ == "__main__":
    sys.exit(main())
