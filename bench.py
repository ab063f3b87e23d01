"""Batched dyad reaction-diffusion benchmark (one GPU).

A "step" advances every site of the batch by one dt: channel kernel (mouth
snapshot, RyR currents, stochastic gating) + the fused state sweep, replayed
from a captured HIP graph.  Inputs are resident in HBM before timing starts.

Workload: 4096 independent release sites, 81 x 81 x 5 nodes each (320 x 320 x
16 nm at h = 4 nm), free Ca2+ + ATP-like + calmodulin-like mobile buffers +
a membrane buffer on the two z-faces, 8..100 RyRs per site in random
clusters, SheepCannell-form gating, dt = 10 ns, FP64 state (FP32 reported
beside it).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated dyad-ms/s on 1 GPU (plain vs surrogate) at ≤2% Ca err; HBM GB/s"
HBM_PEAK_GBS = 8000.0
COPY_GBS = 5314.0  # a plain 1 GiB device-to-device copy on MI355X (read + write), measured with rocprofv3 (kernel trace)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu_leg(workloads, DyadHIP, precision, n_sites, steps, warmup, chunk, retune=False, tuning=None,
                steps_per_pass=2, split=True):
    from casim_amd import autotune
    sp = workloads.batched_dyads(n_sites=n_sites, precision=precision)
    dev = DyadHIP(sp)
    dev.SetSplit(split)
    dev.SetStepsPerPass(steps_per_pass)
    dev.InitOpeningAll()
    ri, rb = sp.rim_arrays()
    dev.Update(ri, rb, sp.jsr_array([sp.meta["jsr_uM"]]))
    dt = sp.meta["dt_ms"]
    tuned = None if retune else autotune.apply(dev, sp)
    if tuning:
        dev.SetTuning(tuning[0], tuning[1], 1)
        tuned = {"planes": tuning[0], "threads": tuning[1], "sweep_ms": None}
    if tuned is None:
        tuned, _ = autotune.tune(dev, sp, dt, log=log)
        # the tuning steps advanced the state: start the timed run from a reset
        dev.Reset()
        dev.InitOpeningAll()
        dev.Update(ri, rb, sp.jsr_array([sp.meta["jsr_uM"]]))
    log("tuning:", {k: tuned[k] for k in ("planes", "threads", "sweep_ms")})
    chunk = min(chunk, steps)
    rem = steps % chunk
    # warmup also captures the graphs used in the timed region
    done = 0
    while done < max(warmup, 1):
        dev.Step(dt, chunk)
        done += chunk
    if rem:
        dev.Step(dt, rem)
    dev.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        dev.Step(dt, chunk)
    if rem:
        dev.Step(dt, rem)
    dev.Synchronize()
    t = time.perf_counter() - t0
    # per-kernel durations, live, with HIP events on the handle's stream
    ms_sweep, ms_chan = dev.StepProfiled(dt, steps)
    per_pass = dev.StepsPerPass()
    launches = steps // per_pass + steps % per_pass
    ca = dev.GetField(0)
    ok = bool(np.isfinite(ca).all())
    bps = dev.BytesPerSiteStep()
    open_frac = float((dev.GetChannelsStates() == 1).sum()) / sum(len(s) for s in sp.sites)
    res = {
        "sites": n_sites, "grid": "%dx%dx%d" % (sp.nx, sp.ny, sp.nz), "dt_ms": dt,
        "wall_s": t, "ms_per_step": 1e3 * t / steps,
        "dyad_ms_per_s": n_sites * steps * dt / t,
        "bytes_per_site_step": bps,
        "sweep_kernel_ms": ms_sweep / steps, "channel_kernel_ms": ms_chan / steps,
        "steps_per_launch": per_pass, "sweep_launch_ms": ms_sweep / launches,
        # one read and one write of the state per launch, whatever the steps it advances
        "sweep_gbs": n_sites * bps / (ms_sweep / launches * 1e-3) / 1e9,
        "finite": ok, "open_fraction": open_frac,
        "workload": sp.meta["workload"],
        "tuning": {"planes": tuned["planes"], "threads": tuned["threads"]},
    }
    dev.close()
    return res


def cpu_baseline(workloads, n_sites=32, steps=5000, single_sites=4, single_steps=250):
    """The oracle (C restatement of DyadRDCPU, OpenMP over sites as
    ParallelHybridSolver::RunSimulationsCPU does, ParallelHybridSolver.cpp:
    408-437, with the serial Update after the parallel loop) on the first
    n_sites sites of the same batch, on every host thread; then the same
    sites on one thread for the single-core cost per site-step."""
    from oracle import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    dt = workloads.DT_MS

    def sample(n):
        sp = workloads.batched_dyads(n_sites=4096)
        sp.sites = sp.sites[:n]
        sp.rng_states = sp.rng_states[:n]
        sp.init_channel = sp.init_channel[:n]
        o = O.Oracle(sp)
        o.init_opening()
        ri, rb = sp.rim_arrays()
        o.update(ri, rb, sp.jsr_array([1000.0]))
        return o

    o = sample(n_sites)
    o.step(dt, 2, threads)
    t0 = time.perf_counter()
    serial = o.step_timed(dt, steps, threads)
    t = time.perf_counter() - t0
    o1 = sample(single_sites)
    o1.step(dt, 2, 1)
    t1 = time.perf_counter()
    o1.step(dt, single_steps, 1)
    t1 = time.perf_counter() - t1
    return {"value": n_sites * steps * dt / t, "unit": "dyad-ms/s", "cores": threads, "kind": "port",
            "sample": "first %d sites of the same 4096-site batch (81x81x5 FP64), %d steps of RunRD + RunMC "
                      "per site in parallel then the serial Update, %.1f s wall" % (n_sites, steps, t),
            "us_per_site_step": 1e6 * t / (n_sites * steps),
            "single_core_us_per_site_step": 1e6 * t1 / (single_sites * single_steps),
            "single_core_sample": "first %d sites, %d steps on one thread (%.1f s)" % (single_sites, single_steps, t1),
            "serial_update_fraction": serial / t}


def surrogate_leg(workloads, DyadHIP, n_sites, t_max, plain_value, heldout_sites=256, ks_sites=1024, ks_tmax=1.0):
    """The surrogate-driven quasi-stationary path (the event-driven release
    site with the learned mouth-[Ca2+] ensemble) on the same sites and
    kinetics as the plain leg: every site from rest, its init channel open,
    to t_max, jSR constant; throughput in the plain leg's unit.  Measured in
    the same run: the ensemble's mouth-[Ca2+] error on held-out geometries
    that no training batch used, solved by the full FP64 solver; and, at a
    t_max the plain path also runs (ks_tmax, depleting jSR, ks_sites sites),
    both paths' throughput and the KS comparison of their release statistics."""
    w = os.path.join(ROOT, "casim_amd", "surrogate_weights.npz")
    if not os.path.exists(w):
        return None
    from casim_amd import autotune, mlp, spark_compare, surrogate, train_surrogate
    net, lin = surrogate.load_model(w)
    spec = workloads.batched_dyads(n_sites=n_sites)
    spark_compare.run_qs(spec, net, 0.05, lin=lin)  # warm-up: workspaces and kernels loaded
    win, full, wall, fb = spark_compare.run_qs(spec, net, t_max, lin=lin)
    meta = mlp.MLP.meta(w)
    value = n_sites * t_max / wall
    out = {"value": value, "unit": "dyad-ms/s", "speedup_vs_plain": value / plain_value, "t_max_ms": t_max,
           "wall_s": wall, "events_per_site": float(full[:, 5].mean()), "fallback_sites": int(fb),
           "rounds": spark_compare.run_qs.last["rounds"], "fallback_s": spark_compare.run_qs.last["fallback_s"],
           "target": meta.get("target")}
    # held-out error, in this run: fresh geometries, open sets and jSR values
    t = autotune.lookup(workloads.batched_dyads(n_sites=heldout_sites, clamped=True))
    tuning = (t["planes"], t["threads"], t.get("prefetch", 1)) if t else None
    t0 = time.perf_counter()
    xh, yh, _ = surrogate.generate(DyadHIP, n_sites=heldout_sites, seed=50_000, tuning=tuning, lin=lin)
    held = train_surrogate.errors(net, xh, yh, lin=lin)
    out["mouth_ca_rel_err_heldout"] = {"sites": heldout_sites, "seconds": round(time.perf_counter() - t0, 1),
                                       **{k: held.get(k) for k in ("rows", "rel_mean", "rel_p50", "rel_p95",
                                                                   "rel_mean_above_5uM", "flagged_frac")}}
    # both paths at one t_max, depleting jSR: throughput and KS of the release statistics
    kspec = spark_compare.workload(ks_sites, kinetics="bench")
    n_ch = np.array([len(st) for st in kspec.sites])
    pools = spark_compare.depleting_pool()
    kwin, _, kwq, kfb = spark_compare.run_qs(kspec, net, ks_tmax, pools=pools, lin=lin)
    jq = spark_compare.run_qs.jsr_low
    kplain, kwp = spark_compare.run_plain(kspec, ks_tmax, pools=pools)
    jp = spark_compare.run_plain.jsr_low
    ks = spark_compare.compare(kplain, kwin, n_ch, ks_tmax)
    jr = spark_compare.compare_jsr_lowest(kplain, kwin, jp, jq)
    out["same_tmax"] = {
        "sites": ks_sites, "t_max_ms": ks_tmax, "jsr": "depleting (SRPool)",
        "plain_dyad_ms_per_s": ks_sites * ks_tmax / kwp, "surrogate_dyad_ms_per_s": ks_sites * ks_tmax / kwq,
        "speedup": kwp / kwq, "fallback_sites": int(kfb),
        "ks_p": {"amplitude": ks["amplitude"]["p"], "last_event": ks["last_event_ms"]["p"],
                 "open_integral": ks["open_integral"]["p"], "n_open_end": ks["n_open_end"]["p"],
                 "open_probability": ks["open_probability"]["p"],
                 "jsr_lowest_other_sites": jr["others"]["p"]},
        "jsr_lowest_held_open": {"weight_p": jr["held_open_weight_p"], "drop_rel_diff": jr.get("held_open_drop_rel_diff")},
        "classes_plain": ks["classes_plain"], "classes_qs": ks["classes_qs"]}
    return out


def coupled_leg(workloads, DyadHIP, n_sites, steps, warmup, chunk=100, tuning=None):
    """The coupled configuration: 2D clefts of 201 x 201 nodes at h = 2 nm,
    each inside a 51^3 cytosol box at h = 20 nm (Ca2+, A, B, troponin; box
    step every 50 dyad steps), a depleting jSR per site, 64 RyRs, dt = 2 ns,
    FP64.  One step = channels + dyad sweep (+ box coupling)."""
    from casim_amd import autotune
    sp = workloads.coupled_dyads(n_sites)
    box = workloads.coupled_box(sp)
    dt = sp.meta["dt_ms"]
    tuned = autotune.lookup(sp)
    if tuning:
        tuned = {"planes": tuning[0], "threads": tuning[1], "prefetch": tuning[2] if len(tuning) > 2 else 1,
                 "sweep_ms": None}
    if tuned is None:
        # the sweep's launch shape, tuned on the same sites and pool without the box
        probe = DyadHIP(sp)
        probe.AttachPools(workloads.coupled_pools())
        probe.InitOpeningAll()
        probe.UpdatePools(*sp.rim_arrays())
        tuned, _ = autotune.tune(probe, sp, dt, log=log)
        probe.close()
    log("coupled tuning:", {k: tuned[k] for k in ("planes", "threads", "prefetch", "sweep_ms")})
    dev = DyadHIP(sp)
    dev.AttachPools(workloads.coupled_pools())
    dev.AttachCytosolBox(box)
    dev.SetTuning(tuned["planes"], tuned["threads"], tuned.get("prefetch", 1))
    dev.InitOpeningAll()
    dev.UpdatePools()
    chunk = max(box.substeps, chunk // box.substeps * box.substeps)
    steps = max(chunk, steps // chunk * chunk)
    done = 0
    while done < max(warmup, 1):
        dev.Step(dt, chunk)
        done += chunk
    dev.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        dev.Step(dt, chunk)
    dev.Synchronize()
    t = time.perf_counter() - t0
    ms_sweep, ms_chan = dev.StepProfiled(dt, chunk)
    jsr = dev.GetPools()[0]
    ca = dev.GetCytosolBox(0)
    res = {"workload": sp.meta["workload"], "sites": n_sites, "dyad_grid": "%dx%d" % (sp.nx, sp.ny),
           "box_grid": "%dx%dx%d" % (box.nx, box.ny, box.nz), "box_substeps": box.substeps, "dt_ns": dt * 1e6,
           "steps": steps, "steps_per_pass": dev.StepsPerPass(),
           "tuning": {k: tuned[k] for k in ("planes", "threads", "prefetch")},
           "wall_s": t, "ms_per_step": 1e3 * t / steps,
           "value": n_sites * steps * dt / t, "unit": "dyad-ms/s",
           "sweep_kernel_ms": ms_sweep / chunk, "channel_kernel_ms": ms_chan / chunk,
           # one read and one write of the state per sweep launch (a launch advances steps_per_pass steps)
           "sweep_launch_ms": ms_sweep / (chunk / dev.StepsPerPass()),
           "sweep_gbs": n_sites * dev.BytesPerSiteStep() / (ms_sweep / (chunk / dev.StepsPerPass()) * 1e-3) / 1e9,
           "sweep_frac_of_hbm_peak": n_sites * dev.BytesPerSiteStep() / (ms_sweep / (chunk / dev.StepsPerPass()) * 1e-3)
           / 1e9 / HBM_PEAK_GBS,
           "jsr_min_uM": float(jsr.min()), "box_ca_max_uM": float(ca.max()),
           "finite": bool(np.isfinite(ca).all() and np.isfinite(dev.GetField(0)).all())}
    dev.close()
    return res


def traffic_from_profiles(kernel):
    """HBM bytes per sweep-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, KiB, scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", "%s_batched4096_81x81x5.md" % kernel)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--sites", type=int, default=4096)
    ap.add_argument("--chunk", type=int, default=100, help="steps per captured HIP graph")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--retune", action="store_true", help="ignore the tuning cache")
    ap.add_argument("--tuning", default=None, help="planes,threads of the state sweep (skips the autotuner)")
    ap.add_argument("--steps-per-pass", type=int, default=2, choices=[1, 2],
                    help="2: pairs of steps per state pass where possible (default); 1: one pass per step")
    ap.add_argument("--no-split", action="store_true",
                    help="one stream for the whole batch (default: two half-batch streams, channel calls of one "
                         "half beside the other half's pass)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--qs-tmax", type=float, default=20.0, help="simulated ms of the surrogate leg")
    ap.add_argument("--no-coupled", action="store_true")
    ap.add_argument("--coupled-sites", type=int, default=512)
    ap.add_argument("--coupled-steps", type=int, default=1000)
    ap.add_argument("--coupled-only", action="store_true",
                    help="run only the coupled dyad + cytosol-box leg and print its JSON (for profiling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: sites are independent and this program has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from casim_amd import abi, workloads
    from casim_amd.dyad import DyadHIP
    abi.load()
    tuning = [int(v) for v in a.tuning.split(",")] if a.tuning else None
    if a.coupled_only:
        print(json.dumps(coupled_leg(workloads, DyadHIP, a.coupled_sites, a.coupled_steps, a.warmup,
                                     tuning=tuning)), flush=True)
        return
    log("fp64 leg: %d sites, %d steps" % (a.sites, a.steps))
    r64 = run_gpu_leg(workloads, DyadHIP, abi.CASIM_FP64, a.sites, a.steps, a.warmup, a.chunk, a.retune, tuning,
                      a.steps_per_pass, not a.no_split)
    log("fp64:", json.dumps(r64))
    r32 = None
    if not a.no_fp32:
        log("fp32 leg")
        r32 = run_gpu_leg(workloads, DyadHIP, abi.CASIM_FP32, a.sites, a.steps, a.warmup, a.chunk, a.retune,
                          tuning, a.steps_per_pass, not a.no_split)
        log("fp32:", json.dumps(r32))
    sur = None
    if not a.no_surrogate:
        log("surrogate leg")
        sur = surrogate_leg(workloads, DyadHIP, a.sites, a.qs_tmax, r64["dyad_ms_per_s"])
        log("surrogate:", json.dumps(sur))
    cpl = None
    if not a.no_coupled:
        log("coupled leg: %d sites" % a.coupled_sites)
        cpl = coupled_leg(workloads, DyadHIP, a.coupled_sites, a.coupled_steps, 100)
        log("coupled:", json.dumps(cpl))
    cpu = None
    if not a.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(workloads)
        log("cpu:", json.dumps(cpu))
    # dominant kernel: k_sweep2 (two steps per launch) or k_sweep (one).  Its
    # algorithmic bytes per launch are one read and one write of every
    # site's state (sites x bytes per site-step): a two-step launch moves the
    # state once, so the steps it advances do not multiply the bytes.
    kernel = "k_sweep2" if r64["steps_per_launch"] == 2 else "k_sweep"
    launch_bytes = a.sites * r64["bytes_per_site_step"]
    achieved = launch_bytes / (r64["sweep_launch_ms"] * 1e-3) / 1e9
    traffic = traffic_from_profiles(kernel)
    out = {
        "metric": METRIC,
        "value": r64["dyad_ms_per_s"],
        "unit": "dyad-ms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": r64["workload"], "sites": a.sites, "grid": r64["grid"], "dt_ns": 10,
                   "path": "plain batched RD, graph of %d steps%s" % (
                       min(a.chunk, a.steps), "" if a.no_split else ", two half-batch streams"),
                   "tuning": r64["tuning"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "steps_per_launch": r64["steps_per_launch"],
                     "launch_ms": r64["sweep_launch_ms"], "algorithmic_bytes_per_launch": launch_bytes,
                     # what HBM actually moves per launch (traffic / launch time), next to the copy rate
                     "physical_gbs": None if traffic is None else traffic / (r64["sweep_launch_ms"] * 1e-3) / 1e9,
                     "copy_gbs": COPY_GBS,
                     # per-step algorithmic rate: the bytes a one-step-per-pass sweep would move
                     # for the same steps, over the same time (not bounded by the HBM peak)
                     "effective_gbs": r64["steps_per_launch"] * achieved},
        "cpu_baseline": cpu,
        # one-step-per-pass bytes of a whole step (channels + sweep) over its time
        "effective_gbs_whole_step": a.sites * r64["bytes_per_site_step"] / (r64["ms_per_step"] * 1e-3) / 1e9,
        "fp32": None if r32 is None else {
            "value": r32["dyad_ms_per_s"], "ms_per_step": r32["ms_per_step"], "sweep_gbs": r32["sweep_gbs"],
            "sweep_kernel_ms": r32["sweep_kernel_ms"]},
        "kernels": {"sweep_ms": r64["sweep_kernel_ms"], "channels_ms": r64["channel_kernel_ms"]},
        "surrogate": sur,
        "coupled": cpl,
        "finite": r64["finite"] and (r32 is None or r32["finite"]) and (cpl is None or cpl["finite"]),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
