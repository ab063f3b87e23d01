"""Train the mouth-[Ca2+] surrogate ensemble on the full solver and measure
its error on held-out geometries.

    python -m casim_amd.train_surrogate --batches 8 --sites 4096 --out casim_amd/surrogate_weights.npz

Each batch is `sites` random dyads (RyR clusters of 8..100 on the 31 nm
lattice, random open sets, jSR 200..1200 uM) relaxed for QS_STEPS clamped
steps of the FP64 solver; the held-out batch uses geometry and open-set
seeds no training batch uses.  Progress and results are JSON lines."""
import argparse
import json
import sys
import time

import numpy as np

from . import mlp, surrogate


def dataset(DyadHIP, batches, sites, seed0=0, tuning=None, log=None, lin=None):
    xs, ys = [], []
    for b in range(batches):
        t0 = time.time()
        x, y, _ = surrogate.generate(DyadHIP, n_sites=sites, seed=seed0 + b, tuning=tuning, lin=lin)
        xs.append(x)
        ys.append(y)
        if log:
            log({"stage": "data", "batch": b, "rows": int(len(y)), "s": round(time.time() - t0, 2)})
    return np.concatenate(xs), np.concatenate(ys)


def errors(net, x, y, threshold=0.02, lin=None):
    """Relative error of the ensemble mean in concentration units, overall
    and on the rows whose spread stays under `threshold`; y is ln [Ca2+]."""
    mean, spread = surrogate.ensemble_mean(net, x, lin)
    true = np.exp(y.astype(np.float64))
    rel = np.abs(mean - true) / true
    ok = spread <= threshold
    out = {"rows": int(len(y)), "rel_mean": float(rel.mean()), "rel_p50": float(np.median(rel)),
           "rel_p95": float(np.quantile(rel, 0.95)), "rel_p99": float(np.quantile(rel, 0.99)),
           "rel_max": float(rel.max()), "flagged_frac": float(1.0 - ok.mean())}
    if ok.any():
        out["rel_mean_unflagged"] = float(rel[ok].mean())
        out["rel_p99_unflagged"] = float(np.quantile(rel[ok], 0.99))
    hi = true > 5.0  # mouths near open channels, where gating rates are not negligible
    if hi.any():
        out["rows_above_5uM"] = int(hi.sum())
        out["rel_mean_above_5uM"] = float(rel[hi].mean())
        out["rel_p95_above_5uM"] = float(np.quantile(rel[hi], 0.95))
    return out


def fit(x, y, width=256, depth=4, members=5, epochs=60, batch=1024, lr=1e-3, seed=1, log=None, schedule=(0.6, 0.85)):
    net = mlp.MLP(x.shape[1], width=width, depth=depth, members=members, seed=seed)
    net.fit_normalization(x, y)
    net.set_data(x, y)
    done = 0
    # step decay: lr, lr/3 after schedule[0] of the epochs, lr/10 after schedule[1];
    # two epochs per call so that progress is reported as it goes
    for frac_end, scale in ((schedule[0], 1.0), (schedule[1], 1.0 / 3.0), (1.0, 0.1)):
        end = max(done + 1, int(round(epochs * frac_end)))
        while done < end:
            n = min(2, end - done)
            t0 = time.time()
            loss = net.train(n, batch=batch, lr=lr * scale, seed=seed + 1000 * done)
            done += n
            if log:
                log({"stage": "train", "epochs": done, "lr": lr * scale, "loss": [float(v) for v in loss],
                     "s": round(time.time() - t0, 2)})
    return net


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--batches", type=int, default=8)
    ap.add_argument("--sites", type=int, default=4096)
    ap.add_argument("--heldout-sites", type=int, default=1024)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--depth", type=int, default=4)
    ap.add_argument("--members", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--data", default=None, help="save / reuse the generated training set (.npz)")
    ap.add_argument("--no-linear-feature", action="store_true",
                    help="56 geometric features only (no superposition estimate as feature 56)")
    ap.add_argument("--absolute", action="store_true",
                    help="with the superposition feature, learn ln [Ca2+] itself instead of its residual")
    a = ap.parse_args(argv)

    def log(d):
        print(json.dumps(d), flush=True)

    from .dyad import DyadHIP
    from . import autotune, workloads
    spec = workloads.batched_dyads(n_sites=a.sites, clamped=True)
    t = autotune.lookup(spec)
    tuning = (t["planes"], t["threads"], t.get("prefetch", 1)) if t else None
    lin = None
    if not a.no_linear_feature:
        from . import linear_baseline
        cal = linear_baseline.calibrate(DyadHIP)
        lin = linear_baseline.feature_table(cal)
        lin.residual = not a.absolute
        log({"stage": "superposition_table", "entries": len(lin.added), "background_uM": lin.background})
    if a.data:
        try:
            z = np.load(a.data, allow_pickle=False)
            x, y = z["x"], z["y"]
            log({"stage": "data", "loaded": a.data, "rows": int(len(y))})
        except FileNotFoundError:
            x, y = dataset(DyadHIP, a.batches, a.sites, 0, tuning, log, lin)
            np.savez(a.data, x=x, y=y)
    else:
        x, y = dataset(DyadHIP, a.batches, a.sites, 0, tuning, log, lin)
    xh, yh, meta_h = surrogate.generate(DyadHIP, n_sites=a.heldout_sites, seed=10_000, tuning=tuning, lin=lin)
    # residual ensembles learn ln [Ca2+] minus the superposition estimate (feature 56)
    yt = y - x[:, surrogate.N_FEAT] if lin is not None and lin.residual else y
    net = fit(x, yt, a.width, a.depth, a.members, a.epochs, a.batch, a.lr, log=log)
    held = errors(net, xh, yh, lin=lin)
    fit_err = errors(net, x[:200_000], y[:200_000], lin=lin)
    log({"stage": "heldout", **held})
    log({"stage": "train_fit", **fit_err})
    if lin is not None:
        # the superposition estimate alone on the same held-out mouths
        true = np.exp(yh.astype(np.float64))
        rel = np.abs(np.exp(xh[:, surrogate.N_FEAT].astype(np.float64)) - true) / true
        log({"stage": "heldout_superposition_feature", "rel_mean": float(rel.mean()),
             "rel_p50": float(np.median(rel)), "rel_p95": float(np.quantile(rel, 0.95))})
    if a.out:
        extra = {} if lin is None else {"lin_added": lin.added, "lin_bg": np.array(lin.background),
                                        "lin_jsr": np.array(lin.jsr_ref), "lin_residual": np.array(lin.residual)}
        net.save(a.out, {"target": "ln mouth [Ca2+] after %d clamped steps of 10 ns from rest" % surrogate.QS_STEPS,
                         "training": "%d batches x %d sites, %d rows; %d epochs, batch %d, Adam lr %g step-decayed"
                                     % (a.batches, a.sites, len(y), a.epochs, a.batch, a.lr),
                         "features": x.shape[1], "heldout": held, "train_fit": fit_err}, **extra)
        log({"stage": "saved", "path": a.out})
    return 0


if __name__ == "__main__":
    sys.exit(main())
