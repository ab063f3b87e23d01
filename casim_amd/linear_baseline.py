"""Non-learned baseline for the surrogate: linear superposition of
single-channel solutions, the structure of the reference's quasi-stationary
diffusion model (StationaryDiffusionModel::GetConcentrations,
StationaryDiffusionModel.cpp:3-11: c_j = background + sum over open i of
added_concentrations[i][j]; LinearNaraghiNeher.cpp:9-75 fills that table from
a linearised point-source solution in free 3D space and sets the open
channel's own mouth to a constant).

Here the table entry for a pair at distance r is the full solver's own
single-channel answer: one open RyR on the 31 nm lattice at the cleft centre,
relaxed like the surrogate's target, read at every other lattice position --
the best a distance-only superposition can do in this cleft, so a surrogate
that beats it beats any such table.  Added [Ca2+] scales with jSR / jSR_cal
(the driving force of the channel current).

    python -m casim_amd.linear_baseline --heldout-sites 1024
"""
import argparse
import json
import sys

import numpy as np

from . import surrogate, workloads

JSR_CAL = 1000.0


def calibrate(DyadHIP, jsr=JSR_CAL, steps=surrogate.QS_STEPS):
    """Added [Ca2+] against lattice distance r^2 (in lattice units squared)
    from one open RyR, and the background [Ca2+] with none open."""
    spec = workloads.batched_dyads(n_sites=2, geometry_seed=1, clamped=True)
    side = 10
    centre = (spec.nx - 1) * spec.dx / 2.0
    cells = [(i, j) for i in range(side) for j in range(side)]
    site = [(centre + surrogate.PITCH_NM * (i - (side - 1) / 2.0), centre + surrogate.PITCH_NM * (j - (side - 1) / 2.0), 0)
            for i, j in cells]
    spec.sites = [site, site]
    spec.init_channel = [0, 0]
    src = cells.index((4, 4))
    st = np.zeros((2, len(site)), dtype=np.int32)
    st[0, src] = 1  # site 0: one open channel; site 1: none
    mouth = surrogate.qs_solve(DyadHIP, spec, st, np.array([jsr, jsr]), steps)
    bg = float(np.median(mouth[1]))
    table = {}
    for k, (i, j) in enumerate(cells):
        r2 = (i - 4) ** 2 + (j - 4) ** 2
        table.setdefault(r2, []).append(mouth[0, k] - bg)
    r2 = np.array(sorted(table))
    added = np.array([np.mean(table[v]) for v in r2])
    return {"r2": r2, "added": added, "background": bg, "jsr": jsr}


def feature_table(cal):
    """The calibration as a dense LinearFeature table over integer r^2
    (linear interpolation between the lattice distances it was read at)."""
    from .surrogate import LinearFeature
    top = int(cal["r2"][-1])
    dense = np.interp(np.arange(top + 1), cal["r2"], cal["added"])
    return LinearFeature(dense, float(cal["background"]), float(cal["jsr"]))


def predict(cal, xy_nm, open_mask, jsr_uM):
    """Mouth [Ca2+] of every RyR of one site by superposition."""
    xy = np.asarray(xy_nm, dtype=np.float64) / surrogate.PITCH_NM
    op = np.asarray(open_mask, dtype=bool)
    d2 = ((xy[:, None, :] - xy[None, op, :]) ** 2).sum(axis=2)
    # lattice positions: r^2 is an integer; beyond the table, the last entry decays as 1/r
    r2 = np.rint(d2).astype(np.int64)
    idx = np.searchsorted(cal["r2"], r2)
    idx = np.minimum(idx, len(cal["r2"]) - 1)
    hit = cal["r2"][idx] == r2
    far = cal["added"][-1] * np.sqrt(cal["r2"][-1] / np.maximum(d2, 1e-12))
    g = np.where(hit, cal["added"][idx], far)
    return cal["background"] + jsr_uM / cal["jsr"] * g.sum(axis=1)


def evaluate(cal, meta):
    spec, states, jsr, mouth = meta["spec"], meta["states"], meta["jsr"], meta["mouth"]
    rel = []
    for s, site in enumerate(spec.sites):
        n = len(site)
        xy = np.array([(c[0], c[1]) for c in site])
        pred = predict(cal, xy, states[s, :n] == 1, jsr[s])
        true = mouth[s, :n]
        rel.append(np.abs(pred - true) / true)
    rel = np.concatenate(rel)
    true = np.concatenate([meta["mouth"][s, :len(site)] for s, site in enumerate(spec.sites)])
    hi = true > 5.0
    return {"rows": int(len(rel)), "rel_mean": float(rel.mean()), "rel_p50": float(np.median(rel)),
            "rel_p95": float(np.quantile(rel, 0.95)), "rel_mean_above_5uM": float(rel[hi].mean()),
            "rel_p95_above_5uM": float(np.quantile(rel[hi], 0.95))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--heldout-sites", type=int, default=1024)
    ap.add_argument("--weights", default="casim_amd/surrogate_weights.npz")
    a = ap.parse_args(argv)

    def log(d):
        print(json.dumps(d), flush=True)

    from .dyad import DyadHIP
    cal = calibrate(DyadHIP)
    log({"stage": "calibrated", "background_uM": cal["background"],
         "added_uM": {int(r): float(v) for r, v in zip(cal["r2"][:8], cal["added"][:8])}})
    net, lin = surrogate.load_model(a.weights)
    xh, yh, meta = surrogate.generate(DyadHIP, n_sites=a.heldout_sites, seed=10_000, lin=lin)
    log({"stage": "linear_superposition", **evaluate(cal, meta)})
    from .train_surrogate import errors
    log({"stage": "surrogate", **errors(net, xh, yh, lin=lin)})
    return 0


if __name__ == "__main__":
    sys.exit(main())
