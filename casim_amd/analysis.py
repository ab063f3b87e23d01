"""Release (spark) statistics and the KS comparison harness.

Definitions follow IndChannelCREAnalyzer (IndChannelCREAnalyzer.cpp:53-138):
per site the amplitude (max open channels), time to peak, time of the last
event, the integral of N dt and its mean over the event time; classes quark
(amplitude 1), blip (amplitude <= nadir) and spark (larger).  The device
accumulates the per-site columns step by step (casim_dyad_get_release_stats);
this module classifies them and compares two populations.
"""
from dataclasses import dataclass

import numpy as np

COLUMNS = ("amp", "ttp", "tte", "q", "q_mean", "n_now")


@dataclass
class Population:
    """Per-site release statistics of one run."""
    stats: np.ndarray        # [S, 6] as returned by DyadHIP.GetReleaseStats
    n_channels: np.ndarray   # [S]
    t_total: float           # simulated time of every site (ms)

    @property
    def amplitude(self):
        return self.stats[:, 0]

    @property
    def duration(self):
        return self.stats[:, 2]

    @property
    def open_probability(self):
        return self.stats[:, 3] / (self.n_channels * self.t_total)

    def classes(self, nadir=4):
        amp = self.amplitude
        quark = amp == 1
        blip = (~quark) & (amp <= nadir)
        spark = amp > nadir
        return {"quarks": int(quark.sum()), "blips": int(blip.sum()), "sparks": int(spark.sum())}


def ks_2samp(a, b):
    """Two-sample Kolmogorov-Smirnov statistic and p-value."""
    from scipy import stats
    r = stats.ks_2samp(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))
    return float(r.statistic), float(r.pvalue)


def compare(pa, pb, alpha=0.01):
    """KS on amplitude, duration and open probability; passes when every
    p-value exceeds alpha."""
    out = {}
    for name in ("amplitude", "duration", "open_probability"):
        d, p = ks_2samp(getattr(pa, name), getattr(pb, name))
        out[name] = {"D": d, "p": p, "pass": p > alpha}
    out["pass"] = all(v["pass"] for v in out.values() if isinstance(v, dict))
    return out


class StatsTracker:
    """The same per-step accumulation from a sequence of channel-state
    arrays (used to check the device statistics against the oracle)."""

    def __init__(self, states0, dt):
        n0 = (np.asarray(states0) == 1).sum(axis=1).astype(np.float64)
        self.dt = dt
        self.steps = 0
        self.n = n0
        self.amp = n0.copy()
        self.ttp = np.zeros_like(n0)
        self.tte = np.zeros_like(n0)
        self.q = np.zeros_like(n0)

    def step(self, states):
        nn = (np.asarray(states) == 1).sum(axis=1).astype(np.float64)
        self.steps += 1
        t = self.steps * self.dt
        self.tte = np.where(nn != self.n, t, self.tte)
        up = nn > self.amp
        self.ttp = np.where(up, t, self.ttp)
        self.amp = np.where(up, nn, self.amp)
        self.q = self.q + nn * self.dt
        self.n = nn

    def table(self):
        qm = np.where(self.tte > 0, self.q / np.where(self.tte > 0, self.tte, 1.0), 0.0)
        return np.stack([self.amp, self.ttp, self.tte, self.q, qm, self.n], axis=1)


def time_histograms(counts, amplitude, n_channels, nadir=4):
    """IndChannelCREAnalyzer's time histograms (IndChannelCREAnalyzer.cpp:53-137,
    CalculateGlobalHist :24-45) from per-bin open counts [bins, S] (255 = bin
    never reached, counted as nothing open): for every bin, the number of
    sites with k channels open, k >= 1, and in column 0 the remaining sites;
    for all sites and separately for quarks (amplitude 1), blips (amplitude
    <= nadir) and sparks."""
    counts = np.asarray(counts)
    amp = np.asarray(amplitude)
    classes = {"all": np.ones_like(amp, dtype=bool), "quarks": amp == 1,
               "blips": (amp != 1) & (amp <= nadir), "sparks": amp > nadir}
    width = {"all": n_channels + 1, "quarks": 2, "blips": nadir + 1, "sparks": n_channels + 1}
    out = {}
    for name, mask in classes.items():
        c = counts[:, mask].astype(np.int64)
        h = np.zeros((counts.shape[0], width[name]), dtype=np.int64)
        for k in range(1, width[name]):
            h[:, k] = (c == k).sum(axis=1)
        h[:, 0] = int(mask.sum()) - h[:, 1:].sum(axis=1)
        out[name] = h
    return out
