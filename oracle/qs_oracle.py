"""CPU restatement of the quasi-stationary release site's event loop
(QuasyStationaryReleaseSite::RunSimulation, QuasyStationaryReleaseSite.cpp:89-133)
and of the release-statistics columns IndChannelCREAnalyzer computes from an
event record (IndChannelCREAnalyzer.cpp:69-117).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
quasi-stationary path.  Pure Python loops over the oracle's MRG32k3a and
channel rates (oracle.rng_draws, oracle.channel_rate); for small cases.

The mouth [Ca2+] of an open set is a parameter (mouth_fn): the reference
takes it from a StationaryDiffusionModel, the device path from the
surrogate ensemble, so the tests pass the device's own values in.
"""
import math

import numpy as np

from . import oracle


class Stream:
    """One MRG32k3a stream drawn one U01d at a time (ParallelRng::runif)."""

    def __init__(self, state):
        self.state = np.ascontiguousarray(state, dtype=np.uint64).copy()

    def runif(self):
        u, self.state = oracle.rng_draws(self.state, 1)
        return float(u[0])

    def rexp(self, theta):
        # ParallelRng.cpp:37-39
        return -math.log(self.runif()) * theta


def init_opening(n, init_channel, stream):
    """QuasyStationaryReleaseSite.cpp:91-97: SetMacroState(0, runif) on every
    channel, then SetMacroState(1, runif) on init_channel (two-state models
    ignore the draw, the stream still advances)."""
    states = np.zeros(n, dtype=np.int32)
    for _ in range(n):
        stream.runif()
    stream.runif()
    states[init_channel] = 1
    return states


class JsrPool:
    """SRPool's jSR (SRPool.cpp:70-89) for the quasi-stationary path: over an
    interval with a fixed open set, dc/dt = ((nsr - c)/tau - c geff Vr)/beta,
    beta = 1 + n K CQ/(K + c)^2, explicit steps of at most 1 us; geff = sum
    over open RyRs of g (1 - c_mouth/c) at the interval's start (mouth [Ca2+]
    proportional to c within it), Vr = V_elem / V_jSR."""

    def __init__(self, c0, g, vr, cq, ncq, kcq, tau, nsr):
        self.c, self.low = float(c0), float(c0)
        self.g, self.vr, self.cq, self.ncq, self.kcq, self.tau, self.nsr = g, vr, cq, ncq, kcq, tau, nsr

    def _run(self, c, geff, dt):
        n = float(math.ceil(dt / 1e-3))
        n = min(max(n, 1.0), 1e6)
        h = dt / n
        for _ in range(int(n)):
            kc = self.kcq + c
            beta = 1 + self.ncq * self.kcq * self.cq / (kc * kc)
            iref = (self.nsr - c) / self.tau
            c = c + h / beta * (iref - c * geff * self.vr)
        return c

    def advance(self, states, ca, dt, t_prev=0.0, t_max=math.inf):
        """Advance over dt from t_prev; `low` keeps the lowest value over
        [0, t_max] (the value at t_max when the interval crosses it)."""
        c0 = self.c
        geff = 0.0
        for i in range(len(states)):
            if states[i] == 1:
                geff += self.g * (1.0 - float(np.float32(ca[i])) / c0)
        self.c = self._run(c0, geff, dt)
        if t_prev + dt <= t_max:
            self.low = min(self.low, self.c)
        elif t_prev < t_max:
            self.low = min(self.low, self._run(c0, geff, t_max - t_prev))


def event(states, ca, model, params, mg, stream, t, pool=None, t_max=math.inf):
    """One Gillespie event (QuasyStationaryReleaseSite.cpp:101-128) with mouth
    [Ca2+] `ca` for the current states; returns (t, chosen channel).  The
    channel flips (TwoStatePoModel / TwoStateModel::MakeTransition).  With a
    JsrPool the jSR advances over the interval before the flip.  With every
    rate zero nothing can happen: the pool advances to t_max and (inf, -1)
    comes back."""
    n = len(states)
    rates = [oracle.channel_rate(model, params, int(states[i]), float(ca[i]), mg) for i in range(n)]
    total = 0.0
    for r in rates:
        total += r
    if not total > 0.0:
        # nothing can happen any more; open channels keep drawing on the jSR to t_max
        if pool is not None and t < t_max:
            pool.advance(states, ca, t_max - t, t, t_max)
        return math.inf, -1
    t_start = t
    dtw = stream.rexp(1.0 / total)
    t = t + dtw
    u = stream.runif()
    p = 0.0
    idx = 0
    for i in range(n):
        w = rates[i] / total
        if p <= u and u <= p + w:
            idx = i
            break
        p += w
    stream.runif()  # MakeTransition(runif())
    if pool is not None:
        pool.advance(states, ca, dtw, t_start, t_max)
    states[idx] ^= 1
    return t, idx


def run_site(n, init_channel, rng_state, model, params, mg, t_max, mouth_fn, pool=None):
    """Whole RunSimulation of one site: the record [(t, states)] and the
    stream state after it.  mouth_fn(states, jsr) gives the mouth [Ca2+]."""
    stream = Stream(rng_state)
    states = init_opening(n, init_channel, stream)
    t = 0.0
    rec = [(t, states.copy())]
    n_open = 1
    while n_open >= 1 and t < t_max:
        jsr = pool.c if pool is not None else None
        t, idx = event(states, mouth_fn(states, jsr), model, params, mg, stream, t, pool, t_max)
        if idx < 0:  # no event can happen any more
            break
        n_open = int(states.sum())
        rec.append((t, states.copy()))
    return rec, stream.state


FAR = 96485.33e-6  # StationaryDiffusionModel.h: Far
PI = math.acos(-1.0)


def linear_naraghi_neher(buffers, current, ca0, dca):
    """LinearNaraghiNeher.cpp:9-64: the linearised steady state of one point
    source with mobile buffers.  buffers: {name: {"kon", "koff", "D",
    "Total Concentration"}} (iterated in name order, like the std::map).
    Per buffer k = B K / (K + ca0)^2, tau = 1 / (koff + kon ca0); with
    Dm = diag(D_1 .. D_n, dca) and AA the (n+1) x (n+1) rate matrix of
    (bound_1 .. bound_n, Ca), B = -Dm^-1 AA; its eigenpairs (mu, V) (real
    parts; the smallest eigenvalue moved last, as the reference does) and
    aa = V^-1 u, u = (0 .. 0, current / (4 pi Far dca)).  Returns (aa, mu, V)."""
    k, tau, d = [], [], []
    for name in sorted(buffers):
        b = buffers[name]
        K = b["koff"] / b["kon"]
        k.append(b["Total Concentration"] * K / ((K + ca0) * (K + ca0)))
        tau.append(1.0 / (b["koff"] + b["kon"] * ca0))
        d.append(b["D"])
    n = len(k)
    Dm = np.zeros((n + 1, n + 1))
    for i in range(n):
        Dm[i, i] = d[i]
    Dm[n, n] = dca
    AA = np.zeros((n + 1, n + 1))
    for i in range(n):
        AA[i, i] = -1.0 / tau[i]
        AA[i, n] = k[i] / tau[i]
        AA[n, i] = -AA[i, i]
        AA[n, n] -= AA[i, n]
    B = -np.linalg.inv(Dm) @ AA
    w, v = np.linalg.eig(B)
    mu, V = np.real(w).copy(), np.real(v).copy()
    j = int(np.argmin(mu))
    mu[[j, n]] = mu[[n, j]]
    V[:, [j, n]] = V[:, [n, j]]
    u = np.zeros(n + 1)
    u[n] = current / (4 * PI * FAR * dca)
    aa = np.linalg.solve(V, u)
    return aa, mu, V


def nn_added(aa, mu, V, r, ca_open):
    """One entry of added_concentrations (LinearNaraghiNeher.cpp:66-71): the
    [Ca2+] an open channel adds at distance r, ca_open at its own mouth."""
    if not r > 1e-6:
        return ca_open
    n = len(aa) - 1
    dca = 0.0
    for ii in range(n + 1):
        dca += 1 / r * aa[ii] * math.exp(-r * math.sqrt(abs(mu[ii])) * 1e-8) * V[n, ii] * 1e-4
    return dca


def stationary_table(xy, aa, mu, V, ca_open):
    """added_concentrations[i][j] over the channel positions xy [n, 2]
    (Point::d with z = 0)."""
    n = len(xy)
    tab = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            dx, dy = xy[i][0] - xy[j][0], xy[i][1] - xy[j][1]
            tab[i, j] = nn_added(aa, mu, V, math.sqrt(dx * dx + dy * dy + 0.0), ca_open)
    return tab


def stationary_concentrations(tab, states, background):
    """StationaryDiffusionModel::GetConcentrations (StationaryDiffusionModel.cpp:3-11)."""
    n = len(states)
    out = [0.0] * n
    for j in range(n):
        for i in range(n):
            out[j] += float(states[i] == 1) * tab[i, j]
        out[j] += background
    return out


def analyze(rec):
    """amplitude, TTP, TTE, Q, Q/TTE and the number of events of one record
    (IndChannelCREAnalyzer.cpp:69-117: amp / ttp from the first strict
    maximum, Q = sum of N_j (t_{j+1} - t_j), TTE = time of the last record)."""
    amp = ttp = q = 0.0
    for j, (t, st) in enumerate(rec):
        no = int(st.sum())
        if no > amp:
            amp, ttp = no, t
        nxt = min(j + 1, len(rec) - 1)
        q += no * (rec[nxt][0] - t)
    tte = rec[-1][0]
    return np.array([amp, ttp, tte, q, q / tte if tte > 0 else 0.0, len(rec) - 1])


def analyze_window(rec, t_max):
    """The record cut at t_max: amplitude, time of the last event, integral of
    N dt and N, over [0, t_max] -- what a fixed-step run of length t_max
    reports for the same history."""
    amp = 0.0
    tte = 0.0
    q = 0.0
    n_end = 0.0
    for j, (t, st) in enumerate(rec):
        if t > t_max:
            break
        no = float(st.sum())
        amp = max(amp, no)
        tte = t
        n_end = no
        t_next = rec[j + 1][0] if j + 1 < len(rec) else t_max
        q += no * (min(t_next, t_max) - t)
    return np.array([amp, tte, q, n_end])


def open_count_bins(rec, bin_dt, n_bins):
    """Per-bin open counts of one record as IndChannelCREAnalyzer fills its
    time histogram (IndChannelCREAnalyzer.cpp:69-84, WriteIntoHistogram
    :18-22): record j covers bins from the first unwritten one up to
    trunc(t_{j+1} / dt) (the last record covers none); 255 = never written."""
    out = np.full(n_bins, 255, dtype=np.uint8)
    cur = 0
    for j, (t, st) in enumerate(rec):
        nxt = rec[min(j + 1, len(rec) - 1)][0]
        nidx = int(nxt / bin_dt)
        for k in range(cur, min(nidx, n_bins - 1) + 1):
            out[k] = int(st.sum())
        cur = nidx + 1
    return out
