"""Mouth-[Ca2+] surrogate: training data from the full RD solver, per-RyR
features, and the quasi-stationary release-site path it drives.

Target.  For a site geometry, a clamped open set and a jSR concentration, the
target of every RyR is ln [Ca2+] at its mouth after a fixed relaxation of the
full reaction-diffusion solver from rest (QS_STEPS steps of dt = 10 ns, 0.5
ms): free Ca2+, the fast ATP-like buffer and most of the slower buffer
loading have settled, which is the quasi-stationary value gating sees while
an open set persists (QuasyStationaryReleaseSite takes a stationary
solution, StationaryDiffusionModel.cpp:3-11).  The same solve is the
ensemble's fallback.

Features of RyR i (lattice units of 31 nm, cleft centre at the origin):
  0  open flag of i            1  jSR [Ca2+] / 1000 uM     2  open RyRs / 100
  3  sum over open j != i of 1/r_ij          4  sum over open j != i of exp(-r_ij)
  5,6  position of i relative to the cleft centre      7  distance of i to the rim
  8 + 3k .. 10 + 3k  (dx, dy, 1/r) to the k-th nearest open RyR j != i,
                     k < 16, zeros when fewer are open
  56 (with a LinearFeature)  ln of the linear-superposition estimate of the
                     mouth [Ca2+]: background + jSR/jSR_ref x sum over open RyRs
                     (i included) of the single-channel table at r^2
"""
import ctypes as C
import dataclasses

import numpy as np

from . import abi, config, mlp, workloads

PITCH_NM = 31.0
N_NEAR = 16
N_FEAT = 8 + 3 * N_NEAR  # 56
QS_STEPS = 50_000
DT_MS = 1e-5


@dataclasses.dataclass
class LinearFeature:
    """Single-channel superposition table for feature 56: added [Ca2+] (uM)
    at integer r^2 (lattice units^2) 0..len-1, interpolated linearly in r^2,
    decaying as 1/r past the end; at jSR jsr_ref."""
    added: np.ndarray
    background: float
    jsr_ref: float
    residual: bool = False  # the ensemble predicts ln [Ca2+] minus feature 56

    def at(self, r2):
        r2 = np.asarray(r2, dtype=np.float64)
        top = len(self.added) - 1
        k = np.minimum(np.floor(r2).astype(np.int64), top - 1)
        w = r2 - k
        inside = self.added[k] + w * (self.added[k + 1] - self.added[k])
        far = self.added[top] * np.sqrt(top / np.maximum(r2, 1e-30))
        return np.where(r2 >= top, far, inside)

    def estimate(self, xy_lattice, open_mask, jsr_uM):
        op = np.asarray(open_mask, dtype=bool)
        d2 = ((xy_lattice[:, None, :] - xy_lattice[None, op, :]) ** 2).sum(axis=2)
        g = self.at(d2)
        return self.background + jsr_uM / self.jsr_ref * g.sum(axis=1)


def site_features(xy_nm, open_mask, jsr_uM, half_extent_nm, lin=None):
    """Feature rows [n, N_FEAT] (N_FEAT + 1 with a LinearFeature) of one site
    (see the module docstring)."""
    xy = (np.asarray(xy_nm, dtype=np.float64) - half_extent_nm) / PITCH_NM
    n = len(xy)
    op = np.asarray(open_mask, dtype=bool)
    f = np.zeros((n, N_FEAT + (1 if lin is not None else 0)))
    if lin is not None:
        f[:, N_FEAT] = np.log(lin.estimate(xy, op, jsr_uM))
    f[:, 0] = op
    f[:, 1] = jsr_uM / 1000.0
    f[:, 2] = op.sum() / 100.0
    f[:, 5:7] = xy
    f[:, 7] = (half_extent_nm / PITCH_NM) - np.abs(xy).max(axis=1)
    oi = np.nonzero(op)[0]
    if len(oi) == 0:
        return f
    d = xy[:, None, :] - xy[None, oi, :]          # [n, n_open, 2]: i minus j
    r = np.sqrt((d ** 2).sum(axis=2))
    self_hit = oi[None, :] == np.arange(n)[:, None]
    r_eff = np.where(self_hit, np.inf, r)
    inv = np.where(self_hit, 0.0, 1.0 / np.where(r > 0, r, 1.0))
    f[:, 3] = inv.sum(axis=1)
    f[:, 4] = np.where(self_hit, 0.0, np.exp(-r)).sum(axis=1)
    order = np.argsort(r_eff, axis=1, kind="stable")[:, :N_NEAR]
    for k in range(min(N_NEAR, order.shape[1])):
        j = order[:, k]
        rk = r_eff[np.arange(n), j]
        ok = np.isfinite(rk)
        f[ok, 8 + 3 * k] = -d[np.arange(n), j, 0][ok]
        f[ok, 9 + 3 * k] = -d[np.arange(n), j, 1][ok]
        f[ok, 10 + 3 * k] = 1.0 / rk[ok]
    return f


def batch_features(spec, states, jsr_site, lin=None):
    half = (spec.nx - 1) * spec.dx / 2.0
    rows = []
    for s, site in enumerate(spec.sites):
        xy = np.array([(c[0], c[1]) for c in site])
        rows.append(site_features(xy, states[s, :len(site)] == 1, jsr_site[s], half, lin))
    return np.concatenate(rows, axis=0)


def random_open_sets(spec, gen):
    """Open fraction uniform in 0..1 per site, at least one RyR open."""
    M = spec.max_channels()
    st = np.full((spec.n_sites, M), -1, dtype=np.int32)
    for s, site in enumerate(spec.sites):
        n = len(site)
        p = gen.uniform(0.0, 1.0)
        o = (gen.random(n) < p).astype(np.int32)
        if o.sum() == 0:
            o[gen.integers(n)] = 1
        st[s, :n] = o
    return st


def jsr_layout(spec, jsr_site):
    """Per-site jSR values -> [ch + n_ch*(site + ion*S)]."""
    M, S = spec.max_channels(), spec.n_sites
    out = np.zeros(M * S * spec.n_ions)
    for s in range(S):
        out[M * s:M * (s + 1)] = jsr_site[s]
    return out


_RATE_PARAMS = {"SheepCannell": ("kopen0", "kclose0"), "TwoStateMWC": ("kclose",), "TwoStateaHTG": ("kclose",)}


def clamped_spec(spec):
    """The same sites with every gating rate zero, so the channels keep the
    states they are given while the solver relaxes."""
    types = []
    for ct in spec.channel_types:
        p = dict(ct.params)
        for k in _RATE_PARAMS[ct.model]:
            p[k] = 0.0
        types.append(dataclasses.replace(ct, params=p))
    return dataclasses.replace(spec, channel_types=types)


def small_batch_tuning(spec, workgroups=1024):
    """Plane chunks small enough that a batch of few sites still spreads over
    about `workgroups` workgroups (a whole-site workgroup marches 79 planes
    one after another)."""
    inner = spec.nx - 2
    chunks = min(inner, -(-workgroups // max(spec.n_sites, 1)))
    if chunks <= 1:
        return None
    return (-(-inner // chunks), 0, 1)


def qs_solve(DyadHIP, spec, states, jsr_site, steps=QS_STEPS, tuning=None):
    """Mouth [Ca2+] of every RyR after `steps` clamped steps from rest: [S, M]."""
    dev = DyadHIP(clamped_spec(spec))
    tuning = tuning or small_batch_tuning(spec)
    if tuning:
        dev.SetTuning(*tuning)
    dev.SetChannelsStates(states)
    ri, rb = spec.rim_arrays()
    dev.Update(ri, rb, jsr_layout(spec, jsr_site))
    done = 0
    while done < steps:
        k = min(1000, steps - done)
        dev.Step(DT_MS, k)
        done += k
    m = dev.GetIonsNearChannels()[0].copy()
    dev.close()
    return m


def generate(DyadHIP, n_sites=4096, seed=0, steps=QS_STEPS, tuning=None, lin=None):
    """One batch of training samples: (features [rows, F], ln Ca [rows], meta)."""
    spec = workloads.batched_dyads(n_sites=n_sites, geometry_seed=10_000 + seed, clamped=True)
    gen = np.random.default_rng(20_000 + seed)
    states = random_open_sets(spec, gen)
    jsr = gen.uniform(200.0, 1200.0, n_sites)
    mouth = qs_solve(DyadHIP, spec, states, jsr, steps, tuning)
    x = batch_features(spec, states, jsr, lin)
    y = np.concatenate([np.log(mouth[s, :len(site)]) for s, site in enumerate(spec.sites)])
    return x.astype(np.float32), y.astype(np.float32), {"spec": spec, "states": states, "jsr": jsr, "mouth": mouth}


def load_model(path):
    """(ensemble, LinearFeature or None) from a file written by
    train_surrogate (mlp.MLP.save with the superposition table)."""
    net = mlp.MLP.load(path)
    z = np.load(path, allow_pickle=False)
    lin = None
    if "lin_added" in z:
        lin = LinearFeature(np.asarray(z["lin_added"], dtype=np.float64), float(z["lin_bg"]), float(z["lin_jsr"]),
                            bool(z["lin_residual"]) if "lin_residual" in z else False)
    return net, lin


def ensemble_mean(net, x, lin=None):
    """Mouth [Ca2+] of feature rows x: the mean over members of exp(output),
    plus feature 56 inside the exponent for a residual ensemble; and the
    relative spread."""
    _, _, raw = net.predict(x)
    raw = raw.astype(np.float64)
    if lin is not None and lin.residual:
        raw = raw + np.asarray(x, dtype=np.float64)[:, N_FEAT][None, :]
    v = np.exp(raw)
    mean = v.mean(axis=0)
    return mean, v.std(axis=0) / mean


class CasimQSConfig(C.Structure):
    _fields_ = [
        ("n_sites", C.c_int32),
        ("site_channel_offset", C.POINTER(C.c_int32)),
        ("channel_x", C.POINTER(C.c_double)),
        ("channel_y", C.POINTER(C.c_double)),
        ("half_extent", C.c_double),
        ("pitch", C.c_double),
        ("model", C.c_int32),
        ("params", C.POINTER(C.c_double)),
        ("mg", C.c_double),
        ("jsr", C.POINTER(C.c_double)),
        ("rng_state", C.POINTER(C.c_uint64)),
        ("init_channel", C.POINTER(C.c_int32)),
        ("t_max", C.c_double),
        ("spread_threshold", C.c_double),
        ("jsr_deplete", C.c_int32),
        ("jsr_volume", C.c_double),
        ("element_volume", C.c_double),
        ("conductance", C.c_double),
        ("jsr_cq_total", C.c_double),
        ("jsr_cq_sites", C.c_double),
        ("jsr_k_cq", C.c_double),
        ("jsr_tau_refill", C.c_double),
        ("nsr", C.c_double),
    ]


class QuasiStationarySites:
    """QuasyStationaryReleaseSite::RunSimulation (QuasyStationaryReleaseSite.cpp:89-133)
    for every site of `spec` at once on the device, the mouth [Ca2+] of each
    open set from the surrogate ensemble `net` (an mlp.MLP with N_FEAT inputs).

    Per site: InitOpening (one draw per channel, one for init_channel), then
    Gillespie events (waiting time, channel choice, MakeTransition's draw)
    until no RyR is open or t >= t_max (ms).  A site whose ensemble spread
    exceeds `spread_threshold` at any RyR takes its next event with the full
    solver's mouth values instead (RunSimulation's `fallback`)."""

    def __init__(self, spec, jsr_site, t_max, net, spread_threshold=0.02, mg=0.0, rng_states=None, pools=None,
                 lin=None, stationary=None):
        """pools: a config.PoolSpec with jsr_constant False makes the jSR
        deplete between events (SRPool's pool, calcium only); jSR starts at
        jsr_site.  stationary: a stationary.StationaryModel (the reference's
        LinearNaraghiNeher) that gives the mouth values instead of the
        ensemble (net may then be None); it does not follow the jSR."""
        if net is None and stationary is None:
            raise ValueError("QuasiStationarySites needs a surrogate ensemble or a stationary model")
        self.lib = mlp.lib()
        self.spec, self.net = spec, net
        self.S = spec.n_sites
        self.jsr = np.ascontiguousarray(jsr_site, dtype=np.float64)
        ct = spec.channel_types[0]
        off = np.zeros(self.S + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(s) for s in spec.sites])
        self.off = off
        self.M = int(max(len(s) for s in spec.sites))
        xs = np.array([c[0] for s in spec.sites for c in s], dtype=np.float64)
        ys = np.array([c[1] for s in spec.sites for c in s], dtype=np.float64)
        par = np.zeros(abi.CASIM_MAX_MODEL_PARAMS)
        for k, pn in enumerate(config.MODEL_PARAMS[ct.model]):
            par[k] = ct.params[pn]
        rs = np.ascontiguousarray(spec.rng_states if rng_states is None else rng_states, dtype=np.uint64)
        init = np.ascontiguousarray(spec.init_channel or [0] * self.S, dtype=np.int32)
        self.half = (spec.nx - 1) * spec.dx / 2.0
        self._keep = (off, xs, ys, par, self.jsr, rs, init)
        dep = pools is not None and not pools.jsr_constant
        cfg = CasimQSConfig(self.S, abi.iptr(off), abi.dptr(xs), abi.dptr(ys), self.half, PITCH_NM,
                            config.MODEL_IDS[ct.model], abi.dptr(par), float(mg), abi.dptr(self.jsr),
                            abi.uptr(rs), abi.iptr(init), float(t_max), float(spread_threshold),
                            int(dep), float(pools.jsr_volume) if dep else 0.0,
                            config.element_volume(spec), float(ct.conductance[0]),
                            float(pools.jsr_cq_total) if dep else 0.0, float(pools.jsr_cq_sites[0]) if dep else 0.0,
                            float(pools.jsr_k_cq[0]) if dep else 0.0, float(pools.jsr_tau_refill[0]) if dep else 1.0,
                            float(pools.nsr_c[0]) if dep else 0.0)
        h = C.c_void_p()
        abi.check(self.lib.casim_qs_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.lin = lin
        if lin is not None:
            tab = np.ascontiguousarray(lin.added, dtype=np.float32)
            abi.check(self.lib.casim_qs_set_linear_feature(self.h, mlp.fptr(tab), len(tab), float(lin.background),
                                                           float(lin.jsr_ref), int(lin.residual)))
        self.stationary = stationary
        if stationary is not None:
            coef = np.ascontiguousarray(stationary.coef, dtype=np.float64)
            decay = np.ascontiguousarray(stationary.decay, dtype=np.float64)
            self._keep_st = (coef, decay)
            abi.check(self.lib.casim_qs_set_stationary_model(self.h, len(coef), abi.dptr(coef), abi.dptr(decay),
                                                             float(stationary.background), float(stationary.ca_open)))
        self.fallback_solves = 0
        self.fallback_s = 0.0
        self.rounds = 0

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "h", None):
            self.lib.casim_qs_destroy(self.h)
            self.h = None

    def Reset(self):
        abi.check(self.lib.casim_qs_reset(self.h))
        self.fallback_solves = 0
        self.fallback_s = 0.0
        self.rounds = 0

    def Round(self, max_rounds=1, wait_batch=1 << 30):
        """Up to max_rounds events per active, non-waiting site;
        (rounds run, active sites, waiting sites)."""
        r, a, w = C.c_int32(), C.c_int32(), C.c_int32()
        net = self.net.h if self.stationary is None else None
        abi.check(self.lib.casim_qs_run(self.h, net, int(max_rounds), int(wait_batch), C.byref(r),
                                        C.byref(a), C.byref(w)))
        return r.value, a.value, w.value

    def Waiting(self):
        flags = np.zeros(self.S, dtype=np.int32)
        states = np.zeros((self.S, self.M), dtype=np.int32)
        abi.check(self.lib.casim_qs_get_waiting(self.h, abi.iptr(flags), abi.iptr(states)))
        return flags, states

    def GetChannelsStates(self):
        return self.Waiting()[1]

    def SetMouths(self, site, ca):
        ca = np.ascontiguousarray(ca, dtype=np.float32)
        abi.check(self.lib.casim_qs_set_mouths(self.h, int(site), mlp.fptr(ca)))

    def GetStats(self):
        """[S, 6]: amplitude, TTP, TTE, Q, Q/TTE, events."""
        out = np.zeros((self.S, 6))
        abi.check(self.lib.casim_qs_get_stats(self.h, abi.dptr(out)))
        return out

    def RecordOpenCounts(self, bin_ms, max_bins):
        """Per-bin open counts for the analyzer's time histograms; takes
        effect from the next Reset."""
        abi.check(self.lib.casim_qs_record_open_counts(self.h, float(bin_ms), int(max_bins)))
        self._bins = max_bins

    def GetOpenCounts(self):
        out = np.empty((self._bins, self.S), dtype=np.uint8)
        abi.check(self.lib.casim_qs_get_open_counts(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def GetWindowStats(self):
        """[S, 4] over [0, t_max]: amplitude, last event time, integral of N dt, N at t_max."""
        out = np.zeros((self.S, 4))
        abi.check(self.lib.casim_qs_get_window_stats(self.h, abi.dptr(out)))
        return out

    def GetFeatures(self):
        out = np.zeros((int(self.off[-1]), N_FEAT + (1 if self.lin is not None else 0)), dtype=np.float32)
        abi.check(self.lib.casim_qs_get_features(self.h, mlp.fptr(out)))
        return out

    def GetJsr(self):
        """(jSR now [S], lowest jSR reached [S])."""
        now, low = np.zeros(self.S), np.zeros(self.S)
        abi.check(self.lib.casim_qs_get_jsr(self.h, abi.dptr(now), abi.dptr(low)))
        return now, low

    def GetMouths(self):
        m = np.zeros(int(self.off[-1]), dtype=np.float32)
        s = np.zeros_like(m)
        abi.check(self.lib.casim_qs_get_mouths(self.h, mlp.fptr(m), mlp.fptr(s)))
        return m, s

    def resolve_waiting(self, tuning=None):
        """Full-solver mouth values for every waiting site (qs_solve on a
        batch of just those sites)."""
        from .dyad import DyadHIP
        flags, states = self.Waiting()
        idx = np.nonzero(flags)[0]
        if len(idx) == 0:
            return 0
        sub = dataclasses.replace(self.spec, sites=[self.spec.sites[s] for s in idx],
                                  rng_states=np.asarray(self.spec.rng_states)[idx],
                                  init_channel=[0] * len(idx))
        st = np.full((len(idx), sub.max_channels()), -1, dtype=np.int32)
        for k, s in enumerate(idx):
            n = len(self.spec.sites[s])
            st[k, :n] = states[s, :n]
        mouth = qs_solve(DyadHIP, sub, st, self.jsr[idx], QS_STEPS, tuning)
        for k, s in enumerate(idx):
            self.SetMouths(s, mouth[k, :len(self.spec.sites[s])])
        self.fallback_solves += len(idx)
        return len(idx)

    def RunSimulation(self, max_rounds=1 << 30, fallback=True, wait_batch=256, tuning=None):
        """Run every site to its end; returns GetStats().  Sites that ask for
        the full solver are resolved in batches of up to wait_batch."""
        import time
        done = 0
        while done < max_rounds:
            r, a, w = self.Round(max_rounds - done, wait_batch)
            done += r
            self.rounds += r
            if a == 0:
                break
            if w:
                if not fallback:
                    raise RuntimeError("%d sites need the full solver and fallback is off" % w)
                t0 = time.perf_counter()
                self.resolve_waiting(tuning)
                self.fallback_s += time.perf_counter() - t0
            elif r == 0:
                break
        return self.GetStats()
