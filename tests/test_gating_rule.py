"""Pins of the oracle's gating draw rule (DyadChannels::InitOpening / RunMC,
DyadChannels.cpp:83-95,97-130, TwoStateModel.cpp:13-15) that do not rely on
the oracle's own reading of it: how far each site's stream advances, checked
against the reference's RngStream (oracle/_ref, rngstream.cpp compiled
unmodified), and the statistics that a two-state Bernoulli chain at constant
rates must show (stationary P_o, geometric dwell times, the transient's
closed form).  Also the one-step closed forms of CytosolPool's leak and SERCA
terms (CytosolPool.cpp:240-256), including the SERCA background read
ions_bg[idx_SERCA]."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from casim_amd import config, rng
from tests.conftest import GOLDEN
from tests.helpers import spec

DT = 1e-5


def sc_params(p_open, p_close, dt=DT):
    """SheepCannell parameters with Ca-independent rates (ncaopen = ncaclose = 0):
    kopen = kopen0 1e-3, kclose = kclose0 1e-3 (SheepCannell.cpp:12-18), chosen
    so that kopen dt = p_open and kclose dt = p_close."""
    return {"kopen0": p_open / dt * 1e3, "ncaopen": 0.0, "kopenmax": 1e300, "kclose0": p_close / dt * 1e3,
            "ncaclose": 0.0}


def tiny(n_ch, n_sites, params, first_stream=0):
    """2D 9 x 9 cleft, n_ch channels on distinct interior nodes, no current
    (g = 0), so the mouth [Ca2+] stays at rest and the rates stay constant."""
    pts = [(4.0 * (1 + (k % 7)), 4.0 * (1 + (k // 7))) for k in range(n_ch)]
    return spec(dim=2, nx=9, ny=9, h=4.0, channels=pts, n_sites=n_sites, g=0.0, params=params, clamped=False,
                first_stream=first_stream)


def ref_state_after(ref_lib, k, n):
    """State of the k-th RngStream after SetPackageSeed(default) once n RandU01 draws are taken."""
    out = np.empty(max(n, 1))
    st = (C.c_ulong * 6)()
    assert ref_lib.ref_stream_draws((C.c_ulong * 6)(*rng.DEFAULT_SEED), k, n, out.ctypes.data_as(
        C.POINTER(C.c_double)), st) == 0
    return np.array(list(st), dtype=np.uint64), out[:n]


def run_oracle(oracle_mod, sp, steps, init_channel=0):
    o = oracle_mod.Oracle(sp)
    o.init_opening([init_channel] * sp.n_sites)
    ri, rb = sp.rim_arrays()
    o.update(ri, rb, sp.jsr_array([1000.0]))
    after_init = o.rng_states().copy()
    if steps:
        o.step(DT, steps, threads=2)
    return o, after_init


@pytest.mark.parametrize("n_ch", [1, 5, 20])
def test_init_opening_takes_n_plus_one_draws(ref_lib, oracle_mod, n_ch):
    """SetMacroState(0, runif()) per channel, then SetMacroState(1, runif())."""
    sp = tiny(n_ch, 3, sc_params(0.0, 0.0), first_stream=4)
    o, after = run_oracle(oracle_mod, sp, 0, init_channel=n_ch - 1)
    for s in range(3):
        want, _ = ref_state_after(ref_lib, 4 + s, n_ch + 1)
        assert np.array_equal(after[s], want)
    st = o.channel_states()
    assert (st[:, :n_ch] == ([0] * (n_ch - 1) + [1])).all()


def test_zero_rate_steps_take_one_draw_per_channel(ref_lib, oracle_mod):
    """rate dt = 0: u1 < 0 never holds, so each step draws u1 only."""
    n, k = 7, 9
    sp = tiny(n, 2, sc_params(0.0, 0.0))
    o, _ = run_oracle(oracle_mod, sp, k)
    for s in range(2):
        want, _ = ref_state_after(ref_lib, s, n + 1 + k * n)
        assert np.array_equal(o.rng_states()[s], want)
    assert (o.channel_states()[:, :n] == [1] + [0] * (n - 1)).all()


def test_certain_toggles_take_two_draws_per_channel(ref_lib, oracle_mod):
    """rate dt >= 1 in both states: every channel toggles every step and
    MakeTransition(runif()) takes the second draw, so k steps take 2 k n
    draws and the states alternate."""
    n = 6
    for k in (1, 2, 5):
        sp = tiny(n, 2, sc_params(1.0, 1.5))
        o, _ = run_oracle(oracle_mod, sp, k)
        for s in range(2):
            want, _ = ref_state_after(ref_lib, s, n + 1 + 2 * k * n)
            assert np.array_equal(o.rng_states()[s], want), k
        init = np.array([1] + [0] * (n - 1))
        assert (o.channel_states()[:, :n] == (init ^ (k % 2))).all(), k


def test_one_sided_rates_count_toggles_only(ref_lib, oracle_mod):
    """kopen dt >= 1, kclose = 0: the first step opens the n - 1 closed
    channels (two draws each) and draws once for the open one; later steps
    draw once per channel.  Total n + 1 + (2 (n - 1) + 1) + (k - 1) n."""
    n, k = 5, 4
    sp = tiny(n, 3, sc_params(2.0, 0.0))
    o, _ = run_oracle(oracle_mod, sp, k)
    total = n + 1 + (2 * (n - 1) + 1) + (k - 1) * n
    for s in range(3):
        want, _ = ref_state_after(ref_lib, s, total)
        assert np.array_equal(o.rng_states()[s], want)
    assert (o.channel_states()[:, :n] == 1).all()


def test_draw_order_against_reference_draws(ref_lib, oracle_mod):
    """rate dt = 1/2 in both states: with the reference's own draw sequence,
    channel i toggles at a step exactly when its u1 < 1/2, and a toggle
    consumes the next draw before channel i + 1 draws (brute force over the
    raw stream, 40 steps of 8 channels)."""
    n, k = 8, 40
    sp = tiny(n, 2, sc_params(0.5, 0.5))
    o, _ = run_oracle(oracle_mod, sp, k, init_channel=3)
    for s in range(2):
        _, u = ref_state_after(ref_lib, s, n + 1 + 2 * k * n)
        pos = n + 1
        state = np.zeros(n, dtype=int)
        state[3] = 1
        toggles = 0
        for _ in range(k):
            for i in range(n):
                u1 = u[pos]
                pos += 1
                if u1 < 0.5:
                    pos += 1
                    state[i] ^= 1
                    toggles += 1
        want, _ = ref_state_after(ref_lib, s, pos)
        assert np.array_equal(o.rng_states()[s], want)
        assert np.array_equal(o.channel_states()[s, :n], state)
        assert 0 < toggles < k * n


def test_constant_rate_chain_statistics(oracle_mod):
    """10^4 two-state chains (2500 sites x 4 channels) at constant rates
    a = kopen dt = 0.02, b = kclose dt = 0.05:
      - P(open after m steps) = P_inf + (P_0 - P_inf)(1 - a - b)^m, P_inf = a/(a+b);
      - the time average over steps 150..400 is kopen / (kopen + kclose);
      - closed and open dwell times are geometric with p = a and p = b: the
        chance of leaving does not depend on the time spent in the state (chi^2)."""
    from scipy import stats
    a, b = 0.02, 0.05
    n_sites, n = 2500, 4
    sp = tiny(n, n_sites, sc_params(a, b))
    o, _ = run_oracle(oracle_mod, sp, 0)
    steps = 400
    hist = np.empty((steps + 1, n_sites, n), dtype=np.int8)
    hist[0] = o.channel_states()[:, :n]
    for m in range(1, steps + 1):
        o.step(DT, 1, threads=4)
        hist[m] = o.channel_states()[:, :n]
    p_inf = a / (a + b)
    frac = hist.reshape(steps + 1, -1).mean(axis=1)
    # transient from the initial open fraction 1/4
    for m in (5, 10, 20, 40):
        want = p_inf + (0.25 - p_inf) * (1 - a - b) ** m
        sd = math.sqrt(want * (1 - want) / (n_sites * n))
        assert abs(frac[m] - want) < 5 * sd, (m, frac[m], want)
    assert abs(frac[150:].mean() - p_inf) < 0.01
    # geometric dwell times = an age-independent hazard: a chain that has
    # spent m steps in state st leaves it at the next step with probability
    # p_st whatever m is (checked per age bin against Binomial(n_m, p_st);
    # ages count from a change inside the record, so no run is truncated)
    h = hist.reshape(steps + 1, -1).astype(np.int64)
    age = np.full(h.shape[1], -1)
    kmax = 60
    n_at = np.zeros((2, kmax + 1))
    x_at = np.zeros((2, kmax + 1))
    for t in range(steps):
        known = age >= 1
        leave = h[t + 1] != h[t]
        a_bin = np.minimum(age, kmax)
        for st in (0, 1):
            sel = known & (h[t] == st)
            np.add.at(n_at[st], a_bin[sel], 1)
            np.add.at(x_at[st], a_bin[sel], leave[sel])
        age = np.where(leave, 1, np.where(known, age + 1, -1))
    chi2, dof = 0.0, 0
    for st, p in ((0, a), (1, b)):
        n_m, x_m = n_at[st, 1:], x_at[st, 1:]
        ok = n_m * p >= 20
        assert ok.sum() >= 20
        chi2 += float((((x_m - n_m * p) ** 2) / (n_m * p * (1 - p)))[ok].sum())
        dof += int(ok.sum())
        # the pooled hazard is the rate itself
        assert abs(x_m.sum() / n_m.sum() - p) < 5 * math.sqrt(p * (1 - p) / n_m.sum())
    assert stats.chi2.sf(chi2, dof) > 1e-3, (chi2, dof)
    # a wrong rule (say, the hazard halving after the first step) is far outside
    assert x_at[0, 1] / n_at[0, 1] > 0.5 * a


def _cyto_pair(oracle_mod, leak, serca, dt):
    sp, j = config.load_json(os.path.join(GOLDEN, "test_input_Parameters.json"), n_sites=1)
    sp.init_channel = [0]
    ps = config.pools_from_json(j, sp)
    ps.cyto_leak_mult = leak
    ps.serca = serca
    h = oracle_mod.OracleHybridRun(sp, ps)
    h.reset()
    h.init_opening()
    return sp, ps, h


def _one_cyto_step(oracle_mod, leak, serca, ions, bound, dt=1e-3):
    sp, ps, h = _cyto_pair(oracle_mod, leak, serca, dt)
    h.set_cytosol(ions, bound)
    h.update()
    h.step(dt, 1, threads=1)
    return sp, ps, h.cytosol()


def _perturbed(oracle_mod):
    sp, ps, h = _cyto_pair(oracle_mod, 0.0, {}, 1e-3)
    ions, bound = h.cytosol()
    ions = ions * np.array([[3.0], [0.7]])
    bound = bound * (1.0 + 0.1 * np.arange(1, bound.shape[0] + 1))[:, None]
    return ions, bound, h


def test_cytosol_leak_one_step(oracle_mod):
    """Leak on minus leak off, one step from the same perturbed state:
    ions -dt L D_j (c_j - c0_j), bound fields -dt L D_b (B_b - B_b,init)
    (CytosolPool.cpp:252-263; onetau = "Leak mult" x D, :18,35,126)."""
    ions, bound, h0 = _perturbed(oracle_mod)
    L, dt = 2e-3, 1e-3
    sp, ps, (i1, b1) = _one_cyto_step(oracle_mod, L, {}, ions, bound, dt)
    _, _, (i0, b0) = _one_cyto_step(oracle_mod, 0.0, {}, ions, bound, dt)
    c0 = np.array([i.c0 for i in sp.ions])
    D = np.array([i.D for i in sp.ions])
    want_i = -dt * L * D * (ions[:, 0] - c0)
    got_i = i1[:, 0] - i0[:, 0]
    assert np.allclose(got_i, want_i, rtol=1e-7, atol=0), (got_i, want_i)
    # bound fields, ion-major over (dyad buffers, then cytosol buffers)
    init = h0.cytosol()[1][:, 0]
    Db = []
    for ion in sp.ions:
        for buf in sp.buffers + ps.cyto_buffers:
            if ion.name in buf.kinetics:
                Db.append(buf.kinetics[ion.name][0])
    Db = np.array(Db)
    assert len(Db) == bound.shape[0]
    want_b = -dt * L * Db * (bound[:, 0] - init)
    got_b = b1[:, 0] - b0[:, 0]
    mobile = Db != 0
    assert np.allclose(got_b[mobile], want_b[mobile], rtol=1e-6, atol=0)
    assert np.all(got_b[~mobile] == 0.0)
    assert mobile.sum() >= 3


def test_cytosol_serca_one_step_reads_background_by_serca_index(oracle_mod):
    """SERCA on minus SERCA off, one step: -dt Jmax (c^2/(c^2+K^2) -
    c_bg^2/(c_bg^2+K^2)) for the pumped ion only, where c_bg is
    ions_bg[idx_SERCA], the background of the idx_SERCA-th ion, not of the
    pumped ion (CytosolPool.cpp:240-250).  With the pump on Magnesium alone
    (ion 1, idx_SERCA 0) the background read is Calcium's 0.1 uM."""
    ions, bound, _ = _perturbed(oracle_mod)
    dt, jmax, K = 1e-3, 0.3, 150.0
    sp, _, (i1, _) = _one_cyto_step(oracle_mod, 0.0, {"Magnesium": (jmax, K)}, ions, bound, dt)
    _, _, (i0, _) = _one_cyto_step(oracle_mod, 0.0, {}, ions, bound, dt)
    assert [i.name for i in sp.ions] == ["Calcium", "Magnesium"]
    c = ions[1, 0]
    bg = sp.ions[0].c0  # Calcium's background
    want = -dt * jmax * (c * c / (c * c + K * K) - bg * bg / (bg * bg + K * K))
    got = i1[:, 0] - i0[:, 0]
    assert got[0] == 0.0
    assert math.isclose(got[1], want, rel_tol=1e-8), (got, want)
    wrong = -dt * jmax * (c * c / (c * c + K * K) - 1000.0 ** 2 / (1000.0 ** 2 + K * K))
    assert not math.isclose(got[1], wrong, rel_tol=1e-3)
    # both ions pumped: the counter reaches 1 for Magnesium, whose own background it reads
    sp, _, (i2, _) = _one_cyto_step(oracle_mod, 0.0, {"Calcium": (0.2, 0.2), "Magnesium": (jmax, K)}, ions, bound,
                                    dt)
    mg_bg = sp.ions[1].c0
    want_mg = -dt * jmax * (c * c / (c * c + K * K) - mg_bg ** 2 / (mg_bg ** 2 + K * K))
    ca = ions[0, 0]
    want_ca = -dt * 0.2 * (ca * ca / (ca * ca + 0.04) - 0.01 / (0.01 + 0.04))
    got2 = i2[:, 0] - i0[:, 0]
    assert math.isclose(got2[1], want_mg, rel_tol=1e-8)
    assert math.isclose(got2[0], want_ca, rel_tol=1e-8)
