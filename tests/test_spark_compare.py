"""The host-side statistics that gate the surrogate path against the plain
path (casim_amd/spark_compare.py), on synthetic release records: equal
populations pass every test, a shifted one fails the test of the shifted
statistic, and the lowest-jSR comparison separates the held-open atom from
the other sites (IndChannelCREAnalyzer.cpp:53-138 defines the statistics)."""
import numpy as np

from casim_amd import spark_compare


def _records(rng, n, t_max=1.0, dt=1e-5):
    n_ch = rng.integers(8, 101, n)
    amp = rng.integers(1, 6, n).astype(float)
    last = np.round(rng.uniform(0, t_max, n) / dt) * dt
    integ = np.round(rng.uniform(0, 0.5, n) * amp / dt) * dt
    n_end = rng.integers(0, 3, n).astype(float)
    plain = np.column_stack([amp, np.zeros(n), last, integ, np.zeros(n), n_end])
    win = np.column_stack([amp, last, integ, n_end])
    return plain, win, n_ch


def test_equal_populations_pass_every_statistic():
    rng = np.random.default_rng(5)
    plain, win, n_ch = _records(rng, 2000)
    out = spark_compare.compare(plain, win, n_ch, 1.0)
    for name in ("amplitude", "last_event_ms", "open_integral", "n_open_end", "open_probability"):
        assert out[name]["p"] > 0.99, name
    assert out["all_pass"]


def test_shifted_open_integral_fails_integral_and_open_probability():
    rng = np.random.default_rng(6)
    plain, win, n_ch = _records(rng, 2000)
    win = win.copy()
    win[:, 2] *= 1.3
    out = spark_compare.compare(plain, win, n_ch, 1.0)
    assert out["open_integral"]["p"] < 1e-6
    assert out["open_probability"]["p"] < 1e-6
    assert out["amplitude"]["p"] > 0.99
    assert not out["all_pass"]


def test_open_probability_is_integral_over_channels_and_time():
    rng = np.random.default_rng(7)
    plain, win, n_ch = _records(rng, 500)
    # the same integrals on sites with twice the channels halve P_o: the
    # per-site normalisation must use each site's own channel count
    out = spark_compare.compare(plain, win, n_ch, 1.0)
    assert out["open_probability"]["mean_plain"] == np.mean(spark_compare.on_grid(plain[:, 3], 1e-5) / n_ch)


def _held_open_records(rng, n_held, n_other):
    plain = np.zeros((n_held + n_other, 6))
    win = np.zeros((n_held + n_other, 4))
    plain[:n_held, 0], plain[:n_held, 5] = 1, 1
    win[:n_held, 0], win[:n_held, 3] = 1, 1
    plain[n_held:, 0] = win[n_held:, 0] = rng.integers(2, 6, n_other)
    plain[n_held:, 2] = win[n_held:, 1] = rng.uniform(0.1, 1.0, n_other)
    return plain, win


def test_jsr_lowest_compares_the_atom_by_parts():
    rng = np.random.default_rng(8)
    plain, win = _held_open_records(rng, 600, 1400)
    others = rng.uniform(900.0, 990.0, 1400)
    lp = np.concatenate([np.full(600, 965.898711), others])
    lq = np.concatenate([np.full(600, 965.897759), rng.permutation(others)])
    jr = spark_compare.compare_jsr_lowest(plain, win, lp, lq)
    assert jr["held_open_plain"] == jr["held_open_qs"] == 600
    assert jr["held_open_spread_uM"] == 0.0
    assert jr["held_open_drop_rel_diff"] < 1e-4
    assert jr["held_open_weight_p"] > 0.99
    assert jr["others"]["p"] > 0.99
    # the whole-population KS sees the two atoms as a jump of the atom's weight
    assert jr["all_sites_ks"]["D"] > 0.25


def test_jsr_lowest_catches_shifted_other_sites():
    rng = np.random.default_rng(9)
    plain, win = _held_open_records(rng, 600, 1400)
    others = rng.uniform(900.0, 990.0, 1400)
    lp = np.concatenate([np.full(600, 965.9), others])
    lq = np.concatenate([np.full(600, 965.9), others - 5.0])
    jr = spark_compare.compare_jsr_lowest(plain, win, lp, lq)
    assert jr["others"]["p"] < 0.01
