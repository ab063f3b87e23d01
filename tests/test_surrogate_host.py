"""Host-side pieces of the surrogate path: features, the linear-superposition
baseline, clamping and launch shapes (no GPU)."""
import math

import numpy as np

from casim_amd import linear_baseline, surrogate, workloads


def test_features_of_three_in_a_row():
    # RyRs at lattice x = -1, 0, 1 (y = 0) around the centre; the outer two open
    half = 160.0
    xy = np.array([[half - 31.0, half], [half, half], [half + 31.0, half]])
    f = surrogate.site_features(xy, [1, 0, 1], 800.0, half)
    assert f.shape == (3, surrogate.N_FEAT)
    assert list(f[:, 0]) == [1, 0, 1]
    assert np.allclose(f[:, 1], 0.8) and np.allclose(f[:, 2], 0.02)
    # middle: two open neighbours at r = 1; ends: one open partner at r = 2
    assert math.isclose(f[1, 3], 2.0) and math.isclose(f[0, 3], 0.5) and math.isclose(f[2, 3], 0.5)
    assert math.isclose(f[1, 4], 2 * math.exp(-1.0)) and math.isclose(f[0, 4], math.exp(-2.0))
    # nearest open partner of the left end: (+2, 0), 1/r = 0.5; no second partner
    assert list(f[0, 8:11]) == [2.0, 0.0, 0.5] and list(f[0, 11:14]) == [0.0, 0.0, 0.0]
    # middle: both at r = 1, the lower index first (stable order)
    assert list(f[1, 8:14]) == [-1.0, 0.0, 1.0, 1.0, 0.0, 1.0]
    assert np.allclose(f[:, 5], [-1, 0, 1]) and np.allclose(f[:, 6], 0)
    assert np.allclose(f[:, 7], 160.0 / 31.0 - np.array([1, 0, 1]))


def test_features_with_nothing_open():
    xy = np.array([[100.0, 100.0], [131.0, 100.0]])
    f = surrogate.site_features(xy, [0, 0], 500.0, 160.0)
    assert np.all(f[:, 3:5] == 0) and np.all(f[:, 8:] == 0) and np.all(f[:, 0] == 0)


def test_linear_superposition_sums_table_entries():
    cal = {"r2": np.array([0, 1, 2, 4]), "added": np.array([300.0, 5.0, 3.0, 2.0]), "background": 0.1,
           "jsr": 1000.0}
    xy = np.array([[0.0, 0.0], [31.0, 0.0], [62.0, 0.0], [31.0, 31.0]])
    open_mask = [True, False, True, False]
    got = linear_baseline.predict(cal, xy, open_mask, 500.0)
    # r^2 to the open pair (0 and 2): ch0 {0, 4}, ch1 {1, 1}, ch2 {4, 0}, ch3 {2, 2}
    want = 0.1 + 0.5 * np.array([300 + 2, 5 + 5, 2 + 300, 3 + 3])
    assert np.allclose(got, want)
    # beyond the table the last entry decays as 1/r
    far = linear_baseline.predict(cal, np.array([[0.0, 0.0], [31.0 * 4, 0.0]]), [True, False], 1000.0)
    assert math.isclose(far[1], 0.1 + 2.0 * 2.0 / 4.0)


def test_clamped_spec_zeroes_rates_only():
    sp = workloads.batched_dyads(n_sites=2)
    c = surrogate.clamped_spec(sp)
    p0, p1 = sp.channel_types[0].params, c.channel_types[0].params
    assert p1["kopen0"] == 0.0 and p1["kclose0"] == 0.0 and p0["kopen0"] > 0
    assert p1["ncaopen"] == p0["ncaopen"] and c.sites is sp.sites


def test_small_batches_get_plane_chunks():
    assert surrogate.small_batch_tuning(workloads.batched_dyads(n_sites=4096)) is None
    planes, _, _ = surrogate.small_batch_tuning(workloads.batched_dyads(n_sites=16))
    assert planes == 2  # 64 chunks of 79 interior planes


def _qs_config(n_sites=2, **over):
    """A valid quasi-stationary configuration (arrays kept alive in the result)."""
    import ctypes as C
    from casim_amd import abi, rng
    off = np.array([0, 3, 5][:n_sites + 1], dtype=np.int32)
    xs = np.array([100.0, 131.0, 162.0, 100.0, 131.0])[:off[-1]]
    ys = np.full(off[-1], 160.0)
    par = np.zeros(abi.CASIM_MAX_MODEL_PARAMS)
    par[:5] = [888888.9, 2.0, 1e300, 500.0, 0.0]
    jsr = np.full(n_sites, 1000.0)
    rs = np.ascontiguousarray(rng.stream_states(rng.DEFAULT_SEED, n_sites), dtype=np.uint64)
    init = np.zeros(n_sites, dtype=np.int32)
    keep = dict(off=off, xs=xs, ys=ys, par=par, jsr=jsr, rs=rs, init=init)
    vals = dict(n_sites=n_sites, site_channel_offset=abi.iptr(off), channel_x=abi.dptr(xs), channel_y=abi.dptr(ys),
                half_extent=160.0, pitch=31.0, model=1, params=abi.dptr(par), mg=0.0, jsr=abi.dptr(jsr),
                rng_state=abi.uptr(rs), init_channel=abi.iptr(init), t_max=1.0, spread_threshold=0.02,
                jsr_deplete=0, jsr_volume=0.0, element_volume=32.0, conductance=1.0, jsr_cq_total=0.0,
                jsr_cq_sites=0.0, jsr_k_cq=0.0, jsr_tau_refill=1.0, nsr=0.0)
    vals.update(over)
    return surrogate.CasimQSConfig(**vals), keep


def test_quasi_stationary_configs_are_validated_on_the_host():
    import ctypes as C
    from casim_amd import abi, mlp
    lib = mlp.lib()

    def create(cfg):
        h = C.c_void_p()
        rc = lib.casim_qs_create(C.byref(cfg), C.byref(h))
        return rc, lib.casim_last_error().decode()

    cfg, keep = _qs_config(model=7)
    assert create(cfg) == (abi.CASIM_E_INVALID, "unknown kinetic model")
    cfg, keep = _qs_config(pitch=0.0)
    assert create(cfg)[0] == abi.CASIM_E_INVALID
    init = np.array([0, 2], dtype=np.int32)  # site 1 has 2 channels
    cfg, keep = _qs_config(init_channel=abi.iptr(init))
    assert create(cfg) == (abi.CASIM_E_RANGE, "init channel out of range")
    bad = np.array(keep_rs := keep["rs"], dtype=np.uint64)
    bad[1, 3:] = 0
    cfg, keep = _qs_config(rng_state=abi.uptr(bad))
    assert "zero component" in create(cfg)[1]
    cfg, keep = _qs_config(jsr_deplete=1, jsr_volume=0.0)
    assert create(cfg) == (abi.CASIM_E_INVALID, "bad jSR pool parameters")
    assert lib.casim_qs_create(None, None) == abi.CASIM_E_INVALID
    del keep_rs


def test_superposition_feature_table_rules():
    lin = surrogate.LinearFeature(np.array([300.0, 5.0, 3.0, 2.0, 1.0]), 0.1, 1000.0)
    assert np.allclose(lin.at([0, 1, 1.5, 4, 16]), [300.0, 5.0, 4.0, 1.0, 0.5])  # 16: 1 * sqrt(4/16)
    xy = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 2.0]])
    est = lin.estimate(xy, [True, False, True], 500.0)
    # open: 0 and 2; r^2 0-2 = 4, 1-0 = 1, 1-2 = 5 (beyond: 1 * sqrt(4/5))
    want = 0.1 + 0.5 * np.array([300.0 + 1.0, 5.0 + np.sqrt(4 / 5), 1.0 + 300.0])
    assert np.allclose(est, want)
    f = surrogate.site_features(xy * surrogate.PITCH_NM + 160.0, [1, 0, 1], 500.0, 160.0, lin)
    assert f.shape == (3, surrogate.N_FEAT + 1) and np.allclose(f[:, -1], np.log(want))
