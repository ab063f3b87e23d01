"""HIP path (through the C ABI) against the oracle on the same seeded inputs.

Bars: RNG states and channel states bit-exact; every grid node, mouth value,
site current and efflux within 1e-9 relative in FP64 and 1e-4 relative in
FP32 (channels clamped for FP32, where rounding of the mouth Ca could
legitimately flip a stochastic decision)."""
import os

import numpy as np
import pytest

from casim_amd import abi, config, workloads
from casim_amd.dyad import DyadHIP
from tests.conftest import GOLDEN
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
FP32_TOL = 1e-4


def say(*a):
    print(*a, flush=True)


def setup_pair(oracle_mod, sp, jsr=1000.0):
    ora = oracle_mod.Oracle(sp)
    dev = DyadHIP(sp)
    ora.init_opening()
    dev.InitOpeningAll()
    ri, rb = sp.rim_arrays()
    j = sp.jsr_array([jsr] + [0.0] * (sp.n_ions - 1))
    ora.update(ri, rb, j)
    dev.Update(ri, rb, j)
    return ora, dev, (ri, rb, j)


def compare(ora, dev, sp, tol, fields=True):
    nf = sp.n_ions + sp.n_bound
    worst = 0.0
    if fields:
        for f in range(nf):
            e = rel_err(dev.GetField(f), ora.field(f))
            worst = max(worst, e)
            assert e < tol, ("field", sp.field_names()[f], e)
    m_dev, m_ora = dev.GetIonsNearChannels(), ora.mouths()
    st = ora.channel_states()
    mask = np.broadcast_to(st >= 0, m_ora.shape)
    assert rel_err(m_dev[mask], m_ora[mask]) < tol
    assert rel_err(dev.GetTotalSRCurrent(), ora.total_sr_current()) < tol
    eo, ed = ora.effluxes(), dev.GetEffluxes()
    # The efflux sums one-sided differences 2u1 - 1.5u0 - 0.5u2 of near-equal
    # rim values; its error is judged against the size of those terms.
    scale = np.maximum(np.abs(eo), efflux_term_scale(ora, sp))
    assert float(np.max(np.abs(ed - eo) / scale)) < max(tol, 1e-8)
    return worst


def efflux_term_scale(ora, sp):
    """[site, field]: D gm s * sum over the efflux stencil of (2|u1| + 1.5|u0| + 0.5|u2|)."""
    gm = 3.0 / (4 * np.pi * sp.radius ** 3)
    if sp.dim == 3:
        sx, sy, npz = sp.dy * sp.dz / sp.dx, sp.dx * sp.dz / sp.dy, sp.nz - 1
    else:
        sx, sy, npz = sp.dy * sp.z / sp.dx, sp.dx * sp.z / sp.dy, 1
    Ds = [i.D for i in sp.ions] + [t[2] for t in sp.bound_fields()]
    out = np.zeros((sp.n_sites, len(Ds)))
    for f, D in enumerate(Ds):
        u = np.abs(ora.field(f))
        m = u.reshape(sp.n_sites, -1).max(axis=1)
        out[:, f] = D * gm * 4 * m * (sx * 2 * (sp.ny - 2) * npz + sy * 2 * sp.nx * npz)
    return np.maximum(out, 1e-300)


def test_init_opening_rng_bit_exact(gpu, oracle_mod):
    sp = workloads.single_dyad(8)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    assert np.array_equal(dev.GetRngStates(), ora.rng_states())
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())
    assert dev.GetChannelsStates()[:, 12].sum() == 8


@pytest.mark.parametrize("precision,tol", [(abi.CASIM_FP64, FP64_TOL), (abi.CASIM_FP32, FP32_TOL)])
def test_clamped_single_dyad(gpu, oracle_mod, precision, tol):
    say("clamped single dyad, precision", precision)
    sp = workloads.single_dyad(4, clamped=True, precision=precision)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt, n = 1e-5, 1500
    dev.Step(dt, n)
    ora.step(dt, n)
    worst = compare(ora, dev, sp, tol)
    say("worst rel err", worst)
    assert np.array_equal(dev.GetRngStates(), ora.rng_states())
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())


def test_stochastic_trajectories_fp64(gpu, oracle_mod):
    """Normal gating: the same channel histories and RNG states, step by step."""
    sp = workloads.single_dyad(8, clamped=False)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = 1e-5
    for block in range(6):
        dev.Step(dt, 250)
        ora.step(dt, 250)
        assert np.array_equal(dev.GetChannelsStates(), ora.channel_states()), block
        assert np.array_equal(dev.GetRngStates(), ora.rng_states()), block
    compare(ora, dev, sp, FP64_TOL)


def test_fast_gating_many_transitions(gpu, oracle_mod):
    """Rates scaled up (kopen x1e6, kclose x1e4: ~500 toggles over 400 steps
    of 16 sites, several per site-step late in the run) so that every branch
    of the ballot scan runs, including repeated re-evaluation in one step."""
    sp = workloads.single_dyad(16, clamped=False)
    sp.channel_types[0].params["kopen0"] *= 1e6
    sp.channel_types[0].params["kclose0"] *= 1e4
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = 1e-5
    prev = ora.channel_states().copy()
    changes = 0
    for block in range(8):
        dev.Step(dt, 50)
        ora.step(dt, 50)
        st = ora.channel_states()
        assert np.array_equal(dev.GetChannelsStates(), st), block
        assert np.array_equal(dev.GetRngStates(), ora.rng_states()), block
        changes += int((st != prev).sum())
        prev = st.copy()
    assert changes > 200
    compare(ora, dev, sp, FP64_TOL)


def test_protocol_path_equals_fused_path(gpu, oracle_mod):
    """Update / RunRD / RunMC one call at a time == the graph-captured fused step."""
    sp = workloads.single_dyad(3, clamped=False)
    _, a, (ri, rb, j) = setup_pair(oracle_mod, sp)
    b = DyadHIP(sp)
    b.InitOpeningAll()
    b.Update(ri, rb, j)
    dt = 1e-5
    for _ in range(120):
        a.RunRD(dt)
        a.RunMC(dt)
        a.Update(ri, rb, j)
    b.Step(dt, 120)
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(a.GetField(f), b.GetField(f))
    assert np.array_equal(a.GetChannelsStates(), b.GetChannelsStates())
    assert np.array_equal(a.GetRngStates(), b.GetRngStates())
    assert np.array_equal(a.GetIonsNearChannels(), b.GetIonsNearChannels())
    assert np.array_equal(a.GetTotalSRCurrent(), b.GetTotalSRCurrent())
    assert np.array_equal(a.GetEffluxes(), b.GetEffluxes())


def test_graph_replays_compose(gpu, oracle_mod):
    sp = workloads.single_dyad(2, clamped=False)
    _, a, (ri, rb, j) = setup_pair(oracle_mod, sp)
    b = DyadHIP(sp)
    b.InitOpeningAll()
    b.Update(ri, rb, j)
    for _ in range(4):
        a.Step(1e-5, 25)
    b.Step(1e-5, 100)
    assert np.array_equal(a.GetField(0), b.GetField(0))
    assert np.array_equal(a.GetRngStates(), b.GetRngStates())


def test_heterogeneous_batch(gpu, oracle_mod):
    """Sites with 8..100 RyRs in random clusters, one batch, stochastic."""
    sp = workloads.batched_dyads(n_sites=6, geometry_seed=7)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = 1e-5
    dev.Step(dt, 400)
    ora.step(dt, 400)
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())
    assert np.array_equal(dev.GetRngStates(), ora.rng_states())
    compare(ora, dev, sp, FP64_TOL)


def test_reference_test_input_2d(gpu, oracle_mod):
    """The reference's own test configuration (CPU2D, Ca + Mg, ATP binding
    both, calmodulin, Fluo-4, sarcolemmal buffer, 20 RyRs, SheepCannell)."""
    sp, j = config.load_json(os.path.join(GOLDEN, "test_input_Parameters.json"), n_sites=3)
    sp.init_channel = [0, 5, 11]
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = float(j["dt, ns"]) * 1e-6
    dev.Step(dt, 300)
    ora.step(dt, 300)
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())
    assert np.array_equal(dev.GetRngStates(), ora.rng_states())
    compare(ora, dev, sp, FP64_TOL)


def test_full_size_batch_sites_match_single_site_oracles(gpu, oracle_mod):
    """4096 heterogeneous sites on the device; sites are independent, so any
    site equals a one-site oracle run with that site's channels and stream."""
    say("full-size batch: 4096 sites")
    sp = workloads.batched_dyads(n_sites=4096)
    dev = DyadHIP(sp)
    dev.InitOpeningAll()
    ri, rb = sp.rim_arrays()
    jj = sp.jsr_array([1000.0])
    dev.Update(ri, rb, jj)
    dt, n = 1e-5, 60
    dev.Step(dt, n)
    ca = dev.GetField(0)
    assert np.isfinite(ca).all()
    st = dev.GetChannelsStates()
    for s in (0, 1234, 4095):
        one = workloads.batched_dyads(n_sites=1)
        one.sites = [sp.sites[s]]
        one.rng_states = sp.rng_states[s:s + 1]
        one.init_channel = [sp.init_channel[s]]
        ora = oracle_mod.Oracle(one)
        ora.init_opening()
        r1, b1 = one.rim_arrays()
        ora.update(r1, b1, one.jsr_array([1000.0]))
        ora.step(dt, n, threads=1)
        assert rel_err(ca[s], ora.field(0)[0]) < FP64_TOL
        n_ch = len(sp.sites[s])
        assert np.array_equal(st[s, :n_ch], ora.channel_states()[0, :n_ch])


@pytest.mark.parametrize("planes,threads,prefetch", [(-1, 256, 0), (0, 128, 1), (0, 448, 1), (10, 256, 2),
                                                     (20, 192, 1), (0, 448, 3), (20, 256, 3)])
def test_sweep_tunings_match_oracle(gpu, oracle_mod, planes, threads, prefetch):
    """Every launch shape of the state sweep gives the oracle's results."""
    sp = workloads.batched_dyads(n_sites=3, geometry_seed=11)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dev.SetTuning(planes, threads, prefetch)
    dt = 1e-5
    dev.Step(dt, 150)
    ora.step(dt, 150)
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())
    compare(ora, dev, sp, FP64_TOL)


def test_sweep_2d_slots(gpu, oracle_mod):
    """2D rows longer than one block: several plane slots per thread."""
    from tests.helpers import spec
    sp = spec(dim=2, nx=40, ny=150, h=2.0, clamped=False, channels=[(40.0, 100.0), (46.0, 100.0)], n_sites=2)
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dev.SetTuning(0, 64)  # 150 nodes per row over 64 threads: 3 slots
    dev.Step(2e-6, 200)  # D dt (2/h^2 + 2/h^2) = 0.44: inside the Euler limit
    ora.step(2e-6, 200)
    compare(ora, dev, sp, FP64_TOL)


def test_release_stats_match_oracle_trajectory(gpu, oracle_mod):
    """On-device release statistics equal the IndChannelCREAnalyzer columns
    accumulated from the oracle's channel states, step by step."""
    from casim_amd.analysis import StatsTracker
    sp = workloads.single_dyad(12, clamped=False)
    sp.channel_types[0].params["kopen0"] *= 1e5
    sp.channel_types[0].params["kclose0"] *= 1e3
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = 1e-5
    tr = StatsTracker(ora.channel_states(), dt)
    for _ in range(200):
        ora.step(dt, 1, threads=4)
        tr.step(ora.channel_states())
    dev.Step(dt, 200)
    got = dev.GetReleaseStats()
    want = tr.table()
    assert np.array_equal(got[:, [0, 5]], want[:, [0, 5]])
    assert np.allclose(got[:, 1:5], want[:, 1:5], rtol=1e-12, atol=1e-15)
    assert want[:, 0].max() > 1


def test_open_count_trace_matches_oracle(gpu, oracle_mod):
    """Per-bin open counts (the analyzer's time-histogram input): bin 0 after
    InitOpening, bin k the count after step k*B - 1, through paired passes."""
    sp = workloads.single_dyad(6, clamped=False)
    sp.channel_types[0].params["kopen0"] *= 1e5
    sp.channel_types[0].params["kclose0"] *= 1e3
    ora = oracle_mod.Oracle(sp)
    dev = DyadHIP(sp)
    B, bins = 7, 40
    dev.RecordOpenCounts(B, bins)
    ora.init_opening()
    dev.InitOpeningAll()
    ri, rb = sp.rim_arrays()
    j = sp.jsr_array([1000.0])
    ora.update(ri, rb, j)
    dev.Update(ri, rb, j)
    dt = 1e-5
    want = np.full((bins, sp.n_sites), 255, dtype=np.uint8)
    want[0] = (ora.channel_states() == 1).sum(axis=1)
    for m in range(1, 200):
        ora.step(dt, 1, threads=4)
        if (m + 1) % B == 0 and (m + 1) // B < bins:
            want[(m + 1) // B] = (ora.channel_states() == 1).sum(axis=1)
    dev.Step(dt, 199)
    got = dev.GetOpenCounts()
    assert np.array_equal(got, want)
    assert len(np.unique(want[want != 255])) > 2


def _pool_pair(oracle_mod, sp, ps, rim=None):
    ora = oracle_mod.OracleHybridRun(sp, ps, rim=rim)
    ora.reset()
    ora.init_opening()
    ora.update()
    dev = DyadHIP(sp)
    dev.AttachPools(ps)
    dev.InitOpeningAll()
    ri, rb = rim if rim is not None else sp.rim_arrays()
    dev.UpdatePools(None if ps.cytosol_pool else ri, None if ps.cytosol_pool else rb)
    return ora, dev


def test_pools_jsr_depletion_fp64(gpu, oracle_mod):
    """Depleting jSR (calsequestrin-buffered SRPool) with a constant rim."""
    from casim_amd.config import PoolSpec
    sp = workloads.single_dyad(4, clamped=False)
    sp.channel_types[0].params["kopen0"] *= 1e4
    ps = PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=2e5, jsr_c0=[1000.0], jsr_cq_sites=[1.0],
                  jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])
    ora, dev = _pool_pair(oracle_mod, sp, ps)
    dt = 1e-5
    for _ in range(4):
        dev.Step(dt, 150)
        ora.step(dt, 150)
        assert np.array_equal(dev.GetChannelsStates(), ora.dyad.channel_states())
    j_dev = dev.GetPools()[0]
    assert rel_err(j_dev, ora.jsr()) < FP64_TOL
    assert ora.jsr().min() < 999.0
    compare(ora.dyad, dev, sp, FP64_TOL)
    assert np.array_equal(dev.GetRngStates(), ora.dyad.rng_states())


def test_two_step_passes_with_jsr_pool(gpu):
    """A depleting jSR pool (no cytosol pool) keeps paired passes, bit for bit
    equal to one step per pass: fields, jSR, channels, RNG, currents."""
    from casim_amd.config import PoolSpec
    sp = workloads.batched_dyads(n_sites=8, geometry_seed=4)
    sp.channel_types[0].params["kopen0"] *= 1e5
    ps = PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=2e5, jsr_c0=[1000.0], jsr_cq_sites=[1.0],
                  jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])
    ri, rb = sp.rim_arrays()
    devs = []
    for steps in (1, 2):
        d = DyadHIP(sp)
        d.AttachPools(ps)
        d.SetStepsPerPass(steps)
        d.InitOpeningAll()
        d.UpdatePools(ri, rb)
        for k in (9, 40, 1, 50):
            d.Step(1e-5, k)
        devs.append(d)
    one, two = devs
    assert one.StepsPerPass() == 1 and two.StepsPerPass() == 2
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(one.GetField(f), two.GetField(f)), f
    j1, j2 = one.GetPools()[0], two.GetPools()[0]
    assert np.array_equal(j1, j2) and j1.min() < 999.9
    assert np.array_equal(one.GetChannelsStates(), two.GetChannelsStates())
    assert np.array_equal(one.GetRngStates(), two.GetRngStates())
    assert np.array_equal(one.GetTotalSRCurrent(), two.GetTotalSRCurrent())
    assert np.array_equal(one.GetIonsNearChannels(), two.GetIonsNearChannels())


def test_pools_reference_test_input_cytosol(gpu, oracle_mod):
    """The reference's test configuration end to end: CytosolPool (leak,
    SERCA, troponin), SRPool with calsequestrin, constant NSR, 2D dyad."""
    sp, j = config.load_json(os.path.join(GOLDEN, "test_input_Parameters.json"), n_sites=3)
    sp.init_channel = [2, 9, 16]
    ps = config.pools_from_json(j, sp)
    ora, dev = _pool_pair(oracle_mod, sp, ps)
    dt = float(j["dt, ns"]) * 1e-6
    for _ in range(3):
        dev.Step(dt, 400)
        ora.step(dt, 400)
        assert np.array_equal(dev.GetChannelsStates(), ora.dyad.channel_states())
    jd, cid, cbd = dev.GetPools()
    cio, cbo = ora.cytosol()
    assert rel_err(jd, ora.jsr()) < FP64_TOL
    assert rel_err(cid, cio) < FP64_TOL
    assert rel_err(cbd, cbo) < FP64_TOL
    assert ora.jsr()[0].min() < 1000.0 and cio[0].max() > 0.1
    compare(ora.dyad, dev, sp, FP64_TOL)


@pytest.mark.parametrize("precision,planes,threads", [(abi.CASIM_FP64, 0, 0), (abi.CASIM_FP64, 10, 256),
                                                      (abi.CASIM_FP32, 0, 448), (abi.CASIM_FP32, 20, 192)])
def test_two_step_passes_equal_single_steps(gpu, precision, planes, threads):
    """Pairs of steps in one state pass give the single-step results bit for
    bit: fields, channel histories, RNG, mouth snapshots, currents, effluxes
    and release statistics, with fast gating so that channels toggle inside
    passes, over odd and even step counts."""
    say("two-step passes", precision, planes, threads)
    sp = workloads.batched_dyads(n_sites=12, geometry_seed=9, precision=precision)
    sp.channel_types[0].params["kopen0"] *= 1e6
    sp.channel_types[0].params["kclose0"] *= 1e4
    ri, rb = sp.rim_arrays()
    j = sp.jsr_array([1000.0])
    devs = []
    for steps in (1, 2):
        d = DyadHIP(sp)
        if planes or threads:
            d.SetTuning(planes, threads, 1)
        d.SetStepsPerPass(steps)
        d.InitOpeningAll()
        d.Update(ri, rb, j)
        for k in (7, 50, 1, 64):
            d.Step(1e-5, k)
        devs.append(d)
    one, two = devs
    assert one.StepsPerPass() == 1 and two.StepsPerPass() == 2
    for f in range(sp.n_ions + sp.n_bound):
        assert np.array_equal(one.GetField(f), two.GetField(f)), f
    st = one.GetChannelsStates()
    assert np.array_equal(st, two.GetChannelsStates())
    assert (st != workloads.batched_dyads(n_sites=12, geometry_seed=9).init_channel[0]).any()
    assert np.array_equal(one.GetRngStates(), two.GetRngStates())
    assert np.array_equal(one.GetIonsNearChannels(), two.GetIonsNearChannels())
    assert np.array_equal(one.GetTotalSRCurrent(), two.GetTotalSRCurrent())
    assert np.array_equal(one.GetEffluxes(), two.GetEffluxes())
    assert np.array_equal(one.GetReleaseStats(), two.GetReleaseStats())


def test_pools_run_one_step_per_pass(gpu):
    sp, j = config.load_json(os.path.join(GOLDEN, "test_input_Parameters.json"), n_sites=2)
    dev = DyadHIP(sp)
    assert dev.StepsPerPass() == 2
    dev.AttachPools(config.pools_from_json(j, sp))
    assert dev.StepsPerPass() == 1


def test_channel_count_edges(gpu, oracle_mod):
    """Sites with 1, 63, 64, 65 and 127 RyRs (the channel kernel works in
    chunks of 64 lanes; 127 is the per-site maximum), stochastic FP64."""
    sp = workloads.batched_dyads(n_sites=5, geometry_seed=3, extent_nm=440.0)
    c = 220.0
    cells = [(c + 31.0 * (i - 5.5), c + 31.0 * (j - 5.5), 0) for i in range(12) for j in range(12)]
    # nearest the centre first, so every site keeps a compact cluster
    cells.sort(key=lambda p: ((p[0] - c) ** 2 + (p[1] - c) ** 2, p[0], p[1]))
    sp.sites = [cells[:n] for n in (1, 63, 64, 65, 127)]
    ora, dev, _ = setup_pair(oracle_mod, sp)
    dt = 1e-5
    dev.Step(dt, 300)
    ora.step(dt, 300)
    assert np.array_equal(dev.GetChannelsStates(), ora.channel_states())
    assert np.array_equal(dev.GetRngStates(), ora.rng_states())
    compare(ora, dev, sp, FP64_TOL)


@pytest.mark.parametrize("n_sites,k", [(5, 40), (64, 41)])
def test_split_pipelining_equals_one_stream(gpu, n_sites, k):
    """Step's split pipelining (two half-batches on two streams, the second one
    channel call behind) gives the one-stream results bit for bit: fields,
    faces, channel and RNG states, mouths, currents, release statistics and
    open-count bins; an odd step count leaves one whole-batch step."""
    sp = workloads.batched_dyads(n_sites=n_sites, geometry_seed=13)
    out = []
    for split in (0, 1):
        dev = DyadHIP(sp)
        dev.SetSplit(split)
        dev.RecordOpenCounts(4, 30)
        dev.InitOpeningAll()
        ri, rb = sp.rim_arrays()
        dev.Update(ri, rb, sp.jsr_array([1000.0]))
        for _ in range(2):  # two graph launches compose
            dev.Step(1e-5, k)
        out.append([dev.GetField(f) for f in range(sp.n_ions + sp.n_bound)] +
                   [dev.GetChannelsStates(), dev.GetRngStates(), dev.GetIonsNearChannels(),
                    dev.GetTotalSRCurrent(), dev.GetReleaseStats(), dev.GetOpenCounts()])
        dev.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert (out[1][sp.n_ions + sp.n_bound] == 1).sum() > 0  # channels open: the runs are not trivial


def test_split_pipelining_with_a_jsr_pool(gpu):
    """The same with a depleting jSR pool (per-site pool state on both streams)."""
    from casim_amd import spark_compare
    sp = workloads.batched_dyads(n_sites=33, geometry_seed=17)
    out = []
    for split in (0, 1):
        dev = DyadHIP(sp)
        dev.SetSplit(split)
        dev.AttachPools(spark_compare.depleting_pool())
        dev.InitOpeningAll()
        dev.UpdatePools(*sp.rim_arrays())
        dev.Step(1e-5, 60)
        out.append([dev.GetField(0), dev.GetChannelsStates(), dev.GetRngStates(), dev.GetPools()[0][0],
                    dev.GetReleaseStats()])
        dev.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
