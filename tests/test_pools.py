"""Pins of the oracle's compartments around the dyad (SRPool, NSRConstPool,
CytosolPool) and of the reference's batch-loop order."""
import math
import os

import numpy as np

from casim_amd import config
from casim_amd.config import PoolSpec
from tests.conftest import GOLDEN
from tests.helpers import spec


def pools(**kw):
    base = dict(jsr_constant=False, jsr_cq_total=800.0, jsr_volume=1e7, jsr_c0=[1000.0], jsr_cq_sites=[15.0],
                jsr_k_cq=[600.0], jsr_tau_refill=[15.0], nsr_c=[1000.0])
    base.update(kw)
    return PoolSpec(**base)


def test_jsr_refill_is_geometric(oracle_mod):
    """No current and no calsequestrin (beta = 1): c_n = nsr + (c0 - nsr)(1 - dt/tau)^n."""
    sp = spec(g=0.0)
    ps = pools(jsr_c0=[500.0], jsr_cq_sites=[0.0])
    h = oracle_mod.OracleHybridRun(sp, ps)
    h.reset()
    h.init_opening([0])
    h.update()
    dt, n = 1e-3, 50
    h.step(dt, n, threads=1)
    # SRPool's ions_NSR starts at the jSR value and takes the NSR value at the
    # first Update, so the refill runs from the first step on
    want = 1000.0 + (500.0 - 1000.0) * (1 - dt / 15.0) ** n
    assert math.isclose(h.jsr()[0, 0], want, rel_tol=1e-12)


def test_jsr_depletion_first_step_buffered(oracle_mod):
    """With the jSR at the NSR value the first step moves it by
    -dt/beta * J Velem / V, beta = 1 + n K CQ / (K + c)^2 (rapid CQ buffering)."""
    g, c0 = 28339.9, 1000.0
    sp = spec(g=g)
    ps = pools()
    h = oracle_mod.OracleHybridRun(sp, ps)
    h.reset()
    h.init_opening([0])
    h.update()
    J = g * (c0 - 0.1)
    assert math.isclose(h.dyad.total_sr_current()[0, 0], J, rel_tol=1e-15)
    dt = 1e-5
    h.step(dt, 1, threads=1)
    beta = 1 + 15.0 * 600.0 * 800.0 / (600.0 + c0) ** 2
    velem = 0.5 * 4.0 ** 3
    assert math.isclose(h.jsr()[0, 0], c0 - dt / beta * (J * velem / 1e7), rel_tol=1e-14)


def test_constant_jsr_stays(oracle_mod):
    sp = spec(g=28339.9)
    h = oracle_mod.OracleHybridRun(sp, pools(jsr_constant=True))
    h.reset()
    h.init_opening([0])
    h.update()
    h.step(1e-5, 20, threads=1)
    assert h.jsr()[0, 0] == 1000.0


def _test_input(n_sites=1, leak=True, serca=True):
    sp, j = config.load_json(os.path.join(GOLDEN, "test_input_Parameters.json"), n_sites=n_sites)
    ps = config.pools_from_json(j, sp)
    if not leak:
        ps.cyto_leak_mult = 0.0
    if not serca:
        ps.serca = {}
    return sp, ps, float(j["dt, ns"]) * 1e-6


def test_cytosol_takes_the_dyad_efflux(oracle_mod):
    """Without leak and SERCA, cytosolic Ca (free + bound) changes per step by
    dt times the Ca and Ca-buffer effluxes of the previous dyad update."""
    sp, ps, dt = _test_input(leak=False, serca=False)
    sp.init_channel = [4]
    h = oracle_mod.OracleHybridRun(sp, ps)
    h.reset()
    h.init_opening()
    h.update()
    h.step(dt, 3000, threads=1)  # let the released Ca reach the rim
    names = sp.field_names()
    ca_fields = [k for k, nm in enumerate(names) if nm.startswith("Calcium")]
    cyto_ca_bound = [0, 1, 2, 3]  # Ca-ATP, Ca-Calmodulin, Ca-Fluo-4, Ca-Troponin (ion-major)
    for _ in range(5):
        eff = h.dyad.effluxes()[0]
        ions, bound = h.cytosol()
        before = ions[0, 0] + bound[cyto_ca_bound, 0].sum()
        h.step(dt, 1, threads=1)
        ions, bound = h.cytosol()
        after = ions[0, 0] + bound[cyto_ca_bound, 0].sum()
        flux = eff[ca_fields].sum()
        assert flux != 0.0
        assert math.isclose(after - before, dt * flux, rel_tol=1e-6, abs_tol=1e-16)


def test_reference_indexing_matches_per_site_for_one_site(oracle_mod):
    """SRPool::Run reads ions[idx + j*S] and ions_NSR[i + idx*nions]
    (SRPool.cpp:75,80); for one site these are the site's own values, so the
    reference-faithful and per-site forms agree on the reference's test input."""
    sp, ps, dt = _test_input()
    sp.init_channel = [7]
    out = []
    for ref in (True, False):
        h = oracle_mod.OracleHybridRun(sp, ps, ref_indexing=ref)
        h.reset()
        h.init_opening()
        h.update()
        h.step(dt, 300, threads=1)
        out.append((h.jsr(), h.cytosol()[0], h.cytosol()[1], h.dyad.field(0), h.dyad.channel_states(),
                    h.dyad.rng_states()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert out[0][0][0, 0] < 1000.0  # the jSR depleted
