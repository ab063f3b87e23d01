"""Output in the reference's HDF5 layout (ParallelHybridSolver.cpp:46-52,
177-188, 234-337).  h5py is not importable in this environment, so the writer
runs on MemoryStore, the in-memory stand-in with h5py's interface subset; the
tests pin the reference's rules for creating, chunking and extending datasets
and placing each (time index, batch) write, the dataset names and shapes of
each group, and -- on the GPU -- that a device run writes the same file as the
oracle's run.  The bytes of a real HDF5 file are parity unpinned here."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import spec  # noqa: E402

from casim_amd import output, workloads  # noqa: E402
from casim_amd.config import Buffer, Ion, PoolSpec  # noqa: E402


class StampSource:
    """A dyad stand-in whose every value is 1000 t + 10 b + (site index / 100),
    so each stored element says where it was written from."""

    def __init__(self, sp, n_cyto_bound=0):
        self.sp, self.t, self.b = sp, 0, 0
        self.S, self.M = sp.n_sites, sp.max_channels()
        self.ncb = n_cyto_bound

    def _v(self, shape):
        site = np.arange(self.S).reshape([self.S] + [1] * (len(shape) - 1)) / 100.0
        return np.broadcast_to(1000.0 * self.t + 10.0 * self.b + site, shape).copy()

    def GetField(self, k):
        return self._v((self.S, self.sp.nx, self.sp.ny, self.sp.nz)) + k * 1e-4

    def GetChannelsStates(self):
        return np.full((self.S, self.M), 1000 * self.t + 10 * self.b, dtype=np.int32)

    def GetIonsNearChannels(self):
        k = self.sp.n_ions + self.sp.n_bound + self.sp.n_buffers
        return np.stack([self._v((self.S, self.M))] * k)

    def GetPools(self):
        n = self.sp.n_ions
        return (self._v((self.S,))[None].repeat(n, 0), self._v((self.S,))[None].repeat(n, 0),
                self._v((self.S,))[None].repeat(max(self.ncb, 1), 0)[:self.ncb])


def small():
    ions = [Ion("Calcium", 220000.0, 0.1)]
    buffers = [Buffer("A", 5000.0, {"Calcium": (140000.0, 0.225, 45.0)}),
               Buffer("F", 1000.0, {"Calcium": (0.0, 0.115, 1.0)}, faces_only=True)]
    return spec(dim=3, nx=7, ny=6, nz=5, h=4.0, ions=ions, buffers=buffers, n_sites=3,
                channels=[(12.0, 8.0), (8.0, 12.0)])


def test_reference_write_rules():
    """Two batches of 3 sites, 10 steps, strides Dyad 5, Cytosol 2, SR 1,
    Channels 2, IonsChannels 5: shapes grow as the reference's
    _ExtendDataset grows them, maxshape and chunks are set at creation, and
    every (time index, batch) lands in [t, b*S:(b+1)*S]."""
    sp = small()
    pools = PoolSpec(jsr_constant=False, jsr_cq_total=1.0, jsr_volume=1.0, jsr_c0=[1000.0], jsr_cq_sites=[1.0],
                     jsr_k_cq=[1.0], jsr_tau_refill=[1.0], nsr_c=[1000.0], cytosol_pool=True,
                     cyto_buffers=[Buffer("Tn", 70.0, {"Calcium": (0.0, 0.0327, 0.0196)})])
    it = {"Dyad": 5, "Cytosol": 2, "SR": 1, "Channels": 2, "IonsChannels": 5}
    store = output.MemoryStore(None)
    out = output.ReferenceOutput(store, sp, ["Calcium", "A", "Tn"], 10, it, 6, pools=pools)
    src = StampSource(sp, n_cyto_bound=3)
    for b in range(2):
        for j in range(11):
            src.t, src.b = j, b
            out.write(j, b, src)
    d = store["Dyad"]
    assert sorted(d.keys()) == ["Calcium", "Calcium-A"]  # F is not in the outputs
    ds = d["Calcium"]
    assert ds.shape == (3, 6, 7, 6, 5) and ds.maxshape == (3, 6, 7, 6, 5) and ds.chunks == (1, 3, 7, 6, 5)
    assert ds.dtype == np.float64
    for t in range(3):
        for b in range(2):
            want = 1000.0 * (5 * t) + 10.0 * b + np.arange(3) / 100.0
            assert np.array_equal(ds[t, 3 * b:3 * b + 3, 2, 3, 1], want)
    ch = store["Channels"]["RyR"]
    assert ch.shape == (6, 6, 2) and ch.dtype == np.int32 and ch.maxshape == (6, 6, 2)
    assert ch[4, 3, 1] == 1000 * 8 + 10
    sr = store["SR"]["Calcium"]
    assert sr.shape == (11, 6, 2) and sr.chunks == (1, 3, 2)
    assert sr[10, 5, 0] == 1000.0 * 10 + 10 + 0.02
    cy = store["Cytosol"]
    assert sorted(cy.keys()) == ["Calcium", "Calcium-A", "Calcium-Tn"]
    assert cy["Calcium-Tn"].shape == (6, 6)
    ic = store["IonsChannels"]
    assert list(ic.keys()) == ["Calcium"] and ic["Calcium"].shape == (3, 6, 2)


def test_resize_beyond_maxshape_refused():
    ds = output.MemoryStore.Dataset((1, 2), (2, 4), (1, 2), np.float64)
    ds.resize((2, 2))
    with pytest.raises(ValueError):
        ds.resize((3, 2))


def test_memory_store_saves_npz(tmp_path):
    sp = small()
    it = {g: 1 for g in output.GROUPS}
    path = str(tmp_path / "Output.h5")
    f = output.open_file(path)
    out = output.ReferenceOutput(f, sp, ["Calcium"], 2, it, 3)
    src = StampSource(sp)
    for j in range(3):
        src.t = j
        out.write(j, 0, src, jsr=np.full((1, 3, 2), 1000.0))
    out.close()
    if isinstance(f, output.MemoryStore):
        z = np.load(str(tmp_path / "Output.npz"))
        assert sorted(z.files) == ["Channels/RyR", "Dyad/Calcium", "IonsChannels/Calcium", "SR/Calcium"]
        assert z["Dyad/Calcium"].shape == (3, 3, 7, 6, 5)
        assert np.all(z["SR/Calcium"] == 1000.0)
    else:  # h5py present: read the real file back
        import h5py
        with h5py.File(path, "r") as h:
            assert h["Dyad/Calcium"].shape == (3, 3, 7, 6, 5)


def test_coupled_box_goes_to_cytosol():
    sp = workloads.coupled_dyads(2)
    box = workloads.coupled_box(sp)
    store = output.MemoryStore(None)
    out = output.ReferenceOutput(store, sp, ["Calcium", "B", "Tn"], 100, {g: 50 for g in output.GROUPS}, 2,
                                 box=box)
    assert out.dims["Cytosol"] == [51, 51, 51] and out.dims["Dyad"] == [201, 201]

    class Src(StampSource):
        def GetCytosolBox(self, f):
            return np.full((2, 51, 51, 51), float(f))
    src = Src(sp)
    out.write(0, 0, src, jsr=np.full((1, 2, 64), 1000.0))
    assert sorted(store["Cytosol"].keys()) == ["Calcium", "Calcium-B", "Calcium-Tn"]
    assert store["Cytosol"]["Calcium-Tn"][0, 1, 0, 0, 0] == 3.0
    assert store["Dyad"]["Calcium"].shape == (1, 2, 201, 201)


class OracleSource:
    """The oracle's batch loop behind the getters ReferenceOutput reads."""

    def __init__(self, ora, sp):
        self.o, self.sp = ora, sp

    def GetField(self, k):
        return self.o.dyad.field(k)

    def GetChannelsStates(self):
        return self.o.dyad.channel_states()

    def GetIonsNearChannels(self):
        k = self.sp.n_ions + self.sp.n_bound + self.sp.n_buffers
        return self.o.dyad.mouths().reshape(k, self.sp.n_sites, -1)

    def GetPools(self):
        ci, cb = self.o.cytosol()
        return self.o.jsr(), ci, cb


@pytest.mark.gpu
def test_device_run_writes_the_oracles_file(gpu, oracle_mod):
    """Two batches of a stochastic 4-site run with a depleting jSR: every
    dataset the device run writes equals the oracle run's (fields 1e-9,
    channel states exactly)."""
    from casim_amd.dyad import DyadHIP
    sp = workloads.single_dyad(4, clamped=False)
    sp.channel_types[0].params["kopen0"] *= 1e4
    pools = PoolSpec(jsr_constant=False, jsr_cq_total=10000.0, jsr_volume=2e5, jsr_c0=[1000.0], jsr_cq_sites=[1.0],
                     jsr_k_cq=[800.0], jsr_tau_refill=[100.0], nsr_c=[1000.0])
    it = {"Dyad": 20, "Cytosol": 10, "SR": 10, "Channels": 5, "IonsChannels": 10}
    n_iter, dt = 40, 1e-5
    stores = []
    for kind in ("dev", "ora"):
        store = output.MemoryStore(None)
        out = output.ReferenceOutput(store, sp, ["Calcium", "A", "B"], n_iter, it, 8, pools=pools)
        if kind == "dev":
            src = DyadHIP(sp)
            src.AttachPools(pools)
        else:
            ora = oracle_mod.OracleHybridRun(sp, pools)
            src = OracleSource(ora, sp)
        for b in range(2):
            if kind == "dev":
                src.Reset()
                src.InitOpeningAll()
                src.UpdatePools(*sp.rim_arrays())
            else:
                ora.reset()
                ora.init_opening()
                ora.update()
            out.write(0, b, src)
            for j in range(1, n_iter + 1):
                if kind == "dev":
                    src.Step(dt, 1)
                else:
                    ora.step(dt, 1)
                out.write(j, b, src)
        stores.append(store)
    dev, ora = stores
    for g in output.GROUPS:
        assert sorted(dev.groups[g].keys()) == sorted(ora.groups[g].keys()), g
        for name, ds in ora.groups[g].items():
            a, b = dev[g + "/" + name][...], ds[...]
            assert a.shape == b.shape, (g, name)
            if g == "Channels":
                assert np.array_equal(a, b), name
            else:
                assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) < 1e-9, (g, name)


def test_reference_parameter_file_drives_the_layout():
    """The reference's own test input (its "Iterations for ..." strides and
    "Outputs"): every group is laid out from it, and the 2D dyad datasets
    named in Outputs are the ones created."""
    import json
    from casim_amd import config
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_input_Parameters.json")
    sp, j = config.load_json(golden, n_sites=2)
    it = output.iterations_from_json(j)
    assert it == {g: 500 for g in output.GROUPS}
    pools = config.pools_from_json(j, sp)
    store = output.MemoryStore(None)
    out = output.ReferenceOutput(store, sp, j["Outputs"], 1000, it, 4, pools=pools)
    assert out.idx_max == {g: 3 for g in output.GROUPS}
    assert out.dims["Dyad"] == [sp.nx, sp.ny] and out.dims["SR"] == [sp.max_channels()]
    src = StampSource(sp, n_cyto_bound=len([1 for ion in sp.ions for b in list(sp.buffers) + list(pools.cyto_buffers)
                                            if ion.name in b.kinetics]))
    out.write(0, 0, src)
    names = sorted(store["Dyad"].keys())
    # a buffer's name may itself hold '-' ("Fluo-4"): the ion is the part before the first one
    assert names and all(n in j["Outputs"] or n.split("-", 1)[1] in j["Outputs"] for n in names)
    assert "Calcium-Fluo-4" in names
    assert store["Dyad"][names[0]].shape == (1, 2, sp.nx, sp.ny)
    assert store["Dyad"][names[0]].maxshape == (3, 4, sp.nx, sp.ny)


def test_dyad_getconcentrations_selects_hyphenated_buffers():
    """DyadHIP.GetConcentrations picks "Ion-Buffer" fields by the buffer's
    full name, also when that name holds '-' (the reference's "Fluo-4")."""
    from casim_amd.dyad import DyadHIP
    from casim_amd import config

    class Fake:
        spec = None

        def GetField(self, k):
            return k

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "test_input_Parameters.json")
    f = Fake()
    f.spec, _ = config.load_json(golden, n_sites=1)
    got = DyadHIP.GetConcentrations(f, ["Fluo-4"])
    assert sorted(got) == ["Calcium-Fluo-4"]
