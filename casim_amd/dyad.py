"""DyadSeries over the C ABI: the dyad compartment of CaSim's batched solver.

DyadHIP mirrors the reference's `DyadSeries` interface (DyadSeries.h:9-44)
method for method, so a ParallelHybridSolver-style driver can call it as it
calls DyadRDCPU / DyadRDMP.  Every call goes to libcasim_dyad.so; there is no
other implementation behind it.
"""
import ctypes as C

import numpy as np

from . import abi


class DyadHIP:
    def __init__(self, spec):
        self.lib = abi.load()
        self.spec = spec
        self._cfg, self._keep = spec.to_struct()
        h = C.c_void_p()
        abi.check(self.lib.casim_dyad_create(C.byref(self._cfg), C.byref(h)))
        self.h = h
        self.S = spec.n_sites
        self.M = self.lib.casim_dyad_max_channels(self.h)
        self.NS = spec.nodes_per_site()
        self.dtype = np.float64 if spec.precision == abi.CASIM_FP64 else np.float32

    def close(self):
        if getattr(self, "h", None):
            self.lib.casim_dyad_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- DyadSeries methods ----
    def UsesGPU(self):
        return True

    def Reset(self):
        abi.check(self.lib.casim_dyad_reset(self.h))

    def InitOpening(self, site, channel):
        abi.check(self.lib.casim_dyad_init_opening(self.h, site, channel, None))

    def InitOpeningAll(self, channels=None):
        chans = channels if channels is not None else self.spec.init_channel
        a = np.ascontiguousarray(chans, dtype=np.int32)
        abi.check(self.lib.casim_dyad_init_opening(self.h, -1, 0, abi.iptr(a)))

    def Update(self, rim_ions, rim_bound, jsr, voltage=0.0):
        rim_ions = np.ascontiguousarray(rim_ions, dtype=np.float64)
        rim_bound = np.ascontiguousarray(rim_bound, dtype=np.float64)
        if rim_bound.size == 0:
            rim_bound = np.zeros(1)
        jsr = np.ascontiguousarray(jsr, dtype=np.float64)
        abi.check(self.lib.casim_dyad_update(self.h, abi.dptr(rim_ions), abi.dptr(rim_bound), abi.dptr(jsr),
                                             voltage))

    def AttachPools(self, pools):
        """jSR / NSR / cytosol compartments on the device (config.PoolSpec)."""
        self._pc, self._pkeep = pools.to_struct(self.spec.ions)
        abi.check(self.lib.casim_dyad_attach_pools(self.h, C.byref(self._pc)))

    def UpdatePools(self, rim_ions=None, rim_bound=None, voltage=0.0):
        """ParallelHybridSolver::Update with the pools attached."""
        ri = None if rim_ions is None else np.ascontiguousarray(rim_ions, dtype=np.float64)
        rb = None if rim_bound is None else np.ascontiguousarray(rim_bound, dtype=np.float64)
        if rb is not None and rb.size == 0:
            rb = np.zeros(1)
        abi.check(self.lib.casim_dyad_update(self.h, None if ri is None else abi.dptr(ri),
                                             None if rb is None else abi.dptr(rb), None, voltage))

    def GetPools(self):
        """(jSR [ion, site], cytosol ions [ion, site], cytosol bound [b, site])."""
        n = self.spec.n_ions * self.S
        nb = self.lib.casim_dyad_cytosol_bound(self.h)
        j, ci, cb = np.empty(n), np.empty(n), np.empty(max(nb, 1) * self.S)
        abi.check(self.lib.casim_dyad_get_pools(self.h, abi.dptr(j), abi.dptr(ci), abi.dptr(cb)))
        return (j.reshape(self.spec.n_ions, -1), ci.reshape(self.spec.n_ions, -1),
                cb[:nb * self.S].reshape(nb, -1))

    def AttachCytosolBox(self, box):
        """3D cytosol box around every site, coupled at the rim (config.CytosolBoxSpec)."""
        self._bc, self._bkeep = box.to_struct(self.spec)
        self._box = box
        abi.check(self.lib.casim_dyad_attach_cytosol_box(self.h, C.byref(self._bc)))

    def CytosolBoxFields(self):
        return int(self.lib.casim_dyad_cytosol_box_fields(self.h))

    def GetCytosolBox(self, f):
        """Box field f (ions, then the box's bound fields) as [site, z, y, x]."""
        b = self._box
        out = np.empty(self.S * b.nz * b.ny * b.nx)
        abi.check(self.lib.casim_dyad_get_cytosol_box(self.h, f, abi.dptr(out)))
        return out.reshape(self.S, b.nz, b.ny, b.nx)

    def SetCytosolBox(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        abi.check(self.lib.casim_dyad_set_cytosol_box(self.h, f, abi.dptr(a)))

    def RunRD(self, dt, site=-1):
        abi.check(self.lib.casim_dyad_run_rd(self.h, dt, site))

    def RunMC(self, dt, site=-1):
        abi.check(self.lib.casim_dyad_run_mc(self.h, dt, site))

    def Step(self, dt, k):
        abi.check(self.lib.casim_dyad_step(self.h, dt, k))

    def StepProfiled(self, dt, k):
        """k fused steps launched eagerly with HIP events around each kernel;
        returns (state-sweep kernel ms, channel kernel ms) summed over k."""
        ms = np.zeros(2)
        abi.check(self.lib.casim_dyad_step_profiled(self.h, dt, k, abi.dptr(ms)))
        return float(ms[0]), float(ms[1])

    def Synchronize(self):
        abi.check(self.lib.casim_dyad_synchronize(self.h))

    def GetElementVolume(self):
        return self.lib.casim_dyad_element_volume(self.h)

    def GetNumSRChannels(self):
        return self.M

    def GetListofIons(self):
        return [i.name for i in self.spec.ions]

    def GetEffluxes(self):
        out = np.empty((self.spec.n_ions + self.spec.n_bound) * self.S)
        abi.check(self.lib.casim_dyad_get_effluxes(self.h, abi.dptr(out)))
        return out.reshape(self.S, -1)

    def GetTotalSRCurrent(self):
        out = np.empty(self.spec.n_ions * self.S)
        abi.check(self.lib.casim_dyad_get_total_sr_current(self.h, abi.dptr(out)))
        return out.reshape(self.spec.n_ions, self.S)

    def GetField(self, f):
        out = np.empty(self.NS * self.S)
        abi.check(self.lib.casim_dyad_get_field(self.h, f, abi.dptr(out)))
        return out.reshape(self.S, self.spec.nx, self.spec.ny, self.spec.nz)

    def SetField(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        abi.check(self.lib.casim_dyad_set_field(self.h, f, abi.dptr(a)))

    def GetConcentrations(self, names=None):
        """{species name: [site][x][y][z]} (DyadRDCPU.cpp:516-539)."""
        out = {}
        for k, nm in enumerate(self.spec.field_names()):
            # bound fields are "Ion-Buffer"; buffer names may hold '-' themselves ("Fluo-4")
            if names is None or nm in names or ("-" in nm and nm.split("-", 1)[1] in names):
                out[nm] = self.GetField(k)
        return out

    def GetChannelsStates(self):
        out = np.empty(self.S * self.M, dtype=np.int32)
        abi.check(self.lib.casim_dyad_get_channel_states(self.h, abi.iptr(out)))
        return out.reshape(self.S, self.M)

    def SetChannelsStates(self, st):
        st = np.ascontiguousarray(st, dtype=np.int32).reshape(-1)
        abi.check(self.lib.casim_dyad_set_channel_states(self.h, abi.iptr(st)))

    def GetIonsNearChannels(self):
        k = self.spec.n_ions + self.spec.n_bound + self.spec.n_buffers
        out = np.empty(self.M * self.S * k)
        abi.check(self.lib.casim_dyad_get_mouths(self.h, abi.dptr(out)))
        return out.reshape(k, self.S, self.M)

    def GetReleaseStats(self):
        """[site, (amp, ttp, tte, Q, Q/tte, N now)] since InitOpening."""
        out = np.empty(self.S * 6)
        abi.check(self.lib.casim_dyad_get_release_stats(self.h, abi.dptr(out)))
        return out.reshape(self.S, 6)

    def GetRngStates(self):
        out = np.empty(self.S * 6, dtype=np.uint64)
        abi.check(self.lib.casim_dyad_get_rng_states(self.h, abi.uptr(out)))
        return out.reshape(self.S, 6)

    def SetRngStates(self, st):
        st = np.ascontiguousarray(st, dtype=np.uint64).reshape(-1)
        abi.check(self.lib.casim_dyad_set_rng_states(self.h, abi.uptr(st)))

    def ChannelNode(self, site, ch):
        return self.lib.casim_dyad_channel_node(self.h, site, ch)

    def BytesPerSiteStep(self):
        return self.lib.casim_dyad_bytes_per_site_step(self.h)

    def SetTuning(self, planes=0, threads=0, prefetch=0):
        abi.check(self.lib.casim_dyad_set_tuning(self.h, planes, threads, prefetch))

    def RecordOpenCounts(self, steps_per_bin, max_bins):
        """Open-channel counts per site every steps_per_bin steps (bin 0:
        after InitOpening), for the analyzer's time histograms."""
        abi.check(self.lib.casim_dyad_record_open_counts(self.h, steps_per_bin, max_bins))
        self._bins = max_bins

    def GetOpenCounts(self):
        """[bins, S] uint8, 255 where a bin was not reached."""
        out = np.empty((self._bins, self.spec.n_sites), dtype=np.uint8)
        abi.check(self.lib.casim_dyad_get_open_counts(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def SetStepsPerPass(self, steps):
        abi.check(self.lib.casim_dyad_set_steps_per_pass(self.h, steps))

    def StepsPerPass(self):
        return self.lib.casim_dyad_steps_per_pass(self.h)

    def SetSplit(self, on):
        """Split pipelining of Step (two half-batches on two streams); the
        results do not depend on it."""
        abi.check(self.lib.casim_dyad_set_split(self.h, int(bool(on))))

    def stream(self):
        return self.lib.casim_dyad_stream(self.h)
