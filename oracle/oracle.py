"""ctypes access to the oracle (liboracle.so) and to oracle/_ref/libcasim_ref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libcasim_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)
_lib = None
_ref = None


def build(ref=False):
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        P = C.c_void_p
        sig = {
            "oracle_rng_check_seed": (C.c_int, [_up]),
            "oracle_rng_stream_states": (None, [_up, C.c_int, C.c_int, _up]),
            "oracle_rng_draws": (None, [_up, C.c_int, _dp]),
            "oracle_rng_substream": (None, [_up]),
            "oracle_rng_jump_matrix": (None, [C.c_int, C.c_int, _dp]),
            "oracle_channel_rate": (C.c_double, [C.c_int, _dp, C.c_int, C.c_double, C.c_double]),
            "oracle_buffer_equilibrium": (None, [C.c_int, _dp, _dp, _dp, C.c_double, _dp]),
            "oracle_create": (P, [C.c_void_p]),
            "oracle_destroy": (None, [P]),
            "oracle_reset": (None, [P]),
            "oracle_init_opening": (None, [P, C.c_int, C.c_int]),
            "oracle_update": (None, [P, _dp, _dp, _dp, C.c_double]),
            "oracle_run_rd": (None, [P, C.c_double, C.c_int]),
            "oracle_run_mc": (None, [P, C.c_double, C.c_int]),
            "oracle_step": (None, [P, C.c_double, C.c_int, C.c_int]),
            "oracle_step_timed": (None, [P, C.c_double, C.c_int, C.c_int, _dp]),
            "oracle_max_channels": (C.c_int, [P]),
            "oracle_site_channels": (C.c_int, [P, C.c_int]),
            "oracle_channel_node": (C.c_int, [P, C.c_int, C.c_int]),
            "oracle_element_volume": (C.c_double, [P]),
            "oracle_bound_init": (None, [P, _dp]),
            "oracle_get_field": (None, [P, C.c_int, _dp]),
            "oracle_set_field": (None, [P, C.c_int, _dp]),
            "oracle_get_free_buffer": (None, [P, C.c_int, _dp]),
            "oracle_get_channel_states": (None, [P, _ip]),
            "oracle_set_channel_states": (None, [P, _ip]),
            "oracle_get_mouths": (None, [P, _dp]),
            "oracle_get_effluxes": (None, [P, _dp]),
            "oracle_get_total_sr_current": (None, [P, _dp]),
            "oracle_get_rng_states": (None, [P, _up]),
            "oracle_set_rng_states": (None, [P, _up]),
            "oracle_hybrid_create": (P, [P, C.c_void_p, _dp, _dp, C.c_int]),
            "oracle_hybrid_destroy": (None, [P]),
            "oracle_hybrid_reset": (None, [P]),
            "oracle_hybrid_update": (None, [P]),
            "oracle_hybrid_step": (None, [P, C.c_double, C.c_int, C.c_int]),
            "oracle_hybrid_get_jsr": (None, [P, _dp]),
            "oracle_hybrid_get_cytosol": (None, [P, _dp, _dp]),
            "oracle_hybrid_cyto_bound": (C.c_int, [P]),
            "oracle_hybrid_set_cytosol": (None, [P, _dp, _dp]),
            "oracle_hybrid_attach_box": (C.c_int, [P, C.c_void_p]),
            "oracle_hybrid_box_fields": (C.c_int, [P]),
            "oracle_hybrid_get_box": (None, [P, C.c_int, _dp]),
            "oracle_hybrid_set_box": (None, [P, C.c_int, _dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(lib, n)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def load_ref():
    """The reference's own RNG / kinetic / Point code (None if not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        lib = C.CDLL(REF_LIB)
        ul = C.POINTER(C.c_ulong)
        lib.ref_stream_draws.argtypes = [ul, C.c_int, C.c_int, _dp, ul]
        lib.ref_stream_substream.argtypes = [ul, C.c_int, ul]
        lib.ref_stream_advance.argtypes = [ul, C.c_long, C.c_long, ul]
        lib.ref_parallel_rng_draws.argtypes = [ul, C.c_int, _dp]
        for n in ("ref_sheep_cannell_rate",):
            getattr(lib, n).argtypes = [_dp, C.c_int, C.c_double]
            getattr(lib, n).restype = C.c_double
        for n in ("ref_mwc_rate", "ref_ahtg_rate"):
            getattr(lib, n).argtypes = [_dp, C.c_int, C.c_double, C.c_double]
            getattr(lib, n).restype = C.c_double
        lib.ref_coords_on_grid.argtypes = [C.c_double] * 6 + [_ip]
        _ref = lib
    return _ref


def d(a):
    return a.ctypes.data_as(_dp)


def u(a):
    return a.ctypes.data_as(_up)


def i32(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """The CPU restatement of DyadRDCPU / DyadRD2DCPU for one batch of sites."""

    def __init__(self, spec):
        self.lib = load()
        self.spec = spec
        self._cfg, self._keep = spec.to_struct()
        self.h = self.lib.oracle_create(C.byref(self._cfg))
        if not self.h:
            raise ValueError("oracle_create rejected the configuration")
        self.S = spec.n_sites
        self.M = self.lib.oracle_max_channels(self.h)
        self.NS = spec.nodes_per_site()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def reset(self):
        self.lib.oracle_reset(self.h)

    def init_opening(self, channels=None):
        chans = channels if channels is not None else self.spec.init_channel
        for s in range(self.S):
            self.lib.oracle_init_opening(self.h, s, int(chans[s]))

    def update(self, rim_ions, rim_bound, jsr, voltage=0.0):
        rim_ions = np.ascontiguousarray(rim_ions, dtype=np.float64)
        rim_bound = np.ascontiguousarray(rim_bound, dtype=np.float64)
        jsr = np.ascontiguousarray(jsr, dtype=np.float64)
        self.lib.oracle_update(self.h, d(rim_ions), d(rim_bound), d(jsr), voltage)

    def run_rd(self, dt, site):
        self.lib.oracle_run_rd(self.h, dt, site)

    def run_mc(self, dt, site):
        self.lib.oracle_run_mc(self.h, dt, site)

    def step(self, dt, k, threads=None):
        self.lib.oracle_step(self.h, dt, k, threads or os.cpu_count() or 1)

    def step_timed(self, dt, k, threads=None):
        """step(); returns the wall seconds spent in the serial Update part."""
        t = np.zeros(1)
        self.lib.oracle_step_timed(self.h, dt, k, threads or os.cpu_count() or 1, d(t))
        return float(t[0])

    def field(self, f):
        out = np.empty(self.NS * self.S)
        self.lib.oracle_get_field(self.h, f, d(out))
        return out.reshape(self.S, self.spec.nx, self.spec.ny, self.spec.nz)

    def set_field(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.lib.oracle_set_field(self.h, f, d(a))

    def free_buffer(self, k):
        out = np.empty(self.NS * self.S)
        self.lib.oracle_get_free_buffer(self.h, k, d(out))
        return out.reshape(self.S, self.spec.nx, self.spec.ny, self.spec.nz)

    def channel_states(self):
        out = np.empty(self.S * self.M, dtype=np.int32)
        self.lib.oracle_get_channel_states(self.h, i32(out))
        return out.reshape(self.S, self.M)

    def set_channel_states(self, st):
        st = np.ascontiguousarray(st, dtype=np.int32).reshape(-1)
        self.lib.oracle_set_channel_states(self.h, i32(st))

    def mouths(self):
        k = self.spec.n_ions + self.spec.n_bound + self.spec.n_buffers
        out = np.empty(self.M * self.S * k)
        self.lib.oracle_get_mouths(self.h, d(out))
        return out.reshape(k, self.S, self.M)

    def effluxes(self):
        out = np.empty((self.spec.n_ions + self.spec.n_bound) * self.S)
        self.lib.oracle_get_effluxes(self.h, d(out))
        return out.reshape(self.S, -1)

    def total_sr_current(self):
        out = np.empty(self.spec.n_ions * self.S)
        self.lib.oracle_get_total_sr_current(self.h, d(out))
        return out.reshape(self.spec.n_ions, self.S)

    def rng_states(self):
        out = np.empty(self.S * 6, dtype=np.uint64)
        self.lib.oracle_get_rng_states(self.h, u(out))
        return out.reshape(self.S, 6)

    def bound_init(self):
        out = np.empty(self.spec.n_bound)
        self.lib.oracle_bound_init(self.h, d(out))
        return out

    def element_volume(self):
        return self.lib.oracle_element_volume(self.h)

    def channel_node(self, s, ch):
        return self.lib.oracle_channel_node(self.h, s, ch)


def rng_draws(state, n):
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    out = np.empty(n)
    load().oracle_rng_draws(u(st), n, d(out))
    return out, st


def stream_states(seed, n, first=0):
    s = np.ascontiguousarray(seed, dtype=np.uint64)
    out = np.empty((n, 6), dtype=np.uint64)
    load().oracle_rng_stream_states(u(s), first, n, u(out))
    return out


def channel_rate(model, params, state, ca, mg=0.0):
    p = np.zeros(16)
    p[:len(params)] = params
    return load().oracle_channel_rate(model, d(p), state, ca, mg)


class OracleHybridRun:
    """The reference's batch loop around the oracle dyad: jSR pool, NSR and
    cytosol pool (or a constant rim), RunSimulationsCPU's step order."""

    def __init__(self, spec, pools, rim=None, ref_indexing=False):
        self.dyad = Oracle(spec)
        self.lib = self.dyad.lib
        self.spec = spec
        self._pc, self._pkeep = pools.to_struct(spec.ions)
        ri, rb = rim if rim is not None else spec.rim_arrays()
        self._ri = np.ascontiguousarray(ri, dtype=np.float64)
        self._rb = np.ascontiguousarray(rb if len(rb) else np.zeros(1), dtype=np.float64)
        self.h = self.lib.oracle_hybrid_create(self.dyad.h, C.byref(self._pc), d(self._ri), d(self._rb),
                                               int(ref_indexing))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_hybrid_destroy(self.h)
            self.h = None

    def reset(self):
        self.lib.oracle_hybrid_reset(self.h)

    def init_opening(self, channels=None):
        self.dyad.init_opening(channels)

    def update(self):
        self.lib.oracle_hybrid_update(self.h)

    def step(self, dt, k, threads=None):
        self.lib.oracle_hybrid_step(self.h, dt, k, threads or os.cpu_count() or 1)

    def jsr(self):
        out = np.empty(self.spec.n_ions * self.spec.n_sites)
        self.lib.oracle_hybrid_get_jsr(self.h, d(out))
        return out.reshape(self.spec.n_ions, self.spec.n_sites)

    def attach_box(self, box):
        """The 3D cytosol box (config.CytosolBoxSpec) around every site."""
        self._bc, self._bkeep = box.to_struct(self.spec)
        self._box = box
        if self.lib.oracle_hybrid_attach_box(self.h, C.byref(self._bc)) != 0:
            raise ValueError("cytosol box refused")

    def box_field(self, f):
        """Box field f as [site, z, y, x]."""
        b = self._box
        out = np.empty(self.spec.n_sites * b.nz * b.ny * b.nx)
        self.lib.oracle_hybrid_get_box(self.h, f, d(out))
        return out.reshape(self.spec.n_sites, b.nz, b.ny, b.nx)

    def set_box_field(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self.lib.oracle_hybrid_set_box(self.h, f, d(a))

    def cytosol(self):
        nb = self.lib.oracle_hybrid_cyto_bound(self.h)
        ions = np.empty(self.spec.n_ions * self.spec.n_sites)
        bound = np.empty(max(nb, 1) * self.spec.n_sites)
        self.lib.oracle_hybrid_get_cytosol(self.h, d(ions), d(bound))
        return ions.reshape(self.spec.n_ions, -1), bound[:nb * self.spec.n_sites].reshape(nb, -1)

    def set_cytosol(self, ions=None, bound=None):
        """Overwrite the cytosol pool's ions [ion, site] and bound fields [b, site]."""
        ia = None if ions is None else np.ascontiguousarray(ions, dtype=np.float64).reshape(-1)
        ba = None if bound is None else np.ascontiguousarray(bound, dtype=np.float64).reshape(-1)
        self.lib.oracle_hybrid_set_cytosol(self.h, None if ia is None else d(ia), None if ba is None else d(ba))
