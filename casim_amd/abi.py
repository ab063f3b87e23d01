"""ctypes mirror of include/casim_dyad.h and the loader of libcasim_dyad.so.

The library is built in-tree by casim_amd/build.py; importing the product
never falls back to anything else: a missing library raises at load time.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcasim_dyad.so")

CASIM_MODEL_SHEEP_CANNELL = 1
CASIM_MODEL_TWO_STATE_MWC = 2
CASIM_MODEL_TWO_STATE_AHTG = 3
CASIM_MAX_MODEL_PARAMS = 16
CASIM_E_INVALID, CASIM_E_DEVICE, CASIM_E_STATE, CASIM_E_RANGE = -1, -2, -3, -4
CASIM_FP64 = 0
CASIM_FP32 = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


class CasimPoolConfig(C.Structure):
    _fields_ = [
        ("jsr_constant", C.c_int32), ("jsr_cq_total", C.c_double), ("jsr_volume", C.c_double),
        ("jsr_c0", _dp), ("jsr_cq_sites", _dp), ("jsr_k_cq", _dp), ("jsr_tau_refill", _dp),
        ("nsr_c", _dp),
        ("cytosol_pool", C.c_int32), ("cyto_leak_mult", C.c_double), ("n_cyto_buffers", C.c_int32),
        ("cyto_buffer_total", _dp), ("cyto_buffer_D", _dp), ("cyto_buffer_kon", _dp), ("cyto_buffer_koff", _dp),
        ("cyto_buffer_binds", _ip),
        ("serca_binds", _ip), ("serca_jmax", _dp), ("serca_kup", _dp),
    ]


class CasimCytosolBoxConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("h", C.c_double), ("origin_x", C.c_double), ("origin_y", C.c_double),
        ("substeps", C.c_int32), ("far_dirichlet", C.c_int32),
        ("ion_D", _dp), ("n_buffers", C.c_int32), ("buffer_total", _dp), ("n_bound", C.c_int32),
        ("bound_ion", _ip), ("bound_buffer", _ip), ("bound_D", _dp), ("bound_kon", _dp), ("bound_koff", _dp),
        ("bound_dyad", _ip), ("serca_jmax", _dp), ("serca_k", _dp),
    ]


class CasimDyadConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("z", C.c_double), ("radius", C.c_double), ("voltage", C.c_double),
        ("n_sites", C.c_int32),
        ("n_ions", C.c_int32), ("n_buffers", C.c_int32), ("n_bound", C.c_int32),
        ("ion_D", _dp), ("ion_c0", _dp), ("buffer_total", _dp), ("buffer_faces_only", _ip),
        ("bound_ion", _ip), ("bound_buffer", _ip),
        ("bound_D", _dp), ("bound_kon", _dp), ("bound_koff", _dp),
        ("start_equilibrium", C.c_int32),
        ("sl_binds", _ip), ("sl_N1", _dp), ("sl_K1", _dp), ("sl_N2", _dp), ("sl_K2", _dp),
        ("calcium_index", C.c_int32), ("magnesium_index", C.c_int32),
        ("n_channel_types", C.c_int32),
        ("type_model", _ip), ("type_params", _dp), ("type_conductance", _dp),
        ("site_channel_offset", _ip), ("channel_x", _dp), ("channel_y", _dp), ("channel_type", _ip),
        ("rng_state", _up),
        ("precision", C.c_int32),
    ]


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def uptr(a):
    return a.ctypes.data_as(_up)


_lib = None


def _declare(lib):
    P = C.c_void_p
    sig = {
        "casim_last_error": (C.c_char_p, []),
        "casim_version": (C.c_char_p, []),
        "casim_rng_stream_states": (C.c_int, [_up, C.c_int32, C.c_int32, _up]),
        "casim_rng_advance": (C.c_int, [_up, C.c_int64]),
        "casim_rng_draws": (C.c_int, [_up, C.c_int32, _dp]),
        "casim_dyad_create": (C.c_int, [C.POINTER(CasimDyadConfig), C.POINTER(P)]),
        "casim_dyad_destroy": (C.c_int, [P]),
        "casim_dyad_reset": (C.c_int, [P]),
        "casim_dyad_init_opening": (C.c_int, [P, C.c_int32, C.c_int32, _ip]),
        "casim_dyad_update": (C.c_int, [P, _dp, _dp, _dp, C.c_double]),
        "casim_dyad_attach_pools": (C.c_int, [P, C.POINTER(CasimPoolConfig)]),
        "casim_dyad_get_pools": (C.c_int, [P, _dp, _dp, _dp]),
        "casim_dyad_cytosol_bound": (C.c_int32, [P]),
        "casim_dyad_attach_cytosol_box": (C.c_int, [P, C.POINTER(CasimCytosolBoxConfig)]),
        "casim_dyad_cytosol_box_fields": (C.c_int32, [P]),
        "casim_dyad_get_cytosol_box": (C.c_int, [P, C.c_int32, _dp]),
        "casim_dyad_set_cytosol_box": (C.c_int, [P, C.c_int32, _dp]),
        "casim_dyad_run_rd": (C.c_int, [P, C.c_double, C.c_int32]),
        "casim_dyad_run_mc": (C.c_int, [P, C.c_double, C.c_int32]),
        "casim_dyad_step": (C.c_int, [P, C.c_double, C.c_int32]),
        "casim_dyad_step_profiled": (C.c_int, [P, C.c_double, C.c_int32, _dp]),
        "casim_dyad_synchronize": (C.c_int, [P]),
        "casim_dyad_get_effluxes": (C.c_int, [P, _dp]),
        "casim_dyad_get_total_sr_current": (C.c_int, [P, _dp]),
        "casim_dyad_get_field": (C.c_int, [P, C.c_int32, _dp]),
        "casim_dyad_set_field": (C.c_int, [P, C.c_int32, _dp]),
        "casim_dyad_get_channel_states": (C.c_int, [P, _ip]),
        "casim_dyad_set_channel_states": (C.c_int, [P, _ip]),
        "casim_dyad_get_mouths": (C.c_int, [P, _dp]),
        "casim_dyad_get_release_stats": (C.c_int, [P, _dp]),
        "casim_dyad_get_rng_states": (C.c_int, [P, _up]),
        "casim_dyad_set_rng_states": (C.c_int, [P, _up]),
        "casim_dyad_element_volume": (C.c_double, [P]),
        "casim_dyad_max_channels": (C.c_int32, [P]),
        "casim_dyad_site_channels": (C.c_int32, [P, C.c_int32]),
        "casim_dyad_channel_node": (C.c_int32, [P, C.c_int32, C.c_int32]),
        "casim_dyad_bytes_per_site_step": (C.c_double, [P]),
        "casim_dyad_set_tuning": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32]),
        "casim_dyad_set_steps_per_pass": (C.c_int, [P, C.c_int32]),
        "casim_dyad_set_split": (C.c_int, [P, C.c_int32]),
        "casim_dyad_record_open_counts": (C.c_int, [P, C.c_int32, C.c_int32]),
        "casim_dyad_get_open_counts": (C.c_int, [P, C.c_void_p]),
        "casim_dyad_steps_per_pass": (C.c_int32, [P]),
        "casim_dyad_stream": (C.c_void_p, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def header_symbols():
    """Names of every function the headers under include/ declare."""
    import glob
    import re
    names = set()
    for hdr in glob.glob(os.path.join(os.path.dirname(HERE), "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
        names |= set(re.findall(r"\b(casim_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def load():
    """Load libcasim_dyad.so (built by casim_amd.build); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libcasim_dyad.so is not built (%s); run casim_amd.build.build() "
                "or __graft_entry__.build()" % LIB_PATH)
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


class CasimError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = load().casim_last_error().decode()
        raise CasimError("casim error %d: %s" % (rc, msg))
    return rc


def rng_stream_states(seed, n, first=0):
    lib = load()
    s = np.ascontiguousarray(seed, dtype=np.uint64)
    out = np.zeros((n, 6), dtype=np.uint64)
    check(lib.casim_rng_stream_states(uptr(s), first, n, uptr(out)))
    return out
