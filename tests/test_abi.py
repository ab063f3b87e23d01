"""The C ABI library loads and exports every symbol include/casim_dyad.h
declares; host-only entry points work without a device; device entry points
fail with a status code (not a crash) when no device is present."""
import ctypes as C
import os

import numpy as np
import pytest

from casim_amd import abi, workloads


def test_library_exports_header_symbols():
    lib = abi.load()
    names = abi.header_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_version_string():
    assert b"gfx950" in abi.load().casim_version()


def test_null_arguments_are_rejected():
    lib = abi.load()
    h = C.c_void_p()
    assert lib.casim_dyad_create(None, C.byref(h)) == -1
    assert b"null" in lib.casim_last_error()
    assert lib.casim_dyad_reset(None) == -1
    assert lib.casim_dyad_destroy(None) == 0


def test_invalid_configs_are_rejected():
    lib = abi.load()
    sp = workloads.single_dyad(1)
    cfg, keep = sp.to_struct()
    cfg.dim = 4
    h = C.c_void_p()
    assert lib.casim_dyad_create(C.byref(cfg), C.byref(h)) == -1
    assert b"dim" in lib.casim_last_error()
    cfg, keep = sp.to_struct()
    cfg.n_ions = 9
    assert lib.casim_dyad_create(C.byref(cfg), C.byref(h)) == -1
    cfg, keep = sp.to_struct()
    cfg.nx, cfg.ny, cfg.nz = 4096, 4096, 20  # 2^28.3 nodes: beyond the sweeps' 32-bit offsets
    assert lib.casim_dyad_create(C.byref(cfg), C.byref(h)) == -1
    assert b"2^28" in lib.casim_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a device is present")
def test_create_without_device_reports_device_error():
    lib = abi.load()
    sp = workloads.single_dyad(1)
    cfg, keep = sp.to_struct()
    h = C.c_void_p()
    assert lib.casim_dyad_create(C.byref(cfg), C.byref(h)) == -2
    assert h.value is None


def test_bytes_per_site_step_formula():
    """The algorithmic traffic of one site-step: every stored node read once,
    every updated node written once (3 grid fields + the membrane buffer on
    its two faces, 61 x 61 x 5 grid, FP64 = 979,472 B)."""
    nx = ny = 61
    nz = 5
    all_, inner = nx * ny * nz, (nx - 2) * (ny - 2) * nz
    fall, finner = 2 * nx * ny, 2 * (nx - 2) * (ny - 2)
    assert 8 * (3 * (all_ + inner) + fall + finner) == 979472
