"""casim_mlp_predict_noalloc, the ensemble prediction that allocates nothing
(its activations live in the library's fixed per-GPU workspace), against
casim_mlp_predict on the same rows: two calls in a row, the second with fewer
rows, on device arrays made through the HIP runtime; and the refusal of a
call beyond the workspace."""
import ctypes as C
import os

import numpy as np
import pytest

from casim_amd import abi, mlp, surrogate
from tests.helpers import DeviceBuffer

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_ROWS = 4096  # CASIM_MLP_NOALLOC_MAX_ROWS


def _declare(lib):
    f = lib.casim_mlp_predict_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    return f


def _rows(net, n, seed):
    """Feature rows around the training distribution of the committed ensemble."""
    z = np.load(os.path.join(ROOT, "casim_amd", "surrogate_weights.npz"), allow_pickle=False)
    g = np.random.default_rng(seed)
    x = z["in_mean"][None, :] + 0.5 * z["in_std"][None, :] * g.standard_normal((n, net.n_in))
    return np.ascontiguousarray(x, dtype=np.float32)


def _noalloc(f, net, x):
    n = x.shape[0]
    dx = DeviceBuffer(x)
    dm = DeviceBuffer(np.full(n, np.nan, dtype=np.float32))
    ds = DeviceBuffer(np.full(n, np.nan, dtype=np.float32))
    abi.check(f(net.h, dx.ptr, n, dm.ptr, ds.ptr, None))
    assert dx.hip.hipDeviceSynchronize() == 0
    return dm.read(), ds.read()


def test_predict_noalloc_equals_predict(gpu):
    f = _declare(mlp.lib())
    net, _ = surrogate.load_model(os.path.join(ROOT, "casim_amd", "surrogate_weights.npz"))
    for n, seed in ((3000, 1), (700, 2)):
        x = _rows(net, n, seed)
        want_m, want_s, _ = net.predict(x)
        got_m, got_s = _noalloc(f, net, x)
        print("rows", n, "max rel diff of the mean", float(np.max(np.abs(got_m - want_m) / want_m)), flush=True)
        assert np.isfinite(want_m).all() and (want_m > 0).all()
        np.testing.assert_allclose(got_m, want_m, rtol=1e-6, atol=0)
        np.testing.assert_allclose(got_s, want_s, rtol=1e-5, atol=1e-7)


def test_predict_noalloc_refuses_more_rows_than_the_workspace(gpu):
    f = _declare(mlp.lib())
    net, _ = surrogate.load_model(os.path.join(ROOT, "casim_amd", "surrogate_weights.npz"))
    # refused on the host before any device work: the pointers are never used
    rc = f(net.h, C.c_void_p(16), MAX_ROWS + 1, None, None, None)
    assert rc == abi.CASIM_E_RANGE
    assert b"workspace" in abi.load().casim_last_error()
