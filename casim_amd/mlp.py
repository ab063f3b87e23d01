"""Ensemble MLP of the mouth-[Ca2+] surrogate over the C ABI
(include/casim_surrogate.h)."""
import ctypes as C

import numpy as np

from . import abi

_fp = C.POINTER(C.c_float)


def _declare(lib):
    P = C.c_void_p
    sig = {
        "casim_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(P)]),
        "casim_mlp_destroy": (C.c_int, [P]),
        "casim_mlp_param_count": (C.c_int64, [P]),
        "casim_mlp_get_params": (C.c_int, [P, C.c_int32, _fp]),
        "casim_mlp_set_params": (C.c_int, [P, C.c_int32, _fp]),
        "casim_mlp_set_normalization": (C.c_int, [P, _fp, _fp, C.c_float, C.c_float]),
        "casim_mlp_set_data": (C.c_int, [P, _fp, _fp, C.c_int64]),
        "casim_mlp_train": (C.c_int, [P, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]),
        "casim_mlp_gradient": (C.c_int, [P, C.c_int32, _fp, _fp, C.c_int64, _fp, _fp]),
        "casim_mlp_predict": (C.c_int, [P, _fp, C.c_int64, _fp, _fp, _fp]),
        "casim_mlp_set_fused_inference": (C.c_int, [P, C.c_int32]),
        "casim_qs_create": (C.c_int, [P, C.POINTER(P)]),
        "casim_qs_destroy": (C.c_int, [P]),
        "casim_qs_reset": (C.c_int, [P]),
        "casim_qs_run": (C.c_int, [P, P, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
        "casim_qs_get_waiting": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "casim_qs_set_mouths": (C.c_int, [P, C.c_int32, _fp]),
        "casim_qs_get_stats": (C.c_int, [P, C.POINTER(C.c_double)]),
        "casim_qs_get_features": (C.c_int, [P, _fp]),
        "casim_qs_set_linear_feature": (C.c_int, [P, _fp, C.c_int32, C.c_float, C.c_float, C.c_int32]),
        "casim_qs_set_stationary_model": (C.c_int, [P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                    C.c_double, C.c_double]),
        "casim_qs_get_window_stats": (C.c_int, [P, C.POINTER(C.c_double)]),
        "casim_qs_get_jsr": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "casim_qs_record_open_counts": (C.c_int, [P, C.c_double, C.c_int32]),
        "casim_qs_get_open_counts": (C.c_int, [P, C.c_void_p]),
        "casim_qs_get_mouths": (C.c_int, [P, _fp, _fp]),
        "casim_gemm_bf16": (C.c_int, [P, C.c_int32, C.c_int32, P, C.c_int32, C.c_int32, P, C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_float, P]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


def lib():
    l = abi.load()
    if not hasattr(l, "_mlp_declared"):
        _declare(l)
        l._mlp_declared = True
    return l


def fptr(a):
    return a.ctypes.data_as(_fp)


class MLP:
    """members x (n_in -> width x depth (SiLU) -> 1), BF16 MFMA GEMMs."""

    def __init__(self, n_in, width=256, depth=4, members=5, seed=1):
        self.lib = lib()
        self.n_in, self.width, self.depth, self.members = n_in, width, depth, members
        h = C.c_void_p()
        abi.check(self.lib.casim_mlp_create(n_in, width, depth, members, seed, C.byref(h)))
        self.h = h
        self.P = self.lib.casim_mlp_param_count(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.casim_mlp_destroy(self.h)
            self.h = None

    def params(self, member):
        out = np.empty(self.P, dtype=np.float32)
        abi.check(self.lib.casim_mlp_get_params(self.h, member, fptr(out)))
        return out

    def set_params(self, member, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        abi.check(self.lib.casim_mlp_set_params(self.h, member, fptr(p)))

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        m = np.ascontiguousarray(in_mean, dtype=np.float32)
        s = np.ascontiguousarray(in_std, dtype=np.float32)
        self.norm = (m.copy(), s.copy(), float(out_mean), float(out_std))
        abi.check(self.lib.casim_mlp_set_normalization(self.h, fptr(m), fptr(s), float(out_mean), float(out_std)))

    def fit_normalization(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        sd = x.std(axis=0)
        sd[sd == 0] = 1.0
        self.set_normalization(x.mean(axis=0), sd, float(np.mean(y)), float(np.std(y) or 1.0))

    def set_data(self, x, y):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        abi.check(self.lib.casim_mlp_set_data(self.h, fptr(self._x), fptr(self._y), len(self._y)))

    def train(self, epochs, batch=2048, lr=1e-3, seed=7):
        loss = np.zeros(self.members, dtype=np.float32)
        abi.check(self.lib.casim_mlp_train(self.h, epochs, batch, lr, seed, fptr(loss)))
        return loss

    def gradient(self, member, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(self.P, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        abi.check(self.lib.casim_mlp_gradient(self.h, member, fptr(x), fptr(y), len(y), fptr(g), fptr(loss)))
        return float(loss[0]), g

    def set_fused(self, on):
        """Fused inference kernel (default) or the per-layer GEMMs; same outputs."""
        abi.check(self.lib.casim_mlp_set_fused_inference(self.h, int(bool(on))))

    def predict(self, x):
        """(ensemble mean of exp(output), relative spread, per-member raw outputs)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        mean = np.empty(n, dtype=np.float32)
        spread = np.empty(n, dtype=np.float32)
        raw = np.empty((self.members, n), dtype=np.float32)
        abi.check(self.lib.casim_mlp_predict(self.h, fptr(x), n, fptr(mean), fptr(spread), fptr(raw)))
        return mean, spread, raw

    def save(self, path, meta=None, **extra):
        """Shape, normalisation, every member's FP32 parameters, an optional
        JSON-able description (training data, held-out errors) and extra
        arrays (.npz)."""
        import json
        m, s, ym, ys = getattr(self, "norm", (np.zeros(self.n_in), np.ones(self.n_in), 0.0, 1.0))
        np.savez(path, shape=np.array([self.n_in, self.width, self.depth, self.members]),
                 in_mean=m, in_std=s, out=np.array([ym, ys]),
                 params=np.stack([self.params(e) for e in range(self.members)]),
                 meta=np.array(json.dumps(meta or {})), **extra)

    @staticmethod
    def meta(path):
        import json
        z = np.load(path, allow_pickle=False)
        return json.loads(str(z["meta"])) if "meta" in z else {}

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        n_in, width, depth, members = (int(v) for v in z["shape"])
        net = cls(n_in, width, depth, members)
        net.set_normalization(z["in_mean"], z["in_std"], float(z["out"][0]), float(z["out"][1]))
        for e in range(members):
            net.set_params(e, z["params"][e])
        return net
