"""Small configurations for closed-form and parity tests."""
import numpy as np

from casim_amd import abi, rng
from casim_amd.config import Buffer, ChannelType, DyadSpec, Ion


def spec(dim=3, nx=21, ny=17, nz=5, h=4.0, ions=None, buffers=None, sl=None, channels=None, n_sites=1,
         clamped=True, precision=abi.CASIM_FP64, g=28000.0, radius=1000.0, z=None, first_stream=0,
         equilibrium=True, model="SheepCannell", params=None):
    ions = ions if ions is not None else [Ion("Calcium", 220000.0, 0.1)]
    buffers = buffers if buffers is not None else []
    if channels is None:
        channels = [(h * (nx // 2), h * (ny // 2))]
    if params is None:
        params = ({"kopen0": 0.0, "ncaopen": 2.0, "kopenmax": 1e300, "kclose0": 0.0, "ncaclose": 0.0} if clamped
                  else {"kopen0": 0.2 / 225.0 * 1e9, "ncaopen": 2.0, "kopenmax": 1e300, "kclose0": 500.0,
                        "ncaclose": 0.0})
    ct = ChannelType("RyR", model, params, [g] + [0.0] * (len(ions) - 1))
    zext = z if z is not None else (h * (nz - 1) if dim == 3 else 15.0)
    return DyadSpec(dim=dim, nx=nx, ny=ny, nz=nz if dim == 3 else 1, dx=h, dy=h, dz=h, z=zext, radius=radius,
                    voltage=0.0, ions=ions, buffers=buffers, sl=sl or {}, start_equilibrium=equilibrium,
                    channel_types=[ct], sites=[[(x, y, 0) for x, y in channels] for _ in range(n_sites)],
                    rng_states=rng.stream_states(rng.DEFAULT_SEED, n_sites, first=first_stream),
                    precision=precision, init_channel=[0] * n_sites)


def grid_coords(sp):
    x = np.arange(sp.nx)[:, None, None]
    y = np.arange(sp.ny)[None, :, None]
    z = np.arange(sp.nz)[None, None, :]
    return np.broadcast_arrays(x, y, z)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0


class DeviceBuffer:
    """Raw device memory through the HIP runtime (for kernel-level tests)."""

    def __init__(self, host):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so")
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.host.nbytes, 1))) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.host.nbytes), 1) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.ptr, self.C.c_size_t(out.nbytes), 2) == 0
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = None
