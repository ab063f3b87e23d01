"""MI355X-native dyadic-cleft reaction-diffusion for CaSim release sites.

The compute path is libcasim_dyad.so (hand-written HIP for gfx950) behind the
C ABI of include/casim_dyad.h; this package holds its ctypes binding, the
DyadSeries-shaped host class, JSON configuration and synthetic workloads.
"""
from . import abi, config, rng, workloads  # noqa: F401

__all__ = ["abi", "config", "rng", "workloads"]
