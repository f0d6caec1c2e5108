"""celeris_amd — MI355X-native Rayleigh–Sommerfeld far-field propagation for CELERIS.

The hot path of the reference's focal, through-focus and chromatic analyses
(src/celeris/cuda/propagate.{hpp,cu}, src/celeris/analysis/focal.cpp) rebuilt as
hand-written HIP kernels for gfx950 behind a C ABI (include/celeris_rs.h).
"""
from .rs import (  # noqa: F401
    KERNEL_REF,
    KERNEL_RSI,
    Context,
    Plane,
    available,
    device_name,
    kernel_name,
    pack_planes,
    propagate_psf,
    propagate_zx,
    set_default_mode,
)

__all__ = [
    "KERNEL_REF",
    "KERNEL_RSI",
    "Context",
    "Plane",
    "available",
    "device_name",
    "kernel_name",
    "pack_planes",
    "propagate_psf",
    "propagate_zx",
    "set_default_mode",
]
