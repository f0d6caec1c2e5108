"""ctypes binding of libceleris_rs.so (the C ABI declared in include/celeris_rs.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m celeris_amd.build``).  Loading fails loudly when it is missing: there
is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libceleris_rs.so")

MODE_FP32C = 0
MODE_FP64 = 1
MODE_FP32 = 2
MODES = {"fp32c": MODE_FP32C, "fp64": MODE_FP64, "fp32": MODE_FP32}

STATUS = {
    0: "CRS_OK",
    1: "CRS_ERR_NO_DEVICE",
    2: "CRS_ERR_INVALID",
    3: "CRS_ERR_ALLOC",
    4: "CRS_ERR_LAUNCH",
    5: "CRS_ERR_CAPTURE",
}


class crs_plane(ctypes.Structure):
    """Mirror of ``crs_plane`` (include/celeris_rs.h)."""

    _fields_ = [
        ("k", c_double),
        ("x0", c_double),
        ("sx", c_double),
        ("y0", c_double),
        ("sy", c_double),
        ("z0", c_double),
        ("sz", c_double),
        ("nx", c_int32),
        ("ny", c_int32),
        ("tset", c_int32),
        ("kernel", c_int32),
        ("offset", c_int64),
    ]


assert ctypes.sizeof(crs_plane) == 7 * 8 + 4 * 4 + 8


class crs_axis_scan_desc(ctypes.Structure):
    """Mirror of ``crs_axis_scan_desc`` (include/celeris_rs.h)."""

    _fields_ = [
        ("tset", c_int32),
        ("nz", c_int32),
        ("wavelength_um", c_double),
        ("zlo", c_double),
        ("zspan", c_double),
        ("z_default", c_double),
    ]


assert ctypes.sizeof(crs_axis_scan_desc) == 2 * 4 + 4 * 8

_DP = POINTER(c_double)

# (name, restype, argtypes) for every entry point of include/celeris_rs.h.
SIGNATURES = [
    ("crs_available", c_int, []),
    ("crs_device_name", c_char_p, []),
    ("crs_propagate_psf", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, c_double, c_int, c_double, _DP]),
    ("crs_propagate_zx", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_int, c_double, c_double, c_int, c_double, _DP]),
    ("crs_set_default_mode", c_int, [c_int]),
    ("crs_context_create", c_int, [c_int, POINTER(c_void_p)]),
    ("crs_context_destroy", None, [c_void_p]),
    ("crs_last_error", c_char_p, []),
    ("crs_forward", c_int,
     [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
      POINTER(crs_plane), c_int, c_void_p, c_void_p]),
    ("crs_adjoint", c_int,
     [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_int,
      POINTER(crs_plane), c_int, c_void_p, c_void_p]),
    ("crs_forward_noalloc", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, POINTER(crs_plane), c_int,
      c_void_p, c_void_p]),
    ("crs_set_tuning", c_int, [c_void_p, c_int, c_int, c_int]),
    ("crs_set_block_shape", c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    ("crs_graph_forward", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, POINTER(crs_plane), c_int,
      c_void_p, c_void_p, POINTER(c_void_p)]),
    ("crs_graph_adjoint", c_int,
     [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, POINTER(crs_plane), c_int, c_void_p,
      c_void_p, POINTER(c_void_p)]),
    ("crs_graph_launch", c_int, [c_void_p, c_void_p]),
    ("crs_graph_destroy", None, [c_void_p]),
    ("crs_forward_host", c_int,
     [c_int, _DP, _DP, _DP, c_int, c_int, POINTER(crs_plane), c_int, _DP, _DP]),
    ("crs_analyze_focus", c_int,
     [_DP, _DP, _DP, _DP, c_int, c_double, c_double, c_double, c_int, _DP, _DP]),
    ("crs_analyze_through_focus", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, c_int, c_int, _DP, _DP, _DP, _DP]),
    ("crs_analyze_chromatic", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, c_double, c_int, c_int, _DP]),
    ("crs_analyze_chromatic_dispersive", c_int,
     [_DP, _DP, _DP, _DP, c_int, c_double, c_double, c_double, c_double, c_int, c_int, _DP]),
    ("crs_analyze_field_grid", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, c_double, c_int, c_int, c_int, _DP]),
    ("crs_analyze_field_of_view", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, _DP, c_int, c_int, _DP]),
    ("crs_analyze_wide_fov", c_int,
     [_DP, _DP, _DP, c_int, c_double, c_double, c_double, c_double, c_double, _DP, c_int, c_int,
      _DP]),
    ("crs_analyze_mtf", c_int,
     [_DP, _DP, c_int, c_double, c_double, c_double, c_double, c_int, POINTER(ctypes.c_float),
      POINTER(ctypes.c_float), POINTER(ctypes.c_float), _DP, POINTER(c_int)]),
    ("crs_analyze_tolerance", c_int,
     [_DP, c_int, c_double, _DP, _DP, _DP, c_int, c_double, c_double, c_double, _DP, c_int,
      c_int, ctypes.c_uint, c_int, _DP, _DP]),
    ("crs_axis_scan", c_int,
     [_DP, _DP, _DP, c_int, c_int, POINTER(crs_axis_scan_desc), c_int, c_int, _DP, _DP, _DP]),
    ("crs_verify_achromatic_focus", c_int,
     [_DP, _DP, _DP, c_int, c_int, _DP, c_int, c_double, c_double, c_int, _DP]),
    ("crs_polar_focus_check", c_int,
     [_DP, _DP, _DP, _DP, c_int, c_double, c_double, c_double, c_int, c_int, _DP]),
    ("crs_set_timing", c_int, [c_void_p, c_int]),
    ("crs_last_kernel_ms", c_int, [c_void_p, c_int, POINTER(ctypes.c_float)]),
    ("crs_kernel_name", c_char_p, [c_int, c_int]),
    ("crs_last_launch", c_int, [c_void_p, c_int, POINTER(c_char_p), POINTER(c_int)]),
]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises LibraryMissing if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    # One HIP runtime per process: torch carries its own libamdhip64 with the same
    # soname as ROCm's, and whichever is loaded first serves every later user.  Load
    # torch's first (when torch is installed) so torch and this library share it.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().crs_last_error().decode()


def check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what} failed: {STATUS.get(status, status)}: {last_error()}")
