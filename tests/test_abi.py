"""CPU checks of the C ABI: the library loads, exports what include/*.h declares, lays
out crs_plane as declared, and keeps the reference's no-device behaviour."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from celeris_amd import Plane, _lib
from celeris_amd import rs as crs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_c_functions():
    hdr = open(os.path.join(ROOT, "include", "celeris_rs.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(crs_[a-z_]+)\s*\(", hdr)))


def test_library_exports_every_declared_function():
    lib = _lib.load()
    names = declared_c_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound


def test_library_exports_reference_cpp_names():
    """include/celeris/cuda/propagate.hpp: the reference's celeris::cuda names."""
    out = subprocess.run(["nm", "-DC", _lib.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    for sig in ("celeris::cuda::available()", "celeris::cuda::device_name()",
                "celeris::cuda::propagate_psf(double const*, double const*, std::complex<double> const*, int, double, double, double, double, int, double, double*)",
                "celeris::cuda::propagate_zx(double const*, double const*, std::complex<double> const*, int, double, double, int, double, double, int, double, double*)"):
        assert sig in out, sig


def test_dropin_header_matches_reference_signatures():
    hdr = open(os.path.join(ROOT, "include", "celeris", "cuda", "propagate.hpp")).read()
    flat = " ".join(hdr.split())
    assert "namespace celeris::cuda" in flat
    assert ("bool propagate_psf(const double* px, const double* py, const std::complex<double>* pt, "
            "int npil, double cx, double cy, double z, double k, int n, double half_window, double* out);") in flat
    assert ("bool propagate_zx(const double* px, const double* py, const std::complex<double>* pt, "
            "int npil, double xmin, double xmax, int nx, double zmin, double zmax, int nz, double k, "
            "double* out);") in flat


def test_plane_struct_layout():
    P = _lib.crs_plane
    assert ctypes.sizeof(P) == 80
    assert P.k.offset == 0 and P.nx.offset == 56 and P.tset.offset == 64 and P.offset.offset == 72


def test_plane_geometry_follows_reference_grid():
    p = Plane.psf(0.5, -0.25, 100.0, 9.9, 161, 5.4)
    assert p.x0 == 0.5 - 5.4 and p.sx == 2 * 5.4 / 160 and p.sz == 0.0
    # n == 1: step = 2W/0 = inf, as propagate.cu:143 and focal.cpp:223 compute it
    assert math.isinf(Plane.psf(0, 0, 10, 10, 1, 1.0).sx)
    zx = Plane.zx(-3, 3, 201, 50, 150, 101, 10.0)
    assert zx.sy == 0 and zx.y0 == 0 and zx.sz == 1.0 and zx.sx == 6 / 200


def test_kernel_names():
    assert crs.kernel_name("fp32c") == "fwd_fp32c"
    assert crs.kernel_name("fp32c", adjoint=True) == "adj_fp32c"
    assert crs.kernel_name("fp64") == "fwd_fp64"


def _has_gpu():
    try:
        return bool(_lib.load().crs_available())
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device path")
def test_no_device_behaviour():
    """No device: available() false, device_name() "", propagate_* false with out
    untouched (propagate.cu:18-31, :137-138, :199-200)."""
    assert crs.available() is False
    assert crs.device_name() == ""
    px = np.zeros(3)
    t = np.ones(3, complex)
    assert crs.propagate_psf(px, px, t, 0, 0, 10, 10, 5, 1.0) is None
    assert crs.propagate_zx(px, px, t, -1, 1, 5, 5, 10, 5, 10) is None
    lib = _lib.load()
    out = (ctypes.c_double * 4)(7, 7, 7, 7)
    dp = ctypes.POINTER(ctypes.c_double)
    pxa = (ctypes.c_double * 3)()
    ta = (ctypes.c_double * 6)(1, 0, 1, 0, 1, 0)
    assert lib.crs_propagate_psf(pxa, pxa, ta, 3, 0.0, 0.0, 10.0, 10.0, 2, 1.0, out) == 0
    assert list(out) == [7, 7, 7, 7]
    h = ctypes.c_void_p()
    assert lib.crs_context_create(0, ctypes.byref(h)) == 1  # CRS_ERR_NO_DEVICE
    assert "device" in _lib.last_error()


def test_set_default_mode_roundtrip():
    prev = crs.set_default_mode("fp64")
    assert crs.set_default_mode(prev) == "fp64"
    assert _lib.load().crs_set_default_mode(7) == -1


def test_host_arrays_of_unequal_length_are_refused():
    """The C entry points read npil elements of every per-pillar array; the Python
    layer refuses mismatched lengths before calling them."""
    from celeris_amd import analysis

    x = np.zeros(4)
    with pytest.raises(ValueError, match="one length"):
        analysis.propagate_pillars(x, x, np.ones(3, complex), 0, 0, 50.0, 0.5, 9, 2.0)
    with pytest.raises(ValueError, match="one length"):
        analysis.analyze_focus_pillars(x, x, np.ones(4, complex), np.zeros(5), 50.0, 0.5, 20.0)
    with pytest.raises(ValueError, match="one length"):
        analysis.polar_focus_check(x, np.zeros(3), np.ones(4, complex), np.ones(4, complex),
                                   0.5, 50.0, 60.0)
    with pytest.raises(ValueError, match="npil"):
        analysis.axis_scan(x, x, np.ones((2, 3), complex), [(0, 10, 0.5, 10.0, 20.0, 10.0)])
    with pytest.raises(ValueError, match="one length"):
        crs.propagate_psf(x, x, np.ones(3, complex), 0, 0, 50.0, 12.0, 9, 2.0)


def test_forward_host_checks_arguments_before_device_work():
    """crs_forward_host refuses bad sizes and planes with CRS_ERR_INVALID (2) before it
    sizes any buffer from them, with or without a device; valid arguments without a
    device give CRS_ERR_NO_DEVICE (1)."""
    lib = _lib.load()
    px = np.zeros(3)
    t = np.ones(3, complex)
    out = np.zeros(25)
    dp = ctypes.POINTER(ctypes.c_double)
    good = crs.pack_planes([Plane.psf(0, 0, 10.0, 10.0, 5, 1.0)])[0]
    bad_tset = crs.pack_planes([Plane.psf(0, 0, 10.0, 10.0, 5, 1.0, tset=1)])[0]

    def call(npil, nsets, planes):
        return lib.crs_forward_host(0, px.ctypes.data_as(dp), px.ctypes.data_as(dp),
                                    t.ctypes.data_as(dp), npil, nsets, planes, 1, None,
                                    out.ctypes.data_as(dp))

    assert call(-5, 1, good) == 2
    assert call(3, -1, good) == 2
    assert call(3, 1, bad_tset) == 2
    assert "tset" in _lib.last_error()
    assert call(3, 1, good) == (0 if crs.available() else 1)


def test_analysis_entry_points_report_invalid_arguments():
    """The analysis entry points clear the thread's last error on entry and name what
    was wrong when they refuse their arguments."""
    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    x = np.zeros(4)
    out = np.zeros(8)
    lib.crs_forward_host(0, x.ctypes.data_as(dp), x.ctypes.data_as(dp), x.ctypes.data_as(dp),
                         -1, 1, None, 0, None, out.ctypes.data_as(dp))  # leaves a message
    assert lib.crs_analyze_focus(x.ctypes.data_as(dp), x.ctypes.data_as(dp), x.ctypes.data_as(dp),
                                 x.ctypes.data_as(dp), 0, 50.0, 0.5, 20.0, 0,
                                 out.ctypes.data_as(dp), None) == 2
    assert "crs_analyze_focus" in _lib.last_error()
    assert lib.crs_polar_focus_check(x.ctypes.data_as(dp), x.ctypes.data_as(dp),
                                     x.ctypes.data_as(dp), x.ctypes.data_as(dp), 2, 0.5, 50.0,
                                     60.0, 0, 0, out.ctypes.data_as(dp)) == 2
    assert "nz < 1" in _lib.last_error()
