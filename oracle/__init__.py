"""The oracle: CPU FP64 restatement of the reference's RS sum, and the reference's own
CPU path compiled from its sources.

TEST INFRASTRUCTURE ONLY — imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package (celeris_amd)
never imports it.

* :func:`forward`, :func:`adjoint`, :func:`forward_ld` wrap oracle/rs_oracle.c
  (built to oracle/_build/liboracle_rs.so by oracle/Makefile).
* :func:`ref_run` runs oracle/_ref/ref_driver, the reference's focal.cpp /
  throughfocus.cpp / chromatic.cpp / mtf.cpp compiled by oracle/ref.mk.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from ctypes import POINTER, c_double, c_int, c_int32, c_int64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle_rs.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")


class orc_plane(ctypes.Structure):
    _fields_ = [
        ("k", c_double), ("x0", c_double), ("sx", c_double), ("y0", c_double),
        ("sy", c_double), ("z0", c_double), ("sz", c_double), ("nx", c_int32),
        ("ny", c_int32), ("tset", c_int32), ("reserved", c_int32), ("offset", c_int64),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        lib = ctypes.CDLL(LIB_PATH)
        dp = POINTER(c_double)
        lib.orc_forward.restype = c_int64
        lib.orc_forward.argtypes = [dp, dp, dp, c_int, POINTER(orc_plane), c_int, dp, dp, c_int, c_int]
        lib.orc_adjoint.restype = c_int64
        lib.orc_adjoint.argtypes = [dp, dp, c_int, c_int, POINTER(orc_plane), c_int, dp, dp, c_int, c_int]
        lib.orc_forward_ld.restype = c_int
        lib.orc_forward_ld.argtypes = [dp, dp, dp, c_int, POINTER(orc_plane), c_int, dp]
        lib.orc_mtf_autocorr.restype = None
        lib.orc_mtf_autocorr.argtypes = [dp, dp, c_int, c_int, dp, dp]
        lib.orc_num_cpus.restype = c_int
        _lib = lib
    return _lib


def _planes(planes):
    arr = (orc_plane * max(1, len(planes)))()
    cursor, total = 0, 0
    for i, p in enumerate(planes):
        off = cursor if p.offset < 0 else p.offset
        arr[i] = orc_plane(p.k, p.x0, p.sx, p.y0, p.sy, p.z0, p.sz, p.nx, p.ny, p.tset, 0, off)
        cursor = off + p.nx * p.ny
        total = max(total, cursor)
    return arr, total


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def forward(px, py, t, planes, threads: int = 0, row_stride: int = 1, intensity: bool = False):
    """Complex field over all planes (complex128, concatenated) [, |E|^2]; t: [nsets, npil]."""
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    t = np.ascontiguousarray(np.atleast_2d(t), np.complex128)
    arr, total = _planes(planes)
    field = np.zeros(total, np.complex128)
    inten = np.zeros(total, np.float64) if intensity else None
    pairs = load().orc_forward(_dp(px), _dp(py), _dp(t), px.size, arr, len(planes), _dp(field),
                               _dp(inten) if inten is not None else None, threads, row_stride)
    if pairs < 0:
        raise RuntimeError("orc_forward failed")
    return (field, inten) if intensity else field


def forward_pairs(px, py, t, planes, threads: int = 0, row_stride: int = 1) -> int:
    """Run the forward on a row sample and return the pairs evaluated (for timing)."""
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    t = np.ascontiguousarray(np.atleast_2d(t), np.complex128)
    arr, total = _planes(planes)
    field = np.zeros(total, np.complex128)
    return int(load().orc_forward(_dp(px), _dp(py), _dp(t), px.size, arr, len(planes),
                                  _dp(field), None, threads, row_stride))


def adjoint(px, py, u, planes, nsets: int = 1, threads: int = 0, pillar_stride: int = 1,
            return_pairs: bool = False):
    """g[s, q] = sum over planes of set s of conj(G_pq) u_p (complex128 [nsets, npil])."""
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    u = np.ascontiguousarray(u, np.complex128)
    arr, total = _planes(planes)
    if u.size < total:
        raise ValueError("u too short")
    g = np.zeros((nsets, px.size), np.complex128)
    pairs = load().orc_adjoint(_dp(px), _dp(py), px.size, nsets, arr, len(planes), _dp(u), _dp(g),
                               threads, pillar_stride)
    if pairs < 0:
        raise RuntimeError("orc_adjoint failed")
    return (g, int(pairs)) if return_pairs else g


def forward_ld(px, py, t, planes):
    """Long-double brute force of :func:`forward` (single thread; tiny inputs only)."""
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    t = np.ascontiguousarray(np.atleast_2d(t), np.complex128)
    arr, total = _planes(planes)
    field = np.zeros(total, np.complex128)
    load().orc_forward_ld(_dp(px), _dp(py), _dp(t), px.size, arr, len(planes), _dp(field))
    return field


def num_cpus() -> int:
    return int(load().orc_num_cpus())


# ----------------------------------------------------------------------------------
# The reference's own CPU path (oracle/_ref/ref_driver).

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def ref_run(command: str, values) -> np.ndarray:
    """Run `ref_driver command` on a flat float64 input; returns its float64 output."""
    if not ref_available():
        raise FileNotFoundError(REF_DRIVER)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.f64"), os.path.join(d, "out.f64")
        np.asarray(values, dtype=np.float64).tofile(fin)
        subprocess.run([REF_DRIVER, command, fin, fout], check=True)
        return np.fromfile(fout, dtype=np.float64)


def ref_propagate_pillars(px, py, t, cx, cy, z, wavelength, n, half_window) -> np.ndarray:
    """celeris::propagate_pillars (focal.cpp:205-238) from the reference build: |E|^2 (n, n)."""
    t = np.asarray(t, np.complex128)
    vals = np.concatenate([[px.size], px, py, np.stack([t.real, t.imag], 1).ravel(),
                           [cx, cy, z, wavelength, n, half_window]])
    return ref_run("pillars", vals).reshape(n, n)


def lens_block(n_cells, period, fill_map, lib_fill, lib_phase, lib_amp):
    return np.concatenate([[n_cells, period], np.asarray(fill_map, np.float64).ravel(),
                           [len(lib_fill)], lib_fill, lib_phase, lib_amp])


def mtf(pupil, pupil0, period_um, focal_um, wavelength_um, diameter_um):
    """analyze_mtf's curve (mtf.cpp:10-63) from an exit pupil P (complex n x n, zero
    outside the aperture) and its mask P0: (freq_cyc_mm, mtf, mtf_ideal) as float32
    and cutoff_cyc_mm, normalised and cast as the reference does (:51-62)."""
    P = np.ascontiguousarray(pupil, np.complex128)
    P0 = np.ascontiguousarray(pupil0, np.float64)
    n = P.shape[0]
    p = float(period_um)
    max_s = int(diameter_um / p) + 1
    ns = min(max_s + 1, n)
    acc = np.empty(2 * ns)
    acc0 = np.empty(ns)
    load().orc_mtf_autocorr(_dp(P.view(np.float64)), _dp(P0), n, ns, _dp(acc), _dp(acc0))
    a = np.abs(acc[0::2] + 1j * acc[1::2])
    a0 = np.abs(acc0)
    normP = a[0] if a[0] > 0 else 1.0
    normP0 = a0[0] if a0[0] > 0 else 1.0
    s = np.arange(ns)
    freq = ((s * p) / (wavelength_um * focal_um) * 1000.0).astype(np.float32)
    return freq, (a / normP).astype(np.float32), (a0 / normP0).astype(np.float32), \
        diameter_um / (wavelength_um * focal_um) * 1000.0


def ref_pgm(nx: int, ny: int, values, gamma: float) -> bytes:
    """The reference's write_pgm (io/image.cpp:10-37) output for a map, as bytes."""
    if not ref_available():
        raise FileNotFoundError(REF_DRIVER)
    vals = np.concatenate([[nx, ny, gamma], np.asarray(values, np.float64).ravel()])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.f64"), os.path.join(d, "out.f64")
        vals.tofile(fin)
        subprocess.run([REF_DRIVER, "pgm", fin, fout], check=True)
        with open(fout + ".pgm", "rb") as f:
            return f.read()
