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
import math
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
        ("ny", c_int32), ("tset", c_int32), ("kernel", c_int32), ("offset", c_int64),
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
        lib.orc_mt19937_nth.restype = ctypes.c_uint32
        lib.orc_mt19937_nth.argtypes = [ctypes.c_uint32, c_int64]
        lib.orc_tolerance_fills.restype = None
        lib.orc_tolerance_fills.argtypes = [ctypes.c_uint32, c_int, c_double, c_double, dp, c_int64, dp]
        _lib = lib
    return _lib


def _planes(planes):
    arr = (orc_plane * max(1, len(planes)))()
    cursor, total = 0, 0
    for i, p in enumerate(planes):
        off = cursor if p.offset < 0 else p.offset
        arr[i] = orc_plane(p.k, p.x0, p.sx, p.y0, p.sy, p.z0, p.sz, p.nx, p.ny, p.tset,
                           getattr(p, "kernel", 0), off)
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


class _AxisPoint:
    """One on-axis sample as a 1 x 1 plane at exactly z (the scans' own z values)."""

    def __init__(self, k, z, tset):
        self.k, self.x0, self.sx, self.y0, self.sy = k, 0.0, 0.0, 0.0, 0.0
        self.z0, self.sz, self.nx, self.ny, self.tset, self.offset = z, 0.0, 1, 1, tset, -1


def axis_intensity(px, py, t_sets, points):
    """I(z) = |sum_q t_q polar(1/R, kR)|^2 at (0, 0, z) for points = [(k, z, tset)], the
    on-axis sums of achromatic.cpp:306-315, pb_achromatic.cpp:275-284,
    cmd_polar.cpp:88-100 / :113-119 and cli_common.cpp:115-123 (R = sqrt(x^2 + y^2 + z^2),
    pillars in list order)."""
    planes = [_AxisPoint(k, z, s) for k, z, s in points]
    f = forward(px, py, t_sets, planes)
    return f.real * f.real + f.imag * f.imag


def axis_scan(px, py, t_sets, scans):
    """scans = [(tset, nz, wavelength, zlo, zspan, z_default)]: samples
    z_j = zlo + zspan * j / (nz - 1) (zlo if nz == 1) and the callers' argmax, the first
    sample with I > every earlier one and > -1 (achromatic.cpp:323-328); returns
    (best_z, best_I, curves) like crs_axis_scan."""
    pts, zs = [], []
    for tset, nz, lam, zlo, zspan, _ in scans:
        k = 2.0 * math.pi / lam
        z = [zlo + zspan * j / (nz - 1) if nz > 1 else zlo for j in range(nz)]
        zs.append(z)
        pts += [(k, zz, tset) for zz in z]
    I = axis_intensity(px, py, t_sets, pts)
    best_z, best_I, off = [], [], 0
    for (tset, nz, lam, zlo, zspan, zdef), z in zip(scans, zs):
        bz, bi = zdef, -1.0
        for j in range(nz):
            if I[off + j] > bi:
                bi, bz = I[off + j], z[j]
        best_z.append(bz)
        best_I.append(bi)
        off += nz
    return np.array(best_z), np.array(best_I), I


def verify_achromatic_focus(px, py, t_bands, wavelengths, center_index, center_wavelength, f):
    """verify_achromatic_focus (achromatic.cpp:288-338) / verify_pb_achromatic_focus
    (pb_achromatic.cpp:255-301) on the aperture cells' per-band transmissions t_bands
    [nb, npil]: nb x (lambda, best z over 200 samples in [0.4 f, 1.8 f], peak / I(f) of
    the centre band at the centre wavelength)."""
    design_peak = axis_intensity(px, py, t_bands, [(2.0 * math.pi / center_wavelength, f,
                                                     center_index)])[0]
    zlo, zhi = 0.4 * f, 1.8 * f
    scans = [(j, 200, lam, zlo, zhi - zlo, f) for j, lam in enumerate(wavelengths)]
    bz, bi, _ = axis_scan(px, py, t_bands, scans)
    return np.array([[lam, z, i / design_peak if design_peak > 0 else 0.0]
                     for lam, z, i in zip(wavelengths, bz, bi)])


def chromatic(px, py, t_sets, design_set, band_sets, design_focal, design_wavelength,
              lambda_min, lambda_max, n):
    """analyze_chromatic (chromatic.cpp:9-58; every band_sets entry the same set) and
    analyze_chromatic_dispersive (chromatic.cpp:60-127; one set per wavelength): the design
    peak I(f0) at lambda0 from set design_set, then per wavelength lambda_i = lmin + (lmax -
    lmin) i / (n - 1) the first maximum of 120 on-axis samples over [0.55, 1.6] f0 lambda0 /
    lambda_i (default z: the expected focus).  Returns n x (lambda, best z, peak / design peak)."""
    design_peak = axis_intensity(px, py, t_sets, [(2.0 * math.pi / design_wavelength, design_focal,
                                                   design_set)])[0]
    scans, lams = [], []
    for i in range(n):
        lam = lambda_min + (lambda_max - lambda_min) * i / (n - 1)
        expected = design_focal * design_wavelength / lam
        zlo, zhi = 0.55 * expected, 1.6 * expected
        scans.append((band_sets[i], 120, lam, zlo, zhi - zlo, expected))
        lams.append(lam)
    bz, bi, _ = axis_scan(px, py, t_sets, scans)
    return np.array([[lam, z, i / design_peak if design_peak > 0 else 0.0]
                     for lam, z, i in zip(lams, bz, bi)])


def polar_focus_check(px, py, tx, ty, wavelength, fx, fy, nz=240):
    """The polarisation check of cmd_polar.cpp:85-126 (nz = 240; gui/workers.cpp:181-197
    uses 200): each channel's on-axis focus over [0.5 min(fx, fy), 1.5 max(fx, fy)] and
    the focal isolations 10 log10(I_own(z) / max(I_other(z), 1e-30)) there (0 when
    |fx - fy| <= 1e-6).  Returns (z_x, z_y, iso_x, iso_y)."""
    zlo, zhi = 0.5 * min(fx, fy), 1.5 * max(fx, fy)
    t2 = np.stack([np.asarray(tx, np.complex128), np.asarray(ty, np.complex128)])
    (zx, zy), _, _ = axis_scan(px, py, t2, [(0, nz, wavelength, zlo, zhi - zlo, zlo),
                                            (1, nz, wavelength, zlo, zhi - zlo, zlo)])
    iso_x = iso_y = 0.0
    if abs(fx - fy) > 1e-6:
        k = 2.0 * math.pi / wavelength
        a = axis_intensity(px, py, t2, [(k, zx, 0), (k, zx, 1), (k, zy, 1), (k, zy, 0)])
        iso_x = 10.0 * math.log10(a[0] / max(a[1], 1e-30))
        iso_y = 10.0 * math.log10(a[2] / max(a[3], 1e-30))
    return zx, zy, iso_x, iso_y


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



def mt19937_nth(seed: int, nth: int) -> int:
    """The nth output of the oracle's MT19937 restatement seeded with seed."""
    return int(load().orc_mt19937_nth(seed & 0xFFFFFFFF, nth))


def tolerance_fills(seed: int, trial: int, sigma_nm: float, period_um: float, fill_map):
    """One trial's jittered fill map of analyze_tolerance (tolerance.cpp:20-31)."""
    f = np.ascontiguousarray(np.asarray(fill_map, np.float64).ravel())
    out = np.empty_like(f)
    load().orc_tolerance_fills(seed & 0xFFFFFFFF, int(trial), float(sigma_nm), float(period_um),
                               _dp(f), f.size, _dp(out))
    return out


def tolerance(n_cells, period, fill_map, lib_fill, lib_phase, lib_amp, f, lam, D, sigmas_nm,
              trials, seed, return_trials=False):
    """analyze_tolerance (tolerance.cpp:12-48) on the oracle: per sigma and trial the
    jittered map, the aperture cells' nearest-fill transmissions (metalens.cpp:112-120,
    first on ties), the 121 x 121 grid at z = f (focal.cpp:61-64, :90-99) and the ideal
    peak (:102), Strehl = peak / peak_ideal (:140); then mean, population std and worst
    in trial order (:37-46)."""
    c = (n_cells - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="ij")
    x, y = ((ix - c) * period).ravel(), ((iy - c) * period).ravel()
    r = np.sqrt(x * x + y * y)
    keep = ~(r > D / 2.0)
    px, py, rk = x[keep], y[keep], r[keep]
    k = 2.0 * math.pi / lam
    phi = -k * (np.sqrt(rk * rk + f * f) - f)
    dl = lam * f / D
    W = max(6.0 * dl, 4.0)
    NG = 121

    class P:
        pass

    grid = P()
    grid.k, grid.x0, grid.sx, grid.y0, grid.sy = k, -W, 2.0 * W / (NG - 1), -W, 2.0 * W / (NG - 1)
    grid.z0, grid.sz, grid.nx, grid.ny, grid.tset, grid.offset = f, 0.0, NG, NG, 0, -1
    ideal = axis_intensity(px, py, np.cos(phi) + 1j * np.sin(phi), [(k, f, 0)])[0]
    lf = np.asarray(lib_fill, np.float64)
    n = max(1, int(trials))
    out, per = [], []
    for s in sigmas_nm:
        st = []
        for t in range(n):
            jf = tolerance_fills(seed, t, s, period, fill_map)[keep]
            e = np.array([_nearest(lf, v) for v in jf])
            tt = np.asarray(lib_amp)[e] * np.cos(np.asarray(lib_phase)[e]) + \
                1j * (np.asarray(lib_amp)[e] * np.sin(np.asarray(lib_phase)[e]))
            _, I = forward(px, py, tt, [grid], intensity=True)
            st.append(I.max() / ideal if ideal > 0 else 0.0)
        s1 = s2 = 0.0
        worst = 1e300
        for v in st:
            s1 += v
            s2 += v * v
            worst = min(worst, v)
        mean = s1 / n
        out.append((s, mean, math.sqrt(max(0.0, s2 / n - mean * mean)), worst))
        per.append(st)
    return (np.array(out), np.array(per)) if return_trials else np.array(out)


def _nearest(lib_fill, v):
    best, best_d = 0, abs(lib_fill[0] - v)
    for i in range(1, lib_fill.size):
        d = abs(lib_fill[i] - v)
        if d < best_d:
            best, best_d = i, d
    return best