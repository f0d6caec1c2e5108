"""The reference's focal-plane analyses, on the MI355X propagation.

Same names, arguments and result fields as the reference's C++ API and its Python
module (src/celeris/analysis/*.hpp; bindings/python/celeris_module.cpp:415-440):

* :func:`analyze_focus`         focal.cpp:32-146      -> :class:`FocalAnalysis`
* :func:`compute_psf`           focal.cpp:148-203     -> :class:`PsfMap`
* :func:`propagate_pillars`     focal.cpp:205-238     -> :class:`PsfMap`
* :func:`compute_psf_field`     focal.cpp:240-311     -> :class:`FieldPsf`
* :func:`analyze_through_focus` throughfocus.cpp:15-132 -> :class:`ThroughFocus`
* :func:`analyze_chromatic`     chromatic.cpp:9-58    -> list of :class:`ChromaticPoint`
* :func:`analyze_chromatic_dispersive`  chromatic.cpp:60-127  -> list of :class:`ChromaticPoint`
* :func:`analyze_field_grid`    field.cpp:150-233     -> :class:`FieldGrid`
* :func:`analyze_field_of_view` field.cpp:12-76       -> list of :class:`FieldPoint`
* :func:`analyze_wide_fov`      field.cpp:78-148      -> list of :class:`FieldPoint`
* :func:`analyze_mtf`           mtf.cpp:10-63         -> :class:`MtfCurve`
* :func:`analyze_tolerance`     tolerance.cpp:12-48   -> list of :class:`ToleranceResult`
* :func:`verify_achromatic_focus`    design/achromatic.cpp:288-338    -> list of
  :class:`AchromaticFocalPoint`; :func:`verify_pb_achromatic_focus`
  design/pb_achromatic.cpp:255-301 (same)
* :func:`polar_focus_check`     cli/cmd_polar.cpp:85-126, gui/workers.cpp:181-197
  -> :class:`PolarFocus`
* :func:`profile_focus_z`       cli/cli_common.cpp:112-133 (focusing profiles)
* :func:`axis_scan`             the batched on-axis scan under the last three

Each runs its sums in one launch through the C ABI (crs_analyze_*, crs_forward_host);
where the reference looped on the CPU (on-axis scans, the ideal peak) the sums are on
the GPU too.  ``mode`` selects the arithmetic ("fp32c" default, "fp64", "fp32").
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Callable, List

import numpy as np

from . import _lib
from .design import (AchromaticDesign, DispersiveLibrary, DispersivePbLibrary, MetalensDesign,
                     PbAchromaticDesign, UnitCellLibrary, achromatic_band_transmissions,
                     aperture_fills, aperture_pillars, exit_pupil,
                     pb_achromatic_band_transmissions)
from .rs import _mode, crs_plane

_DP = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_DP)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pillars(px, py, *per_pillar):
    """px, py as contiguous float64 and each per-pillar array as contiguous complex128
    (transmissions) or float64 (phases), all 1-D of one length: the C entry points read
    npil elements of each."""
    px, py = _d(px), _d(py)
    out = [px, py]
    for a in per_pillar:
        a = np.asarray(a)
        out.append(_c(a) if np.iscomplexobj(a) else _d(a))
    if any(a.ndim != 1 or a.size != px.size for a in out):
        raise ValueError("px, py and the per-pillar arrays must be 1-D arrays of one length")
    return out


@dataclass
class FocalAnalysis:
    strehl: float
    phase_strehl: float
    fwhm_um: float
    diffraction_limit_um: float
    encircled_energy: float
    intensity: np.ndarray = None  # the 121 x 121 |E|^2 grid the metrics come from


@dataclass
class PsfMap:
    n: int
    half_window_um: float
    intensity: np.ndarray  # (n, n), |E|^2, row j = y


@dataclass
class FieldPsf:
    psf: PsfMap
    angle_deg: float
    cx_um: float
    rel_strehl: float


@dataclass
class ThroughFocus:
    z_um: np.ndarray          # float32, 161 axial positions
    intensity: np.ndarray     # float32, on-axis |E|^2 / peak
    z_peak_um: float
    dof_um: float
    caustic_nx: int = 0
    caustic_nz: int = 0
    caustic_xmin: float = 0.0
    caustic_xmax: float = 0.0
    caustic_zmin: float = 0.0
    caustic_zmax: float = 0.0
    caustic: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))


@dataclass
class ChromaticPoint:
    wavelength_um: float
    focal_length_um: float
    rel_peak: float


@dataclass
class FieldGridPoint:
    theta_x_deg: float
    theta_y_deg: float
    rel_strehl: float
    fwhm_x_um: float
    fwhm_y_um: float
    cx_um: float
    cy_um: float


@dataclass
class FieldGrid:
    n: int
    max_angle_deg: float
    points: List[FieldGridPoint]


@dataclass
class FieldPoint:
    """field.hpp:14-18."""

    angle_deg: float
    rel_strehl: float
    spot_shift_um: float


@dataclass
class MtfCurve:
    """mtf.hpp: spatial frequency (cycles/mm), lens MTF and diffraction-limited MTF
    (float32, as the reference's vectors of float) and the cutoff frequency."""

    freq_cyc_mm: np.ndarray
    mtf: np.ndarray
    mtf_ideal: np.ndarray
    cutoff_cyc_mm: float


def _check(status, what):
    _lib.check(status, what)


def analyze_focus(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                  wavelength_um: float, diameter_um: float, mode="fp32c") -> FocalAnalysis:
    px, py, t, phi = aperture_pillars(lens, lib, diameter_um, focal_length_um, wavelength_um)
    return analyze_focus_pillars(px, py, t, phi, focal_length_um, wavelength_um, diameter_um, mode)


def analyze_focus_pillars(px, py, t, phi_ideal, focal_length_um, wavelength_um, diameter_um,
                          mode="fp32c") -> FocalAnalysis:
    px, py, t, phi = _pillars(px, py, _c(t), _d(phi_ideal))
    out = np.zeros(5)
    grid = np.zeros((121, 121))
    _check(_lib.load().crs_analyze_focus(_p(px), _p(py), _p(t), _p(phi), px.size,
                                         focal_length_um, wavelength_um, diameter_um,
                                         _mode(mode), _p(out), _p(grid)), "crs_analyze_focus")
    return FocalAnalysis(*map(float, out), intensity=grid)


def propagate_pillars(px, py, t, cx, cy, z, wavelength_um, n, half_window_um,
                      mode="fp32c") -> PsfMap:
    """focal.cpp:205-238: |E|^2 on an n x n window; an empty list or n <= 0 gives zeros."""
    px, py, t = _pillars(px, py, _c(t))
    n = int(n)
    m = PsfMap(n, half_window_um, np.zeros((max(n, 0), max(n, 0))))
    if px.size == 0 or n <= 0:  # focal.cpp:216
        return m
    k = 2.0 * math.pi / wavelength_um
    with np.errstate(divide="ignore", invalid="ignore"):
        step = float(np.float64(2.0 * half_window_um) / np.float64(n - 1))
    plane = crs_plane(k, cx - half_window_um, step, cy - half_window_um, step, z, 0.0, n, n, 0,
                      0, 0)
    _check(_lib.load().crs_forward_host(_mode(mode), _p(px), _p(py), _p(t), px.size, 1,
                                        ctypes.byref(plane), 1, None, _p(m.intensity)),
           "crs_forward_host")
    return m


def compute_psf(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                wavelength_um: float, diameter_um: float, n: int, half_window_um: float,
                mode="fp32c") -> PsfMap:
    """focal.cpp:148-203: the design's PSF at its focal plane."""
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    return propagate_pillars(px, py, t, 0.0, 0.0, focal_length_um, wavelength_um, n,
                             half_window_um, mode)


def compute_psf_field(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                      wavelength_um: float, diameter_um: float, angle_deg: float, n: int,
                      half_window_um: float, on_axis_peak: float = 0.0,
                      mode="fp32c") -> FieldPsf:
    """focal.cpp:240-311: tilted plane wave exp(i k sin(theta) x), window centred on the
    chief ray x = f tan(theta)."""
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    k = 2.0 * math.pi / wavelength_um
    theta = angle_deg * math.pi / 180.0
    tt = t * (np.cos(k * math.sin(theta) * px) + 1j * np.sin(k * math.sin(theta) * px))
    cx = focal_length_um * math.tan(theta)
    psf = propagate_pillars(px, py, tt, cx, 0.0, focal_length_um, wavelength_um, n,
                            half_window_um, mode)
    peak = float(psf.intensity.max()) if psf.intensity.size else 0.0
    rel = peak / on_axis_peak if on_axis_peak > 0 else 1.0
    return FieldPsf(psf, angle_deg, cx, rel)


def analyze_through_focus(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                          wavelength_um: float, diameter_um: float, caustic_n: int = 201,
                          mode="fp32c") -> ThroughFocus:
    """throughfocus.cpp:15-132.  caustic_n = 201 is the reference's GPU-path caustic
    (throughfocus.cpp:128), 81 its CPU-path one (:140)."""
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    return analyze_through_focus_pillars(px, py, t, focal_length_um, wavelength_um, diameter_um,
                                         caustic_n, mode)


def analyze_through_focus_pillars(px, py, t, focal_length_um, wavelength_um, diameter_um,
                                  caustic_n=201, mode="fp32c") -> ThroughFocus:
    px, py, t = _pillars(px, py, _c(t))
    z = np.zeros(161)
    inten = np.zeros(161)
    sc = np.zeros(6)
    caustic = np.zeros(caustic_n * caustic_n)
    _check(_lib.load().crs_analyze_through_focus(_p(px), _p(py), _p(t), px.size, focal_length_um,
                                                 wavelength_um, diameter_um, int(caustic_n),
                                                 _mode(mode), _p(z), _p(inten), _p(sc),
                                                 _p(caustic)), "crs_analyze_through_focus")
    return ThroughFocus(z.astype(np.float32), inten.astype(np.float32), float(sc[0]), float(sc[1]),
                        caustic_n, caustic_n, float(sc[2]), float(sc[3]), float(sc[4]),
                        float(sc[5]), caustic.astype(np.float32))


def analyze_chromatic(lens: MetalensDesign, lib: UnitCellLibrary, design_focal: float,
                      design_wavelength: float, diameter: float, lambda_min: float,
                      lambda_max: float, n: int, mode="fp32c") -> List[ChromaticPoint]:
    """chromatic.cpp:9-58 (fixed transmission, geometric-phase model)."""
    px, py, t = aperture_pillars(lens, lib, diameter)
    return analyze_chromatic_pillars(px, py, t, design_focal, design_wavelength, lambda_min,
                                     lambda_max, n, mode)


def analyze_chromatic_pillars(px, py, t, design_focal, design_wavelength, lambda_min,
                              lambda_max, n, mode="fp32c") -> List[ChromaticPoint]:
    px, py, t = _pillars(px, py, _c(t))
    out = np.zeros(3 * n)
    _check(_lib.load().crs_analyze_chromatic(_p(px), _p(py), _p(t), px.size, design_focal,
                                             design_wavelength, lambda_min, lambda_max, int(n),
                                             _mode(mode), _p(out)), "crs_analyze_chromatic")
    return [ChromaticPoint(*map(float, out[3 * i:3 * i + 3])) for i in range(n)]


def analyze_chromatic_dispersive(lens: MetalensDesign, build_lib_at: Callable[[float], UnitCellLibrary],
                                 design_focal: float, design_wavelength: float, diameter: float,
                                 lambda_min: float, lambda_max: float, n: int,
                                 mode="fp32c") -> List[ChromaticPoint]:
    """chromatic.cpp:60-127: the unit-cell library is rebuilt at every wavelength by
    build_lib_at(lambda) (the caller's library builder; the RCWA sweep behind it stays on
    the CPU), each aperture cell takes its fill's transmission from that library, and the
    n x 120 on-axis samples plus the design peak run in one launch.  build_lib_at is called
    as the reference calls it: at the design wavelength first, then at each lambda_i."""
    if n < 2:
        raise ValueError("n must be at least 2")
    px, py, t0 = aperture_pillars(lens, build_lib_at(design_wavelength), diameter)
    fills = aperture_fills(lens, diameter)
    bands = []
    for i in range(n):
        lam = lambda_min + (lambda_max - lambda_min) * i / (n - 1)
        bands.append(build_lib_at(lam).transmission_for_fill(fills))
    return analyze_chromatic_dispersive_pillars(px, py, t0, np.array(bands), design_focal,
                                                design_wavelength, lambda_min, lambda_max, n, mode)


def analyze_chromatic_dispersive_pillars(px, py, t_design, t_bands, design_focal, design_wavelength,
                                         lambda_min, lambda_max, n, mode="fp32c") -> List[ChromaticPoint]:
    """The dispersive chromatic scan on given pillars: t_design (npil) at the design
    wavelength, t_bands (n x npil) at lambda_min + (lambda_max - lambda_min) i / (n - 1)."""
    px, py, t0 = _pillars(px, py, _c(t_design))
    tb = np.ascontiguousarray(_c(t_bands).reshape(n, px.size))
    out = np.zeros(3 * n)
    _check(_lib.load().crs_analyze_chromatic_dispersive(
        _p(px), _p(py), _p(t0), _p(tb), px.size, design_focal, design_wavelength, lambda_min,
        lambda_max, int(n), _mode(mode), _p(out)), "crs_analyze_chromatic_dispersive")
    return [ChromaticPoint(*map(float, out[3 * i:3 * i + 3])) for i in range(n)]


def analyze_field_grid(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                       wavelength_um: float, diameter_um: float, max_angle_deg: float,
                       n_half: int, psf_n: int, mode="fp32c") -> FieldGrid:
    """field.cpp:150-233: all (2 n_half + 1)^2 field points plus the on-axis reference
    in one launch."""
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    px, py, t = _pillars(px, py, _c(t))
    g = 2 * n_half + 1
    out = np.zeros(7 * g * g)
    _check(_lib.load().crs_analyze_field_grid(_p(px), _p(py), _p(t), px.size, focal_length_um,
                                              wavelength_um, diameter_um, max_angle_deg,
                                              int(n_half), int(psf_n), _mode(mode), _p(out)),
           "crs_analyze_field_grid")
    pts = [FieldGridPoint(*map(float, out[7 * i:7 * i + 7])) for i in range(g * g)]
    return FieldGrid(g, max_angle_deg, pts)


def analyze_mtf(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                wavelength_um: float, diameter_um: float) -> MtfCurve:
    """mtf.cpp:10-63: MTF along x from the autocorrelation of the exit pupil; the
    pupil is built on the host (exit_pupil), the O(n^3) sums run on the GPU."""
    P, P0 = exit_pupil(lens, lib, focal_length_um, wavelength_um, diameter_um)
    return mtf_from_pupil(P, P0, lens.period_um, focal_length_um, wavelength_um, diameter_um)


def mtf_from_pupil(P, P0, period_um: float, focal_length_um: float, wavelength_um: float,
                   diameter_um: float) -> MtfCurve:
    """The MTF of a given exit pupil P (complex n x n) and aperture mask P0."""
    P = _c(P)
    P0 = _d(P0)
    n = P.shape[0]
    if P.shape != (n, n) or P0.shape != (n, n):
        raise ValueError("pupil and mask must be n x n")
    cap = n
    f32 = ctypes.POINTER(ctypes.c_float)
    freq = np.zeros(cap, np.float32)
    mtf = np.zeros(cap, np.float32)
    ideal = np.zeros(cap, np.float32)
    cutoff = np.zeros(1)
    npts = ctypes.c_int(0)
    _check(_lib.load().crs_analyze_mtf(_p(P.view(np.float64)), _p(P0), n, period_um,
                                       focal_length_um, wavelength_um, diameter_um, cap,
                                       freq.ctypes.data_as(f32), mtf.ctypes.data_as(f32),
                                       ideal.ctypes.data_as(f32), _p(cutoff), ctypes.byref(npts)),
           "crs_analyze_mtf")
    m = npts.value
    return MtfCurve(freq[:m].copy(), mtf[:m].copy(), ideal[:m].copy(), float(cutoff[0]))


def analyze_field_of_view(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                          wavelength_um: float, diameter_um: float, angles_deg,
                          mode="fp32c") -> List[FieldPoint]:
    """field.cpp:12-76: relative peak and spot position per field angle, one launch."""
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    px, py, t = _pillars(px, py, _c(t))
    ang = _d(np.atleast_1d(angles_deg))
    out = np.zeros(3 * max(ang.size, 1))
    _check(_lib.load().crs_analyze_field_of_view(_p(px), _p(py), _p(t), px.size, focal_length_um,
                                                 wavelength_um, diameter_um, _p(ang), ang.size,
                                                 _mode(mode), _p(out)),
           "crs_analyze_field_of_view")
    return [FieldPoint(*map(float, out[3 * i:3 * i + 3])) for i in range(ang.size)]


def analyze_wide_fov(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                     wavelength_um: float, lens_diameter_um: float, stop_diameter_um: float,
                     stop_distance_um: float, angles_deg, mode="fp32c") -> List[FieldPoint]:
    """field.cpp:78-148: field angles seen through a decentred stop patch, one launch."""
    px, py, t = aperture_pillars(lens, lib, lens_diameter_um)
    px, py, t = _pillars(px, py, _c(t))
    ang = _d(np.atleast_1d(angles_deg))
    out = np.zeros(3 * max(ang.size, 1))
    _check(_lib.load().crs_analyze_wide_fov(_p(px), _p(py), _p(t), px.size, focal_length_um,
                                            wavelength_um, lens_diameter_um, stop_diameter_um,
                                            stop_distance_um, _p(ang), ang.size, _mode(mode),
                                            _p(out)),
           "crs_analyze_wide_fov")
    return [FieldPoint(*map(float, out[3 * i:3 * i + 3])) for i in range(ang.size)]


# ---- fabrication tolerance ----------------------------------------------------------------

@dataclass
class ToleranceResult:
    """tolerance.hpp: Strehl statistics over the Monte-Carlo trials of one CD error."""

    sigma_nm: float
    mean_strehl: float
    std_strehl: float
    worst_strehl: float


def analyze_tolerance(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
                      wavelength_um: float, diameter_um: float, sigmas_nm, trials: int,
                      seed: int, mode="fp32c", return_trials: bool = False):
    """tolerance.cpp:12-48: per sigma, max(1, trials) trials of jittered fills (the
    reference's std::mt19937 / std::normal_distribution draws), each trial's Strehl from
    analyze_focus's grid; all trials in batched launches.  With return_trials, also the
    per-trial Strehl ratios [len(sigmas), max(1, trials)]."""
    fm = _d(np.asarray(lens.fill_map, np.float64).ravel())
    if fm.size != lens.n_cells * lens.n_cells:
        raise ValueError("fill_map must hold n_cells^2 fills")
    lf, lp, la = _d(lib.fill), _d(lib.phase), _d(lib.amplitude)
    if not (lf.ndim == lp.ndim == la.ndim == 1 and lf.size == lp.size == la.size > 0):
        raise ValueError("library fill, phase and amplitude must be 1-D arrays of one length")
    sg = _d(np.atleast_1d(sigmas_nm))
    n = max(1, int(trials))
    out = np.zeros(4 * sg.size)
    per = np.zeros(sg.size * n)
    _check(_lib.load().crs_analyze_tolerance(_p(fm), int(lens.n_cells), float(lens.period_um),
                                             _p(lf), _p(lp), _p(la), lf.size, float(focal_length_um),
                                             float(wavelength_um), float(diameter_um), _p(sg),
                                             sg.size, int(trials), int(seed) & 0xFFFFFFFF,
                                             _mode(mode), _p(out), _p(per)),
           "crs_analyze_tolerance")
    res = [ToleranceResult(*map(float, out[4 * i:4 * i + 4])) for i in range(sg.size)]
    return (res, per.reshape(sg.size, n)) if return_trials else res


# ---- on-axis focus scans ----------------------------------------------------------------

@dataclass
class AchromaticFocalPoint:
    """achromatic.hpp:150-154: best on-axis focus per band, peak relative to the design
    point (the centre wavelength at the design focus)."""

    wavelength_um: float
    focal_length_um: float
    rel_peak: float


@dataclass
class PolarFocus:
    """The polarisation-multiplexed lens check (cmd_polar.cpp:85-126): each channel's
    on-axis focus and its focal isolation (dB) over the other channel there."""

    z_x_um: float
    z_y_um: float
    isolation_x_db: float
    isolation_y_db: float


def axis_scan(px, py, t_sets, scans, mode="fp32c"):
    """Batched on-axis scans (crs_axis_scan).  t_sets: [nsets, npil] complex;
    scans: (tset, nz, wavelength_um, zlo, zspan, z_default) each, samples
    z_j = zlo + zspan j / (nz - 1).  Returns (best_z, best_I, curves) with the callers'
    first-maximum rule."""
    px, py = _pillars(px, py)
    t = _c(np.atleast_2d(t_sets))
    if t.ndim != 2 or t.shape[1] != px.size:
        raise ValueError("t_sets must be [nsets, npil]")
    descs = (_lib.crs_axis_scan_desc * len(scans))()
    total = 0
    for i, (tset, nz, lam, zlo, zspan, zdef) in enumerate(scans):
        descs[i] = _lib.crs_axis_scan_desc(int(tset), int(nz), float(lam), float(zlo),
                                           float(zspan), float(zdef))
        total += int(nz)
    bz, bi, curves = np.zeros(len(scans)), np.zeros(len(scans)), np.zeros(total)
    _check(_lib.load().crs_axis_scan(_p(px), _p(py), _p(t), px.size, t.shape[0], descs,
                                     len(scans), _mode(mode), _p(bz), _p(bi), _p(curves)),
           "crs_axis_scan")
    return bz, bi, curves


def verify_achromatic_focus_bands(px, py, t_bands, wavelengths_um, center_index: int,
                                  center_wavelength_um: float, focal_length_um: float,
                                  mode="fp32c") -> List[AchromaticFocalPoint]:
    """The focus check shared by verify_achromatic_focus / verify_pb_achromatic_focus on
    the aperture cells' per-band transmissions t_bands [nb, npil]."""
    px, py = _pillars(px, py)
    t = _c(np.atleast_2d(t_bands))
    lam = _d(wavelengths_um)
    nb = lam.size
    if t.shape != (nb, px.size):
        raise ValueError("t_bands must be [len(wavelengths), npil]")
    out = np.zeros(3 * nb)
    _check(_lib.load().crs_verify_achromatic_focus(_p(px), _p(py), _p(t), px.size, nb, _p(lam),
                                                   int(center_index), float(center_wavelength_um),
                                                   float(focal_length_um), _mode(mode), _p(out)),
           "crs_verify_achromatic_focus")
    return [AchromaticFocalPoint(*map(float, out[3 * j:3 * j + 3])) for j in range(nb)]


def verify_achromatic_focus(lib: DispersiveLibrary, design: AchromaticDesign,
                            mode="fp32c") -> List[AchromaticFocalPoint]:
    """achromatic.cpp:288-338: per band, the on-axis focus over 200 samples in
    [0.4 f, 1.8 f] and its peak relative to the design point."""
    px, py, t = achromatic_band_transmissions(lib, design)
    return verify_achromatic_focus_bands(px, py, t, lib.wavelengths_um, lib.center_index,
                                         lib.center_wavelength_um, design.focal_length_um, mode)


def verify_pb_achromatic_focus(lib: DispersivePbLibrary, design: PbAchromaticDesign,
                               mode="fp32c") -> List[AchromaticFocalPoint]:
    """pb_achromatic.cpp:255-301: as verify_achromatic_focus with the cross-circular
    transmission a_cross(atom) exp(i phi_geo) of each cell."""
    px, py, t = pb_achromatic_band_transmissions(lib, design)
    return verify_achromatic_focus_bands(px, py, t, lib.wavelengths_um, lib.center_index,
                                         lib.center_wavelength_um, design.focal_length_um, mode)


def polar_focus_check(px, py, tx, ty, wavelength_um: float, focal_x_um: float,
                      focal_y_um: float, nz: int = 240, mode="fp32c") -> PolarFocus:
    """cmd_polar.cpp:85-126 (nz = 240; gui/workers.cpp:181-197 uses nz = 200)."""
    px, py, tx, ty = _pillars(px, py, _c(tx), _c(ty))
    out = np.zeros(4)
    _check(_lib.load().crs_polar_focus_check(_p(px), _p(py), _p(tx), _p(ty), px.size,
                                             float(wavelength_um), float(focal_x_um),
                                             float(focal_y_um), int(nz), _mode(mode), _p(out)),
           "crs_polar_focus_check")
    return PolarFocus(*map(float, out))


def profile_focus_z(px, py, t, wavelength_um: float, focal_um: float, nz: int = 240,
                    mode="fp32c") -> float:
    """The focusing-profile proof of cli_common.cpp:112-133: the on-axis focus over
    z = 0.5 f + f j / (nz - 1), j < nz."""
    bz, _, _ = axis_scan(px, py, np.atleast_2d(_c(t)),
                         [(0, nz, wavelength_um, 0.5 * focal_um, focal_um, 0.5 * focal_um)], mode)
    return float(bz[0])
