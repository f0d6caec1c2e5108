"""Lens-description types of the reference and the aperture pillar list built from them.

Mirrors ``UnitCellLibrary`` / ``MetalensDesign`` (src/celeris/design/metalens.hpp:21-47,
113-119) as plain data, and the circular-aperture cell list every reference analysis
builds before propagating (analysis/focal.cpp:42-57, :156-164; throughfocus.cpp:58-66;
chromatic.cpp:20-28).  The RCWA sweep that fills a library stays outside this package:
a library here is given, not computed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class UnitCellLibrary:
    """Sampled pillar fills and the complex transmission each produces."""

    period_um: float
    fill: np.ndarray
    phase: np.ndarray
    amplitude: np.ndarray
    wavelength_um: float = 0.0
    thickness_um: float = 0.0

    def lookup_fill(self, fills: np.ndarray) -> np.ndarray:
        """Index of the nearest sampled fill, first one on ties (metalens.cpp:112-119)."""
        lib = np.asarray(self.fill, np.float64)
        f = np.asarray(fills, np.float64)
        d = np.abs(lib[None, :] - f.reshape(-1, 1))
        return np.argmin(d, axis=1).reshape(f.shape)

    def lookup_weighted(self, targets, amplitude_weight: float) -> np.ndarray:
        """Index of the entry minimising angle_diff(phase, target)^2 +
        w (1 - |t|)^2, the first one on ties (metalens.cpp:58-69)."""
        tg = np.asarray(targets, np.float64)
        ph = np.asarray(self.phase, np.float64)
        amp = np.asarray(self.amplitude, np.float64)
        d = angle_diff(ph[None, :], tg.reshape(-1, 1))
        damp = 1.0 - amp
        cost = d * d + amplitude_weight * damp * damp
        return np.argmin(cost, axis=1).reshape(tg.shape)

    def transmission_for_fill(self, fills) -> np.ndarray:
        """t = |t| e^{i phase} of the nearest sampled pillar (metalens.cpp:112-120)."""
        idx = self.lookup_fill(fills)
        amp = np.asarray(self.amplitude, np.float64)[idx]
        ph = np.asarray(self.phase, np.float64)[idx]
        return amp * np.cos(ph) + 1j * (amp * np.sin(ph))


@dataclass
class MetalensDesign:
    """A square lattice of n_cells x n_cells pillar fills (row-major, iy outer)."""

    n_cells: int
    period_um: float
    fill_map: np.ndarray
    rms_phase_error_deg: float = 0.0
    mean_amplitude: float = 0.0


def angle_diff(a, b) -> np.ndarray:
    """a - b wrapped into (-pi, pi] by whole turns, one turn at a time
    (metalens.cpp:15-20), element-wise."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    d = np.array(d, dtype=np.float64, copy=True)
    if np.isinf(d).any():  # one turn at a time never reaches (-pi, pi] from +-inf
        raise ValueError("angle_diff: infinite angle")
    two_pi = 2.0 * math.pi
    while True:
        m = d > math.pi
        if not m.any():
            break
        d[m] -= two_pi
    while True:
        m = d <= -math.pi
        if not m.any():
            break
        d[m] += two_pi
    return d


def hyperbolic_target(x, y, wavelength_um: float, focal_length_um: float) -> np.ndarray:
    """The focusing profile's phase -k (sqrt(r^2 + f^2) - f) (phase_profile.cpp:36-45)."""
    k = 2.0 * math.pi / wavelength_um
    r = np.sqrt(x * x + y * y)
    return -k * (np.sqrt(r * r + focal_length_um * focal_length_um) - focal_length_um)


def design_metalens(lib: "UnitCellLibrary", focal_length_um: float, diameter_um: float,
                    amplitude_weight: float = 0.25) -> MetalensDesign:
    """design_metalens for a focusing profile (metalens.cpp:230-273): n = max(1,
    round(D / p)) cells a side, each taking the library entry that best trades phase
    error against transmission (lookup_weighted) for the hyperbolic target at its centre
    (every cell of the square, not only the aperture); RMS phase error (degrees) and
    mean |t| over all cells."""
    p = lib.period_um
    n = max(1, int(math.floor(diameter_um / p + 0.5)))  # std::round, positive argument
    c = (n - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = ((ix - c) * p).ravel()
    y = ((iy - c) * p).ravel()
    target = hyperbolic_target(x, y, lib.wavelength_um, focal_length_um)
    idx = lib.lookup_weighted(target, amplitude_weight)
    fill_map = np.asarray(lib.fill, np.float64)[idx].reshape(n, n)
    err = angle_diff(np.asarray(lib.phase, np.float64)[idx], target)
    sq = np.cumsum(err * err)[-1]  # accumulated in cell order
    amp = np.cumsum(np.asarray(lib.amplitude, np.float64)[idx])[-1]
    return MetalensDesign(n, p, fill_map, math.sqrt(sq / idx.size) * 180.0 / math.pi,
                          amp / idx.size)


def aperture_pillars(lens: MetalensDesign, lib: UnitCellLibrary, diameter_um: float,
                     focal_length_um: float | None = None, wavelength_um: float | None = None):
    """Pillars inside the circular aperture, in the reference's cell order.

    Cell (ix, iy) sits at ((ix - c) p, (iy - c) p), c = (n_cells - 1)/2, and is kept
    unless sqrt(x^2 + y^2) > D/2 (focal.cpp:45-50).  Returns px, py, t and, when a
    focal length and wavelength are given, the ideal hyperbolic phase
    -k (sqrt(r^2 + f^2) - f) of each kept cell (focal.cpp:52-54).
    """
    n = lens.n_cells
    p = lens.period_um
    c = (n - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = ((ix - c) * p).ravel()
    y = ((iy - c) * p).ravel()
    r = np.sqrt(x * x + y * y)
    keep = ~(r > diameter_um / 2.0)
    fills = np.asarray(lens.fill_map, np.float64).ravel()[keep]
    t = lib.transmission_for_fill(fills)
    out = [x[keep].copy(), y[keep].copy(), t]
    if focal_length_um is not None and wavelength_um is not None:
        k = 2.0 * math.pi / wavelength_um
        rk = r[keep]
        out.append(-k * (np.sqrt(rk * rk + focal_length_um * focal_length_um) - focal_length_um))
    return tuple(out)


def aperture_fills(lens: MetalensDesign, diameter_um: float) -> np.ndarray:
    """Fill factors of the cells aperture_pillars keeps, in the same order (the per-cell
    input of a library lookup, e.g. chromatic.cpp:75-81 keeps fills and looks them up in
    each wavelength's library)."""
    n = lens.n_cells
    p = lens.period_um
    c = (n - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = ((ix - c) * p).ravel()
    y = ((iy - c) * p).ravel()
    keep = ~(np.sqrt(x * x + y * y) > diameter_um / 2.0)
    return np.asarray(lens.fill_map, np.float64).ravel()[keep]


def exit_pupil(lens: MetalensDesign, lib: UnitCellLibrary, focal_length_um: float,
               wavelength_um: float, diameter_um: float):
    """The exit pupil analyze_mtf correlates (mtf.cpp:20-36): on the full n_cells^2 grid,
    P = t(fill) e^{-i phi_ideal} inside the aperture (r <= D/2) and 0 outside, and the
    aperture mask P0.  Returns (P complex128 [n, n], P0 float64 [n, n])."""
    n = lens.n_cells
    p = lens.period_um
    c = (n - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = (ix - c) * p
    y = (iy - c) * p
    r = np.sqrt(x * x + y * y)
    inside = ~(r > diameter_um / 2.0)
    k = 2.0 * math.pi / wavelength_um
    phi = -k * (np.sqrt(r * r + focal_length_um * focal_length_um) - focal_length_um)
    t = lib.transmission_for_fill(np.asarray(lens.fill_map, np.float64).reshape(n, n))
    rot = np.cos(-phi) + 1j * np.sin(-phi)
    P = np.where(inside, t * rot, 0.0 + 0.0j)
    return P.astype(np.complex128), inside.astype(np.float64)


def aperture_cells(n_cells: int, period_um: float, diameter_um: float):
    """Cells of an n_cells^2 lattice inside the circular aperture, in row-major order:
    (x, y, site index) with x = (ix - c) p, y = (iy - c) p, c = (n_cells - 1)/2, kept
    unless sqrt(x^2 + y^2) > D/2 (achromatic.cpp:294-303, pb_achromatic.cpp:263-272)."""
    c = (n_cells - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n_cells), np.arange(n_cells), indexing="ij")
    x = ((ix - c) * period_um).ravel()
    y = ((iy - c) * period_um).ravel()
    keep = ~(np.sqrt(x * x + y * y) > diameter_um / 2.0)
    return x[keep].copy(), y[keep].copy(), np.flatnonzero(keep)


# ---- dispersive (achromatic) designs: the fields their focus checks read -----------------

@dataclass
class DispersiveAtom:
    """One meta-atom across the band (achromatic.hpp:38-49): |t| and phase per band
    sample.  The RCWA sweep that measures them stays outside this package."""

    amplitude: np.ndarray
    phase_rad: np.ndarray


@dataclass
class DispersiveLibrary:
    """achromatic.hpp:65-73: band samples (ascending), the centre sample, the atoms."""

    wavelengths_um: np.ndarray
    center_wavelength_um: float
    center_index: int
    atoms: list
    period_um: float = 0.0


@dataclass
class AchromaticDesign:
    """achromatic.hpp:112-130: per-site atom choice on an n_cells^2 lattice (row-major)."""

    n_cells: int
    period_um: float
    atom_index: np.ndarray
    focal_length_um: float
    diameter_um: float
    center_wavelength_um: float = 0.0


@dataclass
class DispersivePbAtom:
    """pb_achromatic.hpp:43-51: the spin-flip (cross-circular) transmission a_cross per
    band sample, complex."""

    cross: np.ndarray


@dataclass
class DispersivePbLibrary:
    """pb_achromatic.hpp:53-62."""

    wavelengths_um: np.ndarray
    center_wavelength_um: float
    center_index: int
    atoms: list
    period_um: float = 0.0


@dataclass
class PbAchromaticDesign:
    """pb_achromatic.hpp:72-92: per-site atom and geometric-phase rotation theta; the
    realised transmission is a_cross(atom) exp(i phi_geo), phi_geo = -2 handedness theta."""

    n_cells: int
    period_um: float
    atom_index: np.ndarray
    rotation_rad: np.ndarray
    focal_length_um: float
    diameter_um: float
    handedness: int = 1
    center_wavelength_um: float = 0.0


def achromatic_band_transmissions(lib: DispersiveLibrary, design: AchromaticDesign):
    """Aperture cells and t[j, cell] = polar(amplitude[j], phase[j]) of the cell's atom
    (achromatic.cpp:294-309).  Returns px, py, t [nb, ncells]."""
    x, y, site = aperture_cells(design.n_cells, design.period_um, design.diameter_um)
    atom = np.asarray(design.atom_index).ravel()[site]
    amp = np.stack([np.asarray(a.amplitude, np.float64) for a in lib.atoms])[atom].T
    ph = np.stack([np.asarray(a.phase_rad, np.float64) for a in lib.atoms])[atom].T
    return x, y, amp * np.cos(ph) + 1j * (amp * np.sin(ph))


def pb_achromatic_band_transmissions(lib: DispersivePbLibrary, design: PbAchromaticDesign):
    """Aperture cells and t[j, cell] = a_cross[j](atom) polar(1, phi_geo),
    phi_geo = -2 handedness theta (pb_achromatic.cpp:263-282).  Returns px, py, t."""
    x, y, site = aperture_cells(design.n_cells, design.period_um, design.diameter_um)
    atom = np.asarray(design.atom_index).ravel()[site]
    cross = np.stack([np.asarray(a.cross, np.complex128) for a in lib.atoms])[atom].T
    phi_geo = -2.0 * design.handedness * np.asarray(design.rotation_rad, np.float64).ravel()[site]
    return x, y, cross * (np.cos(phi_geo) + 1j * np.sin(phi_geo))
