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
