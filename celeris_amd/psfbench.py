"""psfbench (cli/bench.cpp:96-170): the reference's PSF propagation benchmark, GPU legs.

Inputs are the reference's: D = 120 um, f = 50 um, lambda = 0.532 um, period 0.35 um, a
161 x 161 focal grid over half-window W = max(6 lambda f / D, 4) um at z = f, and the
circular-aperture pillar list of a ``design_metalens`` lens (343 x 343 cells, 92,361
pillars inside D; bench.cpp:112-137).  The reference fills its 16-sample unit-cell
library by an RCWA sweep of a TiO2-like pillar (bench.cpp:114-117), a solver outside
this package; :func:`psfbench_library` stands in for it with the same 16 fills over
[0.08, 0.92], a phase ramp that covers the circle and |t| between 0.86 and 0.95.  The lens
is then designed by the reference's rule (lookup_weighted, amplitude weight 0.25).

Timing: each arithmetic mode through the drop-in ``propagate_psf`` (the reference's own
call: host arrays in, |E|^2 out, staging and copies included), after a warm-up, median
of ``reps`` calls; plus the pair-sum kernel alone, timed with HIP events on its stream.
The reference's "CPU" figure timed ``compute_psf``, which in a GPU build runs the GPU
path first, before the warm-up (bench.cpp:139-149, focal.cpp:172-186); ``bench.py
--psfbench`` times a real CPU leg beside these.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from . import rs
from .design import MetalensDesign, UnitCellLibrary, aperture_pillars, design_metalens

DIAMETER_UM, FOCAL_UM, WAVELENGTH_UM, PERIOD_UM, GRID = 120.0, 50.0, 0.532, 0.35, 161
FILL_SAMPLES = 16


def psfbench_library(wavelength_um: float = WAVELENGTH_UM, period_um: float = PERIOD_UM,
                     fill_samples: int = FILL_SAMPLES) -> UnitCellLibrary:
    """Stand-in for the reference's RCWA library (bench.cpp:114-117): fills
    linspace(0.08, 0.92), phase rising with fill over 2 pi (wrapped into (-pi, pi]),
    |t| = 0.95 - 0.36 (fill - 0.5)^2."""
    fill = np.linspace(0.08, 0.92, fill_samples)
    ramp = 2.0 * math.pi * (fill - fill[0]) / (fill[-1] - fill[0]) * (fill_samples - 1) / fill_samples
    phase = np.angle(np.exp(1j * (ramp - math.pi + 0.1)))
    amp = 0.95 - 0.36 * (fill - 0.5) ** 2
    return UnitCellLibrary(period_um, fill, phase, amp, wavelength_um=wavelength_um,
                           thickness_um=0.6)


@dataclass
class PsfbenchLens:
    lens: MetalensDesign
    lib: UnitCellLibrary
    px: np.ndarray
    py: np.ndarray
    t: np.ndarray
    focal_um: float
    wavelength_um: float
    diameter_um: float
    grid: int
    half_window_um: float


def psfbench_lens(diameter_um: float = DIAMETER_UM, focal_um: float = FOCAL_UM,
                  grid: int = GRID, wavelength_um: float = WAVELENGTH_UM,
                  period_um: float = PERIOD_UM) -> PsfbenchLens:
    """The lens and aperture pillar list psfbench builds (bench.cpp:118-137)."""
    lib = psfbench_library(wavelength_um, period_um)
    lens = design_metalens(lib, focal_um, diameter_um)
    px, py, t = aperture_pillars(lens, lib, diameter_um)
    dl = wavelength_um * focal_um / diameter_um
    return PsfbenchLens(lens, lib, px, py, t, focal_um, wavelength_um, diameter_um, grid,
                        max(6.0 * dl, 4.0))


def max_normalized_diff(a: np.ndarray, b: np.ndarray) -> float:
    """max |a / max a - b / max b| (bench.cpp:156-162), 0 when either peak is 0."""
    am, bm = float(np.max(a)), float(np.max(b))
    if am <= 0 or bm <= 0:
        return 0.0
    return float(np.max(np.abs(a / am - b / bm)))


def run_gpu(L: PsfbenchLens, modes=("fp32", "fp32c", "fp64"), reps: int = 5) -> Dict[str, dict]:
    """Per mode: the drop-in call's wall time (median of reps after a warm-up), the
    pair kernel's device time, and the |E|^2 map."""
    import torch

    k = 2.0 * math.pi / L.wavelength_um
    out: Dict[str, dict] = {}
    prev = rs.set_default_mode("fp32c")
    try:
        for mode in modes:
            rs.set_default_mode(mode)
            I = rs.propagate_psf(L.px, L.py, L.t, 0.0, 0.0, L.focal_um, k, L.grid,
                                 L.half_window_um)  # warm-up (and the map)
            if I is None:
                raise RuntimeError("propagate_psf returned false")
            wall: List[float] = []
            for _ in range(reps):
                t0 = time.perf_counter()
                rs.propagate_psf(L.px, L.py, L.t, 0.0, 0.0, L.focal_um, k, L.grid,
                                 L.half_window_um)
                wall.append((time.perf_counter() - t0) * 1e3)
            # The pair kernel alone, inputs resident, HIP events on the context's stream.
            dev = torch.device("cuda:0")
            ctx = rs.Context(0)
            ctx.set_timing(True)
            dpx = torch.from_numpy(L.px).to(dev)
            dpy = torch.from_numpy(L.py).to(dev)
            dt = torch.from_numpy(np.ascontiguousarray(L.t[None, :])).to(dev)
            plane = rs.Plane.psf(0.0, 0.0, L.focal_um, k, L.grid, L.half_window_um)
            ks = []
            for _ in range(reps + 1):
                ctx.forward(dpx, dpy, dt, [plane], mode=mode, field=False, intensity=True)
                ks.append(ctx.last_kernel_ms(False))
            ctx.close()
            out[mode] = {"call_ms": float(np.median(wall)), "kernel_ms": float(np.median(ks[1:])),
                         "intensity": I}
    finally:
        rs.set_default_mode(prev)
    return out
