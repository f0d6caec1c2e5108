"""Synthetic metalens inputs for the benchmark and test workloads.

Every random draw comes from a counter-based splitmix64 stream, so a value depends
only on (seed, stream id, index) and is identical in every process and on every
machine (no std::uniform_real_distribution or numpy Generator involved).

Geometry follows the reference's conventions:
  * cell centres ((ix - (N-1)/2) p, (iy - (N-1)/2) p), iy outer, ix inner
    (design/metalens.cpp cell loops; analysis/focal.cpp:45-57)
  * hyperbolic focusing phase phi = -k (sqrt(r^2 + f^2) - f)
    (design/phase_profile.cpp:41-44)
  * NA = sin(atan(R / f)) with R the aperture half-width (analysis/throughfocus.cpp:78)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .rs import Plane

SEED = 20261203
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)

# Stream ids, one per random field.
STREAM_AMP = 1
STREAM_PHASE_NOISE = 2
STREAM_DISPERSION = 3
STREAM_GRID = 4
STREAM_FD_PILLARS = 5
STREAM_AMP_Y = 6
STREAM_U = 7


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finalizer of x + golden gamma, on uint64 arrays (wrapping)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n draws in [0, 1) from stream `stream` of `seed`, indices start..start+n-1."""
    key = splitmix64(np.array([np.uint64(seed) ^ (np.uint64(stream) << np.uint64(40))],
                              dtype=np.uint64))[0]
    idx = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        bits = splitmix64((idx * np.uint64(0xD1B54A32D192ED03)) ^ key)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def square_grid(n_side: int, pitch: float):
    c = (n_side - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    px = ((ix - c) * pitch).ravel().astype(np.float64)
    py = ((iy - c) * pitch).ravel().astype(np.float64)
    return px, py


def hyperbolic_phase(px, py, wavelength: float, focal: float) -> np.ndarray:
    k = 2.0 * math.pi / wavelength
    r2 = px * px + py * py
    return -k * (np.sqrt(r2 + focal * focal) - focal)


def numerical_aperture(half_width: float, focal: float) -> float:
    return math.sin(math.atan(half_width / focal))


@dataclass
class Lens:
    """A pillar list and its transmission sets (t: complex [nsets, npil])."""

    name: str
    px: np.ndarray
    py: np.ndarray
    t: np.ndarray
    planes: List[Plane]
    wavelengths: List[float]
    focal: float
    na: float
    meta: dict = field(default_factory=dict)

    @property
    def npil(self) -> int:
        return int(self.px.size)

    @property
    def pairs(self) -> int:
        return self.npil * sum(p.npix for p in self.planes)


def config1(seed: int = SEED) -> Lens:
    """Plain FP64 CPU case: lambda 0.532, 64x64 pillars at 0.4 um (4,096), f = 50,
    |t| ~ U[0.8, 1.0], 81x81 grid over +-5 um at z = f."""
    lam, p, n_side, f = 0.532, 0.4, 64, 50.0
    px, py = square_grid(n_side, p)
    a = 0.8 + 0.2 * uniform(seed, STREAM_AMP, px.size)
    t = (a * np.exp(1j * hyperbolic_phase(px, py, lam, f)))[None, :]
    k = 2.0 * math.pi / lam
    planes = [Plane.psf(0.0, 0.0, f, k, 81, 5.0)]
    na = numerical_aperture(n_side * p / 2.0, f)
    return Lens("config1", px, py, t, planes, [lam], f, na)


def config2(n: int = 161, seed: int = SEED) -> Lens:
    """Headline shape: lambda 0.633, 303x303 pillars (91,809) at 0.35 um, f = 100,
    hyperbolic phase + U[-0.05, 0.05] rad noise, |t| ~ U[0.7, 1.0]; n x n grid over
    +-4 lambda/NA at z = f."""
    lam, p, n_side, f = 0.633, 0.35, 303, 100.0
    px, py = square_grid(n_side, p)
    a = 0.7 + 0.3 * uniform(seed, STREAM_AMP, px.size)
    noise = (uniform(seed, STREAM_PHASE_NOISE, px.size) - 0.5) * 0.1
    t = (a * np.exp(1j * (hyperbolic_phase(px, py, lam, f) + noise)))[None, :]
    na = numerical_aperture(n_side * p / 2.0, f)
    W = 4.0 * lam / na
    k = 2.0 * math.pi / lam
    planes = [Plane.psf(0.0, 0.0, f, k, n, W)]
    return Lens(f"config2-{n}", px, py, t, planes, [lam], f, na,
                {"amp": a, "phase": hyperbolic_phase(px, py, lam, f) + noise, "W": W})


def config3(seed: int = SEED, n_lambda: int = 8, n_z: int = 31, sigma: float = 0.05,
            grids=(81, 161, 321), n_side: int = 303) -> Lens:
    """Batched through-focus x chromatic: config-2 pillars, per-wavelength t with
    phase scaled by 0.55/lambda plus U[-sigma, sigma] dispersion noise, 8 lambdas in
    [0.45, 0.70] x 31 planes z in f +- 20, grid per instance drawn from `grids`."""
    p, f, lam_d = 0.35, 100.0, 0.55
    px, py = square_grid(n_side, p)
    npil = px.size
    a = 0.7 + 0.3 * uniform(seed, STREAM_AMP, npil)
    phi_d = hyperbolic_phase(px, py, lam_d, f)
    lams = np.linspace(0.45, 0.70, n_lambda)
    t = np.empty((n_lambda, npil), dtype=np.complex128)
    for s, lam in enumerate(lams):
        noise = (uniform(seed, STREAM_DISPERSION, npil, start=s * npil) - 0.5) * 2.0 * sigma
        t[s] = a * np.exp(1j * (phi_d * (lam_d / lam) + noise))
    na = numerical_aperture(n_side * p / 2.0, f)
    zs = np.linspace(f - 20.0, f + 20.0, n_z)
    draw = uniform(seed, STREAM_GRID, n_lambda * n_z)
    planes = []
    for s, lam in enumerate(lams):
        k = 2.0 * math.pi / lam
        W = 4.0 * lam / na
        for iz, z in enumerate(zs):
            n = grids[int(draw[s * n_z + iz] * len(grids))]
            planes.append(Plane.psf(0.0, 0.0, float(z), k, n, W, tset=s))
    return Lens("config3", px, py, t, planes, list(lams), f, na, {"sigma": sigma, "zs": zs})


def config5(seed: int = SEED, n: int = 1024, n_side: int = 1001, lam: float = 0.532) -> Lens:
    """Polarization-multiplexed lens: 1001x1001 pillars (1,002,001) at 0.3 um, X
    channel focusing at f1 = 250, Y channel at f2 = 300, |t| ~ U[0.8, 1.0] per
    channel; an n x n grid per channel at its own focus (t set 0 = X, 1 = Y)."""
    p, f1, f2 = 0.3, 250.0, 300.0
    px, py = square_grid(n_side, p)
    npil = px.size
    ax = 0.8 + 0.2 * uniform(seed, STREAM_AMP, npil)
    ay = 0.8 + 0.2 * uniform(seed, STREAM_AMP_Y, npil)
    t = np.stack([ax * np.exp(1j * hyperbolic_phase(px, py, lam, f1)),
                  ay * np.exp(1j * hyperbolic_phase(px, py, lam, f2))])
    k = 2.0 * math.pi / lam
    half = n_side * p / 2.0
    planes = []
    for s, f in enumerate((f1, f2)):
        W = 4.0 * lam / numerical_aperture(half, f)
        planes.append(Plane.psf(0.0, 0.0, f, k, n, W, tset=s))
    return Lens("config5", px, py, t, planes, [lam], f1, numerical_aperture(half, f1),
                {"f2": f2})


def small_lens(n_side: int = 12, pitch: float = 0.4, lam: float = 0.532, focal: float = 20.0,
               seed: int = SEED, n: int = 9, half_window: Optional[float] = None) -> Lens:
    """A tiny config-1-like lens for tests that must finish in seconds on a CPU."""
    px, py = square_grid(n_side, pitch)
    a = 0.8 + 0.2 * uniform(seed, STREAM_AMP, px.size)
    t = (a * np.exp(1j * hyperbolic_phase(px, py, lam, focal)))[None, :]
    na = numerical_aperture(n_side * pitch / 2.0, focal)
    W = half_window if half_window is not None else 4.0 * lam / na
    k = 2.0 * math.pi / lam
    return Lens("small", px, py, t, [Plane.psf(0.0, 0.0, focal, k, n, W)], [lam], focal, na)
