"""Image output of focal-plane maps: the reference's PGM writer (io/image.hpp:11-16,
io/image.cpp:10-37), with the same normalisation, gamma and rounding, so a PSF or
caustic from this package dumps to the same bytes the reference's CLI writes
(cmd_design.cpp: ``--psf`` with gamma 2.2, ``--report`` caustic with gamma 1.6).
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def pgm_bytes(nx: int, ny: int, values, gamma: float = 1.0) -> Optional[bytes]:
    """The P5 (binary grayscale) image of a row-major nx x ny map, or None where the
    reference returns false for the arguments (non-positive sizes, wrong length).

    Each value is divided by the map's maximum (1 if that is <= 0), clamped to [0, 1],
    raised to 1/gamma when gamma != 1, and rounded to 0..255 half away from zero.
    """
    v = np.asarray(values, dtype=np.float64).ravel()
    if nx <= 0 or ny <= 0 or v.size != nx * ny:
        return None
    maxv = float(np.fmax.reduce(v, initial=0.0))  # NaN entries do not raise the maximum
    if maxv <= 0.0:
        maxv = 1.0
    norm = np.clip(v / maxv, 0.0, 1.0)
    if gamma != 1.0:
        norm = np.power(norm, 1.0 / gamma if gamma != 0.0 else 1.0)
    y = norm * 255.0
    whole = np.trunc(y)
    px = (whole + ((y - whole) >= 0.5)).astype(np.uint8)
    return b"P5\n%d %d\n255\n" % (nx, ny) + px.tobytes()


def write_pgm(path: str, nx: int, ny: int, values, gamma: float = 1.0) -> bool:
    """write_pgm (io/image.hpp:15-16): False on bad arguments or a file error."""
    data = pgm_bytes(nx, ny, values, gamma)
    if data is None:
        return False
    try:
        with open(path, "wb") as f:
            f.write(data)
    except OSError:
        return False
    return True
