"""The lens-description helpers (celeris_amd.design) against the reference's own code.

design_metalens (metalens.cpp:230-273) from the reference build writes
tests/golden/design_metalens.npz: psfbench's lens (D 120, f 50, lambda 0.532, p 0.35, the
stand-in library of celeris_amd.psfbench) and a seeded library at amplitude weights 0
and 1.  The restatement must pick the same library entry in every cell (bit-exact fill
map) and reproduce the RMS phase error and mean |t| to roundoff.
"""
import os

import numpy as np
import pytest

from celeris_amd.design import UnitCellLibrary, angle_diff, aperture_pillars, design_metalens

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "design_metalens.npz")


def _cases():
    d = np.load(GOLDEN)
    return [(n, d) for n in sorted({k.split("__")[0] for k in d.files})]


@pytest.mark.parametrize("name,d", _cases(), ids=lambda c: c if isinstance(c, str) else "")
def test_design_metalens_matches_reference(name, d):
    fill, phase, amp = d[f"{name}__lib"]
    lam, per, f, D, w = d[f"{name}__args"]
    lib = UnitCellLibrary(per, fill, phase, amp, wavelength_um=lam)
    L = design_metalens(lib, f, D, w)
    n_ref, rms_ref, amp_ref = d[f"{name}__n_rms_amp"]
    assert L.n_cells == int(n_ref)
    assert np.array_equal(L.fill_map, fill[d[f"{name}__fill_index"]])
    assert L.rms_phase_error_deg == pytest.approx(rms_ref, rel=1e-12)
    assert L.mean_amplitude == pytest.approx(amp_ref, rel=1e-12)
    if name == "psfbench":
        px, _, _ = aperture_pillars(L, lib, D)
        assert px.size == 92361  # the pillar count psfbench reports for D = 120, p = 0.35


def test_angle_diff_wraps_by_whole_turns():
    """(-pi, pi] with pi itself kept and -pi sent to +pi; many turns away still lands in
    range; the result differs from a - b by whole turns."""
    a = np.array([np.pi, -np.pi, 0.0, 7.5, -40.0, 3.0])
    b = np.array([0.0, 0.0, 0.0, 0.0, 0.0, -3.5])
    d = angle_diff(a, b)
    assert d[0] == np.pi and d[1] == np.pi and d[2] == 0.0
    assert np.all((d > -np.pi) & (d <= np.pi))
    turns = (a - b - d) / (2 * np.pi)
    assert np.allclose(turns, np.round(turns), atol=1e-12)


def test_psfbench_lens_inputs():
    """psfbench's inputs (bench.cpp:96-137): 343 x 343 cells, the aperture's 92,361
    pillars, W = max(6 lambda f / D, 4) = 4 um, a 161 grid."""
    from celeris_amd.psfbench import psfbench_lens

    L = psfbench_lens()
    assert L.lens.n_cells == 343 and L.px.size == 92361
    assert L.half_window_um == 4.0 and L.grid == 161
    assert np.all(np.abs(L.t) > 0.8)


def test_angle_diff_refuses_infinite_angles():
    """Wrapping one turn at a time (metalens.cpp:15-20) cannot finish from +-inf; NaN
    passes through as NaN."""
    with pytest.raises(ValueError, match="infinite"):
        angle_diff(np.array([0.0, np.inf]), 0.0)
    assert np.isnan(angle_diff(np.nan, 0.0))
