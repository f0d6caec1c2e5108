"""The analysis layer (celeris_amd.analysis -> crs_analyze_*) against golden vectors
written by the reference's own analysis code (tests/golden/lens_analyses.npz, from
oracle/_ref/ref_driver: focal.cpp, throughfocus.cpp, chromatic.cpp, field.cpp).

In FP64 mode every sum matches the reference's FP64 CPU loops to roundoff, so the
derived metrics match to 1e-9 relative; discrete outputs (FWHM in whole pixel steps,
argmax positions, z grids) must match exactly.  FP32C metrics match to 1e-5.
"""
import math
import os

import numpy as np
import pytest

from celeris_amd import analysis
from celeris_amd.design import MetalensDesign, UnitCellLibrary
from conftest import dispersive_golden

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lens_analyses.npz")


@pytest.fixture(scope="module")
def lens(torch_cuda):
    d = np.load(GOLDEN)
    L = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    return d, L, lib, float(d["f"]), float(d["lam"]), float(d["D"])


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-9), ("fp32c", 1e-5)])
def test_analyze_focus(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    print(f"[analysis] analyze_focus {mode}", flush=True)
    a = analysis.analyze_focus(L, lib, f, lam, D, mode=mode)
    ref = d["focus"]
    got = np.array([a.strehl, a.phase_strehl, a.fwhm_um, a.diffraction_limit_um, a.encircled_energy])
    print(mode, got, ref)
    assert got[3] == ref[3]  # lambda f / D
    W = max(6 * ref[3], 4.0)
    slack = 0 if mode == "fp64" else 2 * W / 120 * 1.0001  # one pixel step (see field grid)
    assert abs(got[2] - ref[2]) <= slack
    assert np.allclose(got[[0, 1, 4]], ref[[0, 1, 4]], rtol=tol, atol=0)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_compute_psf(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    n, W = int(d["psf_n"][0]), float(d["psf_n"][1])
    m = analysis.compute_psf(L, lib, f, lam, D, n, W, mode=mode)
    ref = d["psf"].reshape(n, n)
    assert np.linalg.norm(m.intensity - ref) / np.linalg.norm(ref) <= tol


@pytest.mark.parametrize("mode,tol", [("fp64", 2e-7), ("fp32c", 1e-5)])
def test_analyze_through_focus(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    tf_ref = d["throughfocus"]
    N = int(tf_ref[0])
    z_ref, I_ref = tf_ref[1:1 + N], tf_ref[1 + N:1 + 2 * N]
    zpk_ref, dof_ref = tf_ref[1 + 2 * N], tf_ref[2 + 2 * N]
    cnx = int(tf_ref[3 + 2 * N])
    caustic_ref = tf_ref[5 + 2 * N:]
    print(f"[analysis] through focus {mode}", flush=True)
    tf = analysis.analyze_through_focus(L, lib, f, lam, D, caustic_n=cnx, mode=mode)
    assert np.array_equal(tf.z_um.astype(np.float64), z_ref)
    assert np.max(np.abs(tf.intensity - I_ref)) <= tol
    assert tf.z_peak_um == pytest.approx(zpk_ref, abs=1e-6)
    assert tf.dof_um == pytest.approx(dof_ref, abs=1e-5)
    assert np.max(np.abs(tf.caustic - caustic_ref)) <= tol
    # the reference's GPU path samples the caustic on 201 x 201
    tf201 = analysis.analyze_through_focus(L, lib, f, lam, D, mode=mode)
    assert tf201.caustic.size == 201 * 201 and tf201.caustic.max() == pytest.approx(1.0)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-9), ("fp32c", 1e-5)])
def test_analyze_chromatic(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    lmin, lmax, n = d["chromatic_args"]
    pts = analysis.analyze_chromatic(L, lib, f, lam, D, lmin, lmax, int(n), mode=mode)
    ref = d["chromatic"].reshape(-1, 3)
    got = np.array([[p.wavelength_um, p.focal_length_um, p.rel_peak] for p in pts])
    print(mode, got, ref)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1], ref[:, 1], rtol=1e-12, atol=0)  # argmax on the same z grid
    assert np.allclose(got[:, 2], ref[:, 2], rtol=tol, atol=0)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-9), ("fp32c", 1e-5)])
def test_analyze_chromatic_dispersive(torch_cuda, mode, tol):
    """analyze_chromatic_dispersive (chromatic.cpp:60-127) against the reference's own output
    (tests/golden/chromatic_dispersive.npz): per-wavelength libraries from the caller's
    builder, asked for in the reference's order (design wavelength first), the whole scan in
    one launch."""
    lens_, build, f, lam, D, lmin, lmax, n, ref, calls = dispersive_golden()
    pts = analysis.analyze_chromatic_dispersive(lens_, build, f, lam, D, lmin, lmax, n, mode=mode)
    got = np.array([[p.wavelength_um, p.focal_length_um, p.rel_peak] for p in pts])
    print(mode, got, ref)
    assert calls[0] == lam and np.allclose(calls[1:], [lmin + (lmax - lmin) * i / (n - 1) for i in range(n)])
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1], ref[:, 1], rtol=1e-12, atol=0)
    assert np.allclose(got[:, 2], ref[:, 2], rtol=tol, atol=0)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-9), ("fp32c", 1e-5)])
def test_analyze_field_grid(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    max_angle, n_half, psf_n = d["fieldgrid_args"]
    g = analysis.analyze_field_grid(L, lib, f, lam, D, max_angle, int(n_half), int(psf_n), mode=mode)
    ref = d["fieldgrid"]
    assert g.n == int(ref[0])
    ref = ref[1:].reshape(-1, 7)
    got = np.array([[p.theta_x_deg, p.theta_y_deg, p.rel_strehl, p.fwhm_x_um, p.fwhm_y_um,
                     p.cx_um, p.cy_um] for p in g.points])
    print(mode, got[:, 2], ref[:, 2])
    assert np.array_equal(got[:, [0, 1, 5, 6]], ref[:, [0, 1, 5, 6]])
    # FWHM counts whole pixel steps above half maximum; in FP32C a pixel lying within
    # roundoff of the threshold may fall on the other side (one step, 2W/(psf_n-1)).
    W = max(6 * lam * f / D, 4.0)
    step = 2 * W / (psf_n - 1)
    slack = 0 if mode == "fp64" else step * 1.0001
    assert np.max(np.abs(got[:, [3, 4]] - ref[:, [3, 4]])) <= slack
    assert np.allclose(got[:, 2], ref[:, 2], rtol=tol, atol=0)
    # selftest [21] invariants (selftest.cpp:908-929): centre node is the reference,
    # 4-corner symmetry of a rotationally symmetric lens
    # (FP32C: the reference peak and the centre node come from calls of different shapes,
    # whose pillar ranges are cut differently, so they agree to the mode's tolerance.)
    c = g.n // 2
    assert got[c * g.n + c, 2] == pytest.approx(1.0, abs=1e-9 if mode == "fp64" else tol)
    corners = got[[0, g.n - 1, (g.n - 1) * g.n, g.n * g.n - 1], 2]
    assert corners.max() - corners.min() < 1e-3


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_compute_psf_field(lens, mode, tol):
    d, L, lib, f, lam, D = lens
    angle, n, W, on_axis_peak = d["psffield_args"]
    p = analysis.compute_psf_field(L, lib, f, lam, D, angle, int(n), W, on_axis_peak, mode=mode)
    ref = d["psffield"]
    assert p.angle_deg == ref[0] and p.cx_um == pytest.approx(ref[1], rel=1e-15)
    assert p.rel_strehl == pytest.approx(ref[2], rel=tol)
    I_ref = ref[3:].reshape(int(n), int(n))
    assert np.linalg.norm(p.psf.intensity - I_ref) / np.linalg.norm(I_ref) <= tol


@pytest.mark.parametrize("name", ["lens_analyses.npz", "mtf_large.npz"])
def test_analyze_mtf_golden(torch_cuda, name):
    """analyze_mtf on the GPU (FP64 sums) against the reference's own float curve."""
    d = np.load(os.path.join(os.path.dirname(GOLDEN), name))
    L = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    ref = d["mtf"]
    k = int(ref[0])
    print(f"[analysis] analyze_mtf {name}", flush=True)
    c = analysis.analyze_mtf(L, lib, float(d["f"]), float(d["lam"]), float(d["D"]))
    assert c.freq_cyc_mm.size == k
    assert np.array_equal(c.freq_cyc_mm, ref[1:1 + k].astype(np.float32))
    # FP64 sums in another order than the reference's loop: equal up to the float cast
    # (at most one float ulp where a value sits on a rounding boundary).
    for got, want in ((c.mtf, ref[1 + k:1 + 2 * k]), (c.mtf_ideal, ref[1 + 2 * k:1 + 3 * k])):
        assert np.all(np.abs(got.astype(np.float64) - want) <= 1.2e-7 * np.abs(want) + 1e-30)
    assert c.cutoff_cyc_mm == ref[1 + 3 * k]


def test_analyze_mtf_vs_oracle_large(torch_cuda):
    """A 201 x 201-cell pupil with random phases: GPU curve vs the oracle's."""
    import oracle

    n, p, f, lam, D = 201, 0.3, 40.0, 0.55, 58.0
    rng = np.random.default_rng(11)
    c = (n - 1) / 2
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    P0 = (~(np.hypot((ix - c) * p, (iy - c) * p) > D / 2)).astype(np.float64)
    P = P0 * rng.uniform(0.6, 1.0, (n, n)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n)))
    cv = analysis.mtf_from_pupil(P, P0, p, f, lam, D)
    fr, m, mi, cut = oracle.mtf(P, P0, p, f, lam, D)
    assert np.array_equal(cv.freq_cyc_mm, fr) and cv.cutoff_cyc_mm == cut
    assert np.all(np.abs(cv.mtf.astype(np.float64) - m) <= 1.2e-7 * np.abs(m) + 1e-30)
    assert np.all(np.abs(cv.mtf_ideal.astype(np.float64) - mi) <= 1.2e-7 * np.abs(mi) + 1e-30)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-9), ("fp32c", 1e-5)])
def test_analyze_field_of_view_and_wide_fov(torch_cuda, mode, tol):
    """analyze_field_of_view / analyze_wide_fov (field.cpp:12-148) against the
    reference's own outputs (tests/golden/field_of_view.npz)."""
    d = np.load(os.path.join(os.path.dirname(GOLDEN), "field_of_view.npz"))
    L = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    f, lam, D = float(d["f"]), float(d["lam"]), float(d["D"])
    print(f"[analysis] field of view {mode}", flush=True)
    pts = analysis.analyze_field_of_view(L, lib, f, lam, D, d["fov_angles"], mode=mode)
    ref = d["fov"].reshape(-1, 3)
    got = np.array([[p.angle_deg, p.rel_strehl, p.spot_shift_um] for p in pts])
    print(mode, got, ref)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1], ref[:, 1], rtol=tol, atol=0)
    # the spot position is a pixel of the 81 x 81 window: equal, or one pixel step away
    # where two pixels tie to within the mode's roundoff
    W = np.array([max(3 * lam * f / D, 2.0) + abs(f * math.tan(math.radians(a)) - f * math.sin(math.radians(a)))
                  for a in ref[:, 0]])
    slack = 1e-9 if mode == "fp64" else 2 * W / 80 * 1.0001
    assert np.all(np.abs(got[:, 2] - ref[:, 2]) <= slack)
    Dstop, sdist = d["widefov_args"]
    pts = analysis.analyze_wide_fov(L, lib, f, lam, D, Dstop, sdist, d["widefov_angles"], mode=mode)
    ref = d["widefov"].reshape(-1, 3)
    got = np.array([[p.angle_deg, p.rel_strehl, p.spot_shift_um] for p in pts])
    print(mode, got, ref)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, 1], ref[:, 1], rtol=tol, atol=0)
    dls = lam * f / Dstop
    W = np.array([max(3 * dls, 2.0) + 0.5 * abs(f * math.tan(math.radians(a)) - f * math.sin(math.radians(a)))
                  for a in ref[:, 0]])
    slack = 1e-9 if mode == "fp64" else 2 * W / 80 * 1.0001
    assert np.all(np.abs(got[:, 2] - ref[:, 2]) <= slack)


TOL_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tolerance.npz")


@pytest.mark.parametrize("mode,rtol,stdtol", [("fp64", 1e-9, 3e-8), ("fp32c", 1e-5, 2e-6)])
@pytest.mark.parametrize("name", ["cli", "single"])
def test_analyze_tolerance_vs_reference(torch_cuda, name, mode, rtol, stdtol):
    """analyze_tolerance (tolerance.cpp:12-48) against the reference's own sweep: the
    same jitter draws (std::mt19937 / std::normal_distribution seeded per trial), every
    trial's Strehl from the batched grids.  The spread is the reference's
    sqrt(max(0, sum2/n - mean^2)), which cancels to a few ulp of mean^2 (~0.5): a zero
    spread (sigma 0, or one trial) reads anywhere up to sqrt(4 * 2^-53 * 0.5) ~ 1.5e-8,
    hence 3e-8 absolute in FP64."""
    import oracle

    d = np.load(TOL_GOLDEN)
    L = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    trials, seed = map(int, d[f"{name}__args"])
    print(f"[analysis] tolerance {name} {mode}", flush=True)
    res, per = analysis.analyze_tolerance(L, lib, float(d["f"]), float(d["lam"]), float(d["D"]),
                                          d[f"{name}__sigmas"], trials, seed, mode=mode,
                                          return_trials=True)
    got = np.array([[r.sigma_nm, r.mean_strehl, r.std_strehl, r.worst_strehl] for r in res])
    ref = d[f"{name}__out"]
    print(mode, name, got, ref)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, [1, 3]], ref[:, [1, 3]], rtol=rtol, atol=0)
    assert np.allclose(got[:, 2], ref[:, 2], rtol=0, atol=stdtol)
    _, oper = oracle.tolerance(int(d["n_cells"]), float(d["period"]), d["fill_map"], d["lib_fill"],
                               d["lib_phase"], d["lib_amp"], float(d["f"]), float(d["lam"]),
                               float(d["D"]), d[f"{name}__sigmas"], trials, seed,
                               return_trials=True)
    assert np.allclose(per, oper, rtol=rtol, atol=0)


def test_analyze_tolerance_many_trials_batches(torch_cuda):
    """300 trials of one sigma run in two launches (256 + 44 trials); trials on both
    sides of the boundary match the oracle's, and the statistics are the per-trial
    ratios' own."""
    import oracle

    d = np.load(TOL_GOLDEN)
    L = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    f, lam, D = float(d["f"]), float(d["lam"]), float(d["D"])
    res, per = analysis.analyze_tolerance(L, lib, f, lam, D, [15.0], 300, 777, mode="fp64",
                                          return_trials=True)
    assert per.shape == (1, 300)
    s = per[0]
    assert res[0].mean_strehl == pytest.approx(s.mean(), rel=1e-12)
    assert res[0].worst_strehl == s.min()
    assert res[0].std_strehl == pytest.approx(s.std(), rel=1e-6)
    # oracle Strehl of single trials
    n = int(d["n_cells"])
    c = (n - 1) / 2.0
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x, y = ((ix - c) * L.period_um).ravel(), ((iy - c) * L.period_um).ravel()
    r = np.sqrt(x * x + y * y)
    keep = ~(r > D / 2)
    k = 2 * math.pi / lam
    phi = -k * (np.sqrt(r[keep] ** 2 + f * f) - f)
    ideal = oracle.axis_intensity(x[keep], y[keep], np.exp(1j * phi), [(k, f, 0)])[0]
    W = max(6 * lam * f / D, 4.0)
    from celeris_amd import Plane
    grid = Plane.psf(0.0, 0.0, f, k, 121, W)
    for trial in (0, 1, 255, 256, 299):
        jf = oracle.tolerance_fills(777, trial, 15.0, L.period_um, d["fill_map"])[keep]
        t = lib.transmission_for_fill(jf)
        _, I = oracle.forward(x[keep], y[keep], t, [grid], intensity=True)
        assert s[trial] == pytest.approx(I.max() / ideal, rel=1e-9), trial
