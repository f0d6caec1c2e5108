"""Pins the FP64 oracle (oracle/rs_oracle.c) before anything is compared against it.

Checks, each chosen so that a plausible slip (a dropped 1/R, a sign of the phase, a
transposed index, a wrong pixel step, a missing conjugate) fails at least one:
  * the reference's own CPU code (oracle/_ref, compiled from its sources), directly
    where it is built and through the committed golden vectors everywhere;
  * closed forms: one pillar, the in-phase focus of a hyperbolic lens;
  * long-double brute force on tiny inputs;
  * linearity and mirror symmetry;
  * the adjoint: dot-product identity and column-by-column construction;
  * the on-axis focus scans: the reference's verify_achromatic_focus /
    verify_pb_achromatic_focus outputs, and closed forms for the scan rule and the
    polarisation isolation.
"""
import math
import os
from dataclasses import replace

import numpy as np
import pytest

import oracle
from celeris_amd import KERNEL_RSI, Plane, configs
from celeris_amd.design import MetalensDesign, UnitCellLibrary, aperture_pillars
from conftest import achromatic_golden, dispersive_golden, rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def plane_for(cx, cy, z, lam, n, W):
    return Plane.psf(cx, cy, z, 2 * math.pi / lam, int(n), W)


def load_pillar_golden():
    d = np.load(os.path.join(GOLDEN, "propagate_pillars.npz"))
    names = sorted({k.split("__")[0] for k in d.files})
    return [(nm, d[f"{nm}__px"], d[f"{nm}__py"], d[f"{nm}__t"], d[f"{nm}__args"], d[f"{nm}__I"])
            for nm in names]


@pytest.mark.parametrize("case", load_pillar_golden(), ids=lambda c: c[0])
def test_oracle_matches_reference_golden(case):
    """Golden |E|^2 written by celeris::propagate_pillars (focal.cpp:205-238)."""
    name, px, py, t, args, I_ref = case
    cx, cy, z, lam, n, W = args
    f, I = oracle.forward(px, py, t[None], [plane_for(cx, cy, z, lam, n, W)], intensity=True)
    err = np.max(np.abs(I - I_ref.ravel())) / np.max(np.abs(I_ref))
    assert err <= 1e-13, err


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_oracle_matches_reference_build_random():
    rng = np.random.default_rng(123)
    for trial in range(4):
        npil = int(rng.integers(1, 60))
        px = rng.uniform(-5, 5, npil)
        py = rng.uniform(-5, 5, npil)
        t = rng.uniform(0.2, 1, npil) * np.exp(1j * rng.uniform(-4, 4, npil))
        cx, cy = rng.uniform(-2, 2, 2)
        z = rng.uniform(2, 40)
        lam = rng.uniform(0.4, 0.8)
        n = int(rng.integers(2, 12))
        W = rng.uniform(0.5, 4)
        I_ref = oracle.ref_propagate_pillars(px, py, t, cx, cy, z, lam, n, W)
        _, I = oracle.forward(px, py, t[None], [plane_for(cx, cy, z, lam, n, W)], intensity=True)
        assert np.max(np.abs(I - I_ref.ravel())) / np.max(I_ref) <= 1e-13


def test_oracle_psf_of_lens_matches_reference_compute_psf():
    """compute_psf (focal.cpp:148-203) of a fill-map lens, through aperture_pillars."""
    d = np.load(os.path.join(GOLDEN, "lens_analyses.npz"))
    lens = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    px, py, t = aperture_pillars(lens, lib, float(d["D"]))
    n, W = int(d["psf_n"][0]), float(d["psf_n"][1])
    _, I = oracle.forward(px, py, t[None], [plane_for(0, 0, float(d["f"]), float(d["lam"]), n, W)],
                          intensity=True)
    assert np.max(np.abs(I - d["psf"])) / np.max(d["psf"]) <= 1e-13


def test_oracle_zx_and_axis_match_reference_through_focus():
    """analyze_through_focus (throughfocus.cpp:15-132): the 161-point on-axis scan and
    the 81x81 CPU caustic, both normalized to their peak and stored as float."""
    d = np.load(os.path.join(GOLDEN, "lens_analyses.npz"))
    lens = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    f, lam, D = float(d["f"]), float(d["lam"]), float(d["D"])
    px, py, t = aperture_pillars(lens, lib, D)
    tf = d["throughfocus"]
    N = int(tf[0])
    z_ref, I_ref = tf[1:1 + N], tf[1 + N:1 + 2 * N]
    cnx, cnz = int(tf[3 + 2 * N]), int(tf[4 + 2 * N])
    caustic_ref = tf[5 + 2 * N:5 + 2 * N + cnx * cnz]
    k = 2 * math.pi / lam
    na = math.sin(math.atan((D / 2) / f))
    half = max(2.0 * lam / (na * na + 1e-6), 0.25 * f)
    zlo, zhi = max(0.1 * f, f - half), f + half
    axis = oracle.forward(px, py, t[None], [Plane.axis(zlo, zhi, N, k)])
    Ia = np.abs(axis) ** 2
    assert np.max(np.abs(Ia / Ia.max() - I_ref)) <= 2e-7  # float storage
    assert np.max(np.abs(zlo + (zhi - zlo) * np.arange(N) / (N - 1) - z_ref)) <= 1e-5
    Wx = max(8.0 * lam * f / D, 3.0)
    zx = oracle.forward(px, py, t[None], [Plane.zx(-Wx, Wx, cnx, zlo, zhi, cnz, k)])
    Iz = np.abs(zx) ** 2
    assert np.max(np.abs(Iz / Iz.max() - caustic_ref)) <= 2e-7


def test_single_pillar_closed_form():
    k = 2 * math.pi / 0.55
    plane = Plane.psf(0.3, -0.1, 12.0, k, 7, 2.0)
    a, phi, x, y = 0.7, -2.1, 1.5, 0.4
    f = oracle.forward(np.array([x]), np.array([y]), np.array([[a * np.exp(1j * phi)]]), [plane])
    i = np.arange(7)
    fx = plane.x0 + i * plane.sx
    fy = plane.y0 + i * plane.sy
    R = np.sqrt((fx[None, :] - x) ** 2 + (fy[:, None] - y) ** 2 + 144.0)
    exact = a * np.exp(1j * (phi + k * R)) / R
    assert rel_l2(f, exact.ravel()) <= 1e-14


def test_inphase_focus_identity():
    L = configs.config1()
    k = 2 * math.pi / 0.532
    f = oracle.forward(L.px, L.py, L.t, [Plane.psf(0, 0, L.focal, k, 3, 1.0)])
    R = np.sqrt(L.px ** 2 + L.py ** 2 + L.focal ** 2)
    exact = np.exp(1j * k * L.focal) * np.sum(np.abs(L.t[0]) / R)
    assert abs(f[4] - exact) / abs(exact) <= 1e-12


def test_long_double_brute_force():
    L = configs.small_lens(n_side=9, n=5)
    f = oracle.forward(L.px, L.py, L.t, L.planes)
    fl = oracle.forward_ld(L.px, L.py, L.t, L.planes)
    assert rel_l2(f, fl) <= 1e-12


def test_linearity_and_mirror_symmetry():
    px, py = configs.square_grid(8, 0.5)
    rng = np.random.default_rng(9)
    t1 = rng.standard_normal(px.size) + 1j * rng.standard_normal(px.size)
    t2 = rng.standard_normal(px.size) + 1j * rng.standard_normal(px.size)
    k = 2 * math.pi / 0.6
    plane = Plane.psf(0, 0, 9.0, k, 9, 2.0)
    e1 = oracle.forward(px, py, t1[None], [plane])
    e2 = oracle.forward(px, py, t2[None], [plane])
    e12 = oracle.forward(px, py, (2 * t1 - 1j * t2)[None], [plane])
    assert rel_l2(e12, 2 * e1 - 1j * e2) <= 1e-13
    # Mirror the aperture in x (square grid is x-symmetric): the field mirrors in x.
    order = np.lexsort((-px, py))  # pillar that sits at (-x, y)
    tm = t1[order]
    em = oracle.forward(px, py, tm[None], [plane]).reshape(9, 9)
    assert rel_l2(em, e1.reshape(9, 9)[:, ::-1]) <= 1e-12


def test_adjoint_dot_product_and_columns():
    px, py = configs.square_grid(6, 0.4)
    k = 2 * math.pi / 0.5
    planes = [Plane.psf(0, 0, 8.0, k, 5, 1.0), Plane.zx(-1, 1, 4, 6, 10, 3, k)]
    npix = 25 + 12
    rng = np.random.default_rng(4)
    u = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    g = oracle.adjoint(px, py, u, planes)[0]
    # column q of G is the field of a unit transmission at pillar q
    for q in (0, 7, 35):
        e = np.zeros(px.size, complex)
        e[q] = 1.0
        col = oracle.forward(px, py, e[None], planes)
        assert abs(g[q] - np.vdot(col, u)) <= 1e-12 * np.abs(g).max()
    t = rng.standard_normal(px.size) + 1j * rng.standard_normal(px.size)
    lhs = np.vdot(u, oracle.forward(px, py, t[None], planes))
    rhs = np.vdot(g, t)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_sets_follow_planes():
    px, py = configs.square_grid(5, 0.4)
    k = 2 * math.pi / 0.5
    planes = [Plane.psf(0, 0, 8.0, k, 3, 1.0, tset=1), Plane.psf(0, 0, 9.0, k, 3, 1.0, tset=1)]
    u = np.ones(18, complex)
    g = oracle.adjoint(px, py, u, planes, nsets=2)
    assert np.all(g[0] == 0) and np.any(g[1] != 0)
    g_a = oracle.adjoint(px, py, u[:9], planes[:1], nsets=2)[1]
    g_b = oracle.adjoint(px, py, u[9:], [Plane.psf(0, 0, 9.0, k, 3, 1.0, tset=1)], nsets=2)[1]
    assert rel_l2(g[1], g_a + g_b) <= 1e-14


def test_row_stride_sample_matches_full():
    L = configs.small_lens(n_side=10, n=9)
    full = oracle.forward(L.px, L.py, L.t, L.planes)
    part = oracle.forward(L.px, L.py, L.t, L.planes, row_stride=4)
    rows = np.zeros((9, 9), bool)
    rows[::4] = True
    assert np.array_equal(part[rows.ravel()], full[rows.ravel()])
    assert np.all(part[~rows.ravel()] == 0)


def _mtf_golden(name):
    d = np.load(os.path.join(GOLDEN, name))
    lens = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    ref = d["mtf"]
    k = int(ref[0])
    return lens, lib, float(d["f"]), float(d["lam"]), float(d["D"]), \
        (ref[1:1 + k], ref[1 + k:1 + 2 * k], ref[1 + 2 * k:1 + 3 * k], ref[1 + 3 * k])


@pytest.mark.parametrize("name", ["lens_analyses.npz", "mtf_large.npz"])
def test_oracle_mtf_matches_reference(name):
    """analyze_mtf (mtf.cpp:10-63): the reference's float curve, reproduced bit for bit
    from exit_pupil + the oracle's autocorrelation."""
    from celeris_amd.design import exit_pupil

    lens, lib, f, lam, D, (rf, rm, ri, rc) = _mtf_golden(name)
    P, P0 = exit_pupil(lens, lib, f, lam, D)
    fr, m, mi, cut = oracle.mtf(P, P0, lens.period_um, f, lam, D)
    assert fr.size == rf.size
    assert np.array_equal(fr, rf.astype(np.float32))
    assert np.array_equal(m, rm.astype(np.float32))
    assert np.array_equal(mi, ri.astype(np.float32))
    assert cut == rc


def test_oracle_mtf_invariants():
    """A perfect pupil's MTF is the mask's; a phase ramp along x, e^{i a x}, leaves the
    MTF along x unchanged (|acc(s)| gains only e^{-i a s}); a uniform disc's MTF falls
    from 1 and reaches 0 past the aperture width (no overlap)."""
    n, p, f, lam, D = 41, 0.5, 20.0, 0.5, 10.0
    c = (n - 1) / 2
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    r = np.hypot((ix - c) * p, (iy - c) * p)
    P0 = (~(r > D / 2)).astype(np.float64)
    fr, m, mi, cut = oracle.mtf(P0.astype(np.complex128), P0, p, f, lam, D)
    assert np.array_equal(m, mi) and m[0] == 1.0 and np.all(np.diff(mi) <= 0)
    assert mi[-1] == 0.0 and cut == D / (lam * f) * 1000.0
    rng = np.random.default_rng(3)
    P = P0 * np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n)))
    _, m1, _, _ = oracle.mtf(P, P0, p, f, lam, D)
    _, m2, _, _ = oracle.mtf(P * np.exp(1j * 0.7 * ix), P0, p, f, lam, D)
    assert np.allclose(m1, m2, rtol=1e-6, atol=1e-7)
    # ... but not under a conjugate-free correlation (catches a dropped conj)
    acc = [np.sum(P[:, :n - s] * P[:, s:]) for s in range(5)]
    acc2 = [np.sum((P * np.exp(1j * 0.7 * ix))[:, :n - s] * (P * np.exp(1j * 0.7 * ix))[:, s:])
            for s in range(5)]
    assert not np.allclose(np.abs(acc), np.abs(acc2), rtol=1e-3)


def _fov_windows(px, py, t, f, lam, spots, r2):
    """The oracle's version of field.cpp's tilted-spot search: per spot (theta, xc, W,
    x_patch) an 81 x 81 window at z = f of the lit cells under exp(i k sin(theta) x)."""
    k = 2 * math.pi / lam
    NG = 81
    out = []
    for theta, xc, W, xp in spots:
        lit = np.ones(px.size, bool) if r2 < 0 else ~((px - xp) ** 2 + py ** 2 > r2)
        ts = np.where(lit, t * np.exp(1j * k * math.sin(theta) * px), 0)
        pl = Plane(k, xc - W, 2 * W / (NG - 1), -W, 2 * W / (NG - 1), f, 0.0, NG, NG)
        I = np.abs(oracle.forward(px, py, ts[None], [pl])) ** 2
        j, i = np.unravel_index(np.argmax(I), (NG, NG))  # first maximum, rows outer
        out.append((I.max(), xc - W + 2.0 * W * i / (NG - 1)))
    return out


def test_oracle_field_of_view_matches_reference():
    """analyze_field_of_view / analyze_wide_fov (field.cpp:12-148) restated as tilted
    transmission sets on 81 x 81 windows: the reference's rel_strehl and spot shifts."""
    d = np.load(os.path.join(GOLDEN, "field_of_view.npz"))
    lens = MetalensDesign(int(d["n_cells"]), float(d["period"]), d["fill_map"])
    lib = UnitCellLibrary(float(d["period"]), d["lib_fill"], d["lib_phase"], d["lib_amp"])
    f, lam, D = float(d["f"]), float(d["lam"]), float(d["D"])
    px, py, t = aperture_pillars(lens, lib, D)
    dl = lam * f / D
    sp = [(0.0, 0.0, max(3 * dl, 2.0), 0.0)]
    for a in d["fov_angles"]:
        th = math.radians(a)
        sp.append((th, f * math.tan(th), max(3 * dl, 2.0) + abs(f * math.tan(th) - f * math.sin(th)), 0.0))
    res = _fov_windows(px, py, t, f, lam, sp, -1.0)
    ref = d["fov"].reshape(-1, 3)
    for a, (pk, x), r in zip(d["fov_angles"], res[1:], ref):
        assert r[0] == a and pk / res[0][0] == pytest.approx(r[1], rel=1e-12)
        assert x == pytest.approx(r[2], abs=1e-12)
    Dstop, sdist = d["widefov_args"]
    dls = lam * f / Dstop
    sp = []
    for a in np.concatenate([[0.0], d["widefov_angles"]]):
        th = math.radians(a)
        xs, xt = f * math.sin(th), f * math.tan(th)
        sp.append((th, 0.5 * (xs + xt), max(3 * dls, 2.0) + 0.5 * abs(xt - xs), sdist * math.tan(th)))
    res = _fov_windows(px, py, t, f, lam, sp, (Dstop / 2) ** 2)
    ref = d["widefov"].reshape(-1, 3)
    for a, (pk, x), r in zip(d["widefov_angles"], res[1:], ref):
        assert r[0] == a and pk / res[0][0] == pytest.approx(r[1], rel=1e-12)
        assert x == pytest.approx(r[2], abs=1e-12)


@pytest.mark.parametrize("case", achromatic_golden(), ids=lambda c: f"{c[0]}-{c[1]}")
def test_oracle_achromatic_focus_matches_reference(case):
    """verify_achromatic_focus (achromatic.cpp:288-338) and verify_pb_achromatic_focus
    (pb_achromatic.cpp:255-301): the restated scan picks the reference's z sample in
    every band and its relative peak to 1e-12 (the transmissions are rebuilt with
    numpy's sin/cos, the reference's with glibc's)."""
    name, kind, px, py, t, lam, ci, lc, f, ref = case[:10]
    got = oracle.verify_achromatic_focus(px, py, t, lam, ci, lc, f)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.array_equal(got[:, 1], ref[:, 1])
    assert np.allclose(got[:, 2], ref[:, 2], rtol=1e-12, atol=0)


def test_oracle_axis_scan_closed_forms():
    """A single pillar at the origin: I(z) = a^2 / z^2 exactly decreasing, so every scan
    stops at its first sample; a pillar at (0, 0) with t = 0 gives I = 0 everywhere, and
    the callers' rule (I > best_I from -1) still takes the first sample; nz = 1 is z = zlo;
    the sample grid is zlo + zspan j / (nz - 1)."""
    px, py = np.zeros(1), np.zeros(1)
    t = np.array([[0.8 + 0.6j], [0.0j]])
    scans = [(0, 7, 0.5, 3.0, 9.0, -5.0), (1, 4, 0.6, 2.0, 1.5, -5.0), (0, 1, 0.5, 4.0, 0.0, -5.0)]
    bz, bi, I = oracle.axis_scan(px, py, t, scans)
    assert list(bz) == [3.0, 2.0, 4.0]
    assert bi[0] == pytest.approx(1.0 / 9.0, rel=1e-15) and bi[1] == 0.0
    z = 3.0 + 9.0 * np.arange(7) / 6
    assert np.allclose(I[:7], 1.0 / z ** 2, rtol=1e-15)


def test_oracle_polar_focus_closed_form():
    """cmd_polar.cpp:85-126 on one pillar with t_x / t_y = 2: both channels peak at the
    first sample (I = |t|^2 / z^2 decreases) and each isolation is 10 log10(|t_x / t_y|^2)
    = +-6.0206 dB; equal focal lengths give no isolation; swapping the channels swaps
    the answers."""
    px, py = np.zeros(1), np.zeros(1)
    zx, zy, ix, iy = oracle.polar_focus_check(px, py, [1.0 + 0j], [0.5j], 0.532, 20.0, 30.0)
    assert zx == zy == 10.0
    assert ix == pytest.approx(10 * math.log10(4.0), rel=1e-14)
    assert iy == pytest.approx(-10 * math.log10(4.0), rel=1e-14)
    assert oracle.polar_focus_check(px, py, [1.0 + 0j], [0.5j], 0.532, 20.0, 20.0)[2:] == (0.0, 0.0)


def test_oracle_polar_focus_bifocal_lens():
    """A bifocal lens (X channel hyperbolic to f_x, Y to f_y): each channel focuses near
    its own plane, before the other channel's, and is brighter there; swapping the
    channels swaps the foci and the isolations."""
    px, py = configs.square_grid(41, 0.4)
    lam, fx, fy = 0.532, 30.0, 45.0
    tx = np.exp(1j * configs.hyperbolic_phase(px, py, lam, fx))
    ty = np.exp(1j * configs.hyperbolic_phase(px, py, lam, fy))
    zx, zy, ix, iy = oracle.polar_focus_check(px, py, tx, ty, lam, fx, fy, nz=120)
    assert 0.8 * fx < zx < 1.05 * fx and 0.8 * fy < zy < 1.05 * fy
    assert ix > 3.0 and iy > 3.0
    zy2, zx2, iy2, ix2 = oracle.polar_focus_check(px, py, ty, tx, lam, fx, fy, nz=120)
    assert (zx2, zy2) == (zx, zy) and ix2 == pytest.approx(ix, rel=1e-12)


def test_oracle_mt19937_known_answer():
    """C++ [rand.predef]: the 10000th output of std::mt19937 from the default seed 5489 is
    4123659995 (the restated engine under the tolerance jitter)."""
    assert oracle.mt19937_nth(5489, 10000) == 4123659995


def tolerance_golden():
    d = np.load(os.path.join(GOLDEN, "tolerance.npz"))
    return d, sorted({k.split("__")[0] for k in d.files if k.endswith("__out")})


@pytest.mark.parametrize("name", tolerance_golden()[1])
def test_oracle_tolerance_matches_reference(name):
    """analyze_tolerance (tolerance.cpp:12-48) from the reference build: the restated
    mt19937 / normal_distribution draws give the reference's Strehl mean and worst to
    1e-12 and its spread to 1e-8 absolute (the reference computes the variance
    sum2/n - mean^2 with a fused multiply-add, so a single trial's spread is ~1e-9, not 0)."""
    d, _ = tolerance_golden()
    trials, seed = d[f"{name}__args"]
    got = oracle.tolerance(int(d["n_cells"]), float(d["period"]), d["fill_map"], d["lib_fill"],
                           d["lib_phase"], d["lib_amp"], float(d["f"]), float(d["lam"]),
                           float(d["D"]), d[f"{name}__sigmas"], int(trials), int(seed))
    ref = d[f"{name}__out"]
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[:, [1, 3]], ref[:, [1, 3]], rtol=1e-12, atol=0)
    assert np.allclose(got[:, 2], ref[:, 2], rtol=1e-9, atol=1e-8)
    if name == "cli":
        assert np.all(ref[1:, 2] > 1e-4)  # the jitter really spreads the trials


def test_oracle_tolerance_jitter_properties():
    """The jitter at sigma = 0 leaves every fill in [0.05, 0.95] unchanged and clamps the
    rest; a different trial or sigma gives a different map; draws are N(0, sigma_fill):
    over 40,000 cells the sample mean and spread match within 4 standard errors."""
    n = 200
    fm = np.linspace(0.0, 1.0, n * n)
    same = oracle.tolerance_fills(12345, 0, 0.0, 0.4, fm)
    assert np.array_equal(same, np.clip(fm, 0.05, 0.95))
    mid = np.full(n * n, 0.5)
    a = oracle.tolerance_fills(12345, 0, 10.0, 0.4, mid) - 0.5
    b = oracle.tolerance_fills(12345, 1, 10.0, 0.4, mid) - 0.5
    c = oracle.tolerance_fills(12345, 0, 11.0, 0.4, mid) - 0.5
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    sf = 0.010 / 0.4
    assert abs(a.mean()) < 4 * sf / math.sqrt(a.size)
    assert abs(a.std() / sf - 1) < 4 / math.sqrt(2 * a.size)


def test_oracle_chromatic_dispersive_matches_reference():
    """analyze_chromatic_dispersive (chromatic.cpp:60-127) through the reference's own code
    (tests/golden/chromatic_dispersive.npz): the restated scan, on the transmissions each
    wavelength's library gives the aperture cells, picks the reference's z sample in every
    band and its relative peak to 1e-12."""
    from celeris_amd.design import aperture_fills, aperture_pillars
    lens, build, f, lam, D, lmin, lmax, n, ref, calls = dispersive_golden()
    px, py, t0 = aperture_pillars(lens, build(lam), D)
    fills = aperture_fills(lens, D)
    bands = [build(lmin + (lmax - lmin) * i / (n - 1)).transmission_for_fill(fills) for i in range(n)]
    t_sets = np.vstack([t0[None], np.array(bands)])
    got = oracle.chromatic(px, py, t_sets, 0, list(range(1, n + 1)), f, lam, lmin, lmax, n)
    assert np.array_equal(got[:, 0], ref[:, 0]) and np.array_equal(got[:, 1], ref[:, 1])
    assert np.allclose(got[:, 2], ref[:, 2], rtol=1e-12, atol=0)
    # the fixed-transmission scan differs (the libraries really disperse)
    fixed = oracle.chromatic(px, py, t0[None], 0, [0] * n, f, lam, lmin, lmax, n)
    assert np.max(np.abs(fixed[:, 2] - ref[:, 2])) > 1e-3


# ---- the opt-in RS-I kernel exp(ikR)/R (z/R)(1/R - ik) (no reference counterpart) ----

def _rsi(p):
    return replace(p, kernel=KERNEL_RSI)


def test_rsi_single_pillar_on_axis_closed_form():
    # Pixel straight above the pillar: R = z, z/R = 1, so G = exp(ikz) (1/z - ik) / z.
    k = 2 * math.pi / 0.633
    x, y, z, t = 1.3, -0.7, 37.0, 0.8 * np.exp(0.3j)
    plane = _rsi(Plane(k, x, 0.0, y, 0.0, z, 0.0, 1, 1))
    e = oracle.forward(np.array([x]), np.array([y]), np.array([[t]]), [plane])
    exact = t * np.exp(1j * k * z) * (1.0 / z - 1j * k) / z
    assert abs(e[0] - exact) <= 1e-14 * abs(exact)
    # Far from the axis the factor tends to -ik z/R: |G| R / (k z/R) -> 1.
    far = _rsi(Plane(k, x + 4000.0, 0.0, y, 0.0, z, 0.0, 1, 1))
    g = oracle.forward(np.array([x]), np.array([y]), np.array([[1.0]]), [far])[0]
    R = math.hypot(4000.0, z)
    assert abs(abs(g) * R / (k * z / R) - 1.0) <= 1e-6


def test_rsi_is_minus_z_derivative_of_reference_kernel():
    # G_RSI = -d/dz [exp(ikR)/R]: central differences of the reference-kernel sum in the
    # plane depth (tilted plane: every row moves by the same dz).
    L = configs.small_lens(n_side=7, n=5)
    base = Plane(2 * math.pi / 0.55, -1.0, 0.5, -0.8, 0.4, 20.0, 0.3, 5, 4)
    h = 1e-4
    ep = oracle.forward(L.px, L.py, L.t, [replace(base, z0=base.z0 + h)])
    em = oracle.forward(L.px, L.py, L.t, [replace(base, z0=base.z0 - h)])
    rsi = oracle.forward(L.px, L.py, L.t, [_rsi(base)])
    assert rel_l2(rsi, -(ep - em) / (2 * h)) <= 1e-6


def test_rsi_long_double_brute_force_and_adjoint():
    L = configs.small_lens(n_side=9, n=5)
    planes = [_rsi(p) for p in L.planes]
    f = oracle.forward(L.px, L.py, L.t, planes)
    fl = oracle.forward_ld(L.px, L.py, L.t, planes)
    assert rel_l2(f, fl) <= 1e-12
    # dot-product identity with the adjoint on the same RS-I planes
    rng = np.random.default_rng(11)
    u = rng.standard_normal(f.size) + 1j * rng.standard_normal(f.size)
    g = oracle.adjoint(L.px, L.py, u, planes)[0]
    assert abs(np.vdot(u, f) - np.vdot(g, L.t[0])) <= 1e-12 * abs(np.vdot(u, f))
