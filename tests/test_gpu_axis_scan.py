"""On-axis focus scans on the GPU (crs_axis_scan, crs_verify_achromatic_focus,
crs_polar_focus_check) against the reference's verify_achromatic_focus /
verify_pb_achromatic_focus outputs (tests/golden/achromatic.npz) and the oracle's
restatement of the scans (oracle.axis_scan, oracle.polar_focus_check).

FP64: the scan samples match the reference's z exactly and every intensity to 1e-9.
FP32C: intensities to 1e-5; a focus may land on a neighbouring sample only where the
FP64 curve has a near-tie there (within 2e-5 of the maximum).
"""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import analysis, configs
from conftest import achromatic_golden, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-9, "fp32c": 1e-5}


def _same_focus(mode, z_gpu, z_ref, curve_z, curve_I):
    """z_gpu is z_ref, or (FP32C) a sample whose FP64 intensity ties the maximum."""
    if z_gpu == z_ref:
        return True
    if mode == "fp64":
        return False
    j = int(np.argmin(np.abs(curve_z - z_gpu)))
    return curve_z[j] == z_gpu and curve_I[j] >= (1 - 2e-5) * curve_I.max()


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
@pytest.mark.parametrize("case", achromatic_golden(), ids=lambda c: f"{c[0]}-{c[1]}")
def test_verify_achromatic_focus_vs_reference(torch_cuda, case, mode):
    name, kind, px, py, t, lam, ci, lc, f, ref, lib, des = case
    print(f"[axis] verify {kind} {name} {mode}", flush=True)
    fn = analysis.verify_achromatic_focus if kind == "achro" else analysis.verify_pb_achromatic_focus
    pts = fn(lib, des, mode=mode)
    got = np.array([[p.wavelength_um, p.focal_length_um, p.rel_peak] for p in pts])
    assert np.array_equal(got[:, 0], ref[:, 0])
    zlo, zhi = 0.4 * f, 1.8 * f
    zs = zlo + (zhi - zlo) * np.arange(200) / 199
    for j in range(len(lam)):
        if got[j, 1] != ref[j, 1]:
            _, _, curve = oracle.axis_scan(px, py, t, [(j, 200, lam[j], zlo, zhi - zlo, f)])
            assert _same_focus(mode, got[j, 1], ref[j, 1], zs, curve), (j, got[j], ref[j])
        else:
            assert got[j, 2] == pytest.approx(ref[j, 2], rel=TOL[mode])
    print(mode, kind, name, "max rel err", np.max(np.abs(got[:, 2] / ref[:, 2] - 1)))


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_axis_scan_vs_oracle(torch_cuda, mode):
    """Several transmission sets and wavelengths in one call, nz = 1 scans, a scan that
    starts at the lens (z = 0.5 um) and a long one."""
    px, py = configs.square_grid(33, 0.4)
    lam0, f = 0.532, 18.0
    u = configs.uniform(configs.SEED, configs.STREAM_U, 3 * px.size).reshape(3, -1)
    t = np.stack([np.exp(1j * configs.hyperbolic_phase(px, py, lam0, f)),
                  (0.5 + 0.5 * u[0]) * np.exp(1j * configs.hyperbolic_phase(px, py, 0.633, 25.0)),
                  (0.7 + 0.3 * u[1]) * np.exp(2j * math.pi * u[2])])
    scans = [(0, 161, lam0, 4.5, 31.5, 0.0), (1, 97, 0.633, 10.0, 30.0, 1.0),
             (2, 33, 0.45, 0.5, 20.0, 2.0), (0, 1, lam0, f, 0.0, 3.0), (1, 1, 0.7, 22.0, 0.0, 4.0),
             (2, 400, 0.6, 5.0, 60.0, 5.0)]
    bz, bi, curves = analysis.axis_scan(px, py, t, scans, mode=mode)
    rbz, rbi, rcurves = oracle.axis_scan(px, py, t, scans)
    e = rel_l2(curves, rcurves)
    print(f"[axis] axis_scan {mode} curves relL2 {e:.3e}", flush=True)
    assert e <= TOL[mode]
    off = 0
    for s, (tset, nz, lam, zlo, zspan, zd) in enumerate(scans):
        zs = zlo + zspan * np.arange(nz) / max(nz - 1, 1)
        assert _same_focus(mode, bz[s], rbz[s], zs, rcurves[off:off + nz])
        assert bi[s] == pytest.approx(rcurves[off:off + nz].max(), rel=TOL[mode] * 10)
        off += nz


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
@pytest.mark.parametrize("nz", [240, 200])
def test_polar_focus_check_vs_oracle(torch_cuda, mode, nz):
    """The bifocal lens of cmd_polar.cpp's check: foci and isolations."""
    px, py = configs.square_grid(61, 0.35)
    lam, fx, fy = 0.532, 40.0, 55.0
    a = 0.8 + 0.2 * configs.uniform(configs.SEED, configs.STREAM_AMP, px.size)
    tx = a * np.exp(1j * configs.hyperbolic_phase(px, py, lam, fx))
    ty = a[::-1] * np.exp(1j * configs.hyperbolic_phase(px, py, lam, fy))
    got = analysis.polar_focus_check(px, py, tx, ty, lam, fx, fy, nz=nz, mode=mode)
    zx, zy, ix, iy = oracle.polar_focus_check(px, py, tx, ty, lam, fx, fy, nz=nz)
    print(f"[axis] polar {mode} nz={nz}: {got} vs {(zx, zy, ix, iy)}", flush=True)
    zlo, zhi = 0.5 * fx, 1.5 * fy
    zs = zlo + (zhi - zlo) * np.arange(nz) / (nz - 1)
    t2 = np.stack([tx, ty])
    _, _, cur = oracle.axis_scan(px, py, t2, [(0, nz, lam, zlo, zhi - zlo, zlo),
                                              (1, nz, lam, zlo, zhi - zlo, zlo)])
    assert _same_focus(mode, got.z_x_um, zx, zs, cur[:nz])
    assert _same_focus(mode, got.z_y_um, zy, zs, cur[nz:])
    if got.z_x_um == zx and got.z_y_um == zy:
        # isolation in dB: 10 log10 of a ratio known to TOL relative
        assert got.isolation_x_db == pytest.approx(ix, abs=10 * TOL[mode])
        assert got.isolation_y_db == pytest.approx(iy, abs=10 * TOL[mode])
    same = analysis.polar_focus_check(px, py, tx, ty, lam, fx, fx, nz=nz, mode=mode)
    assert same.isolation_x_db == 0.0 and same.isolation_y_db == 0.0


@pytest.mark.parametrize("mode", ["fp64", "fp32c"])
def test_profile_focus_z_vs_oracle(torch_cuda, mode):
    """cli_common.cpp:112-133: 240 samples over [0.5 f, 1.5 f], z = 0.5 f + f j / 239."""
    px, py = configs.square_grid(51, 0.4)
    lam, f = 0.633, 35.0
    t = np.exp(1j * configs.hyperbolic_phase(px, py, lam, f))
    z = analysis.profile_focus_z(px, py, t, lam, f, mode=mode)
    rbz, _, cur = oracle.axis_scan(px, py, t, [(0, 240, lam, 0.5 * f, f, 0.5 * f)])
    zs = 0.5 * f + f * np.arange(240) / 239
    assert _same_focus(mode, z, rbz[0], zs, cur)
    assert 0.8 * f < z < 1.05 * f


def test_axis_scan_rejects_bad_scans(torch_cuda):
    px, py = np.zeros(4), np.arange(4.0)
    t = np.ones((1, 4), np.complex128)
    for bad in [(1, 10, 0.5, 1.0, 1.0, 0.0), (0, 0, 0.5, 1.0, 1.0, 0.0), (0, 5, 0.0, 1.0, 1.0, 0.0)]:
        with pytest.raises(RuntimeError):
            analysis.axis_scan(px, py, t, [bad])
