"""GPU parity of the FP32C forward on the matrix cores (crs::fwd_mfma).

A plane whose 16 x 32 pixel blocks meet the pair-error bound (rs_internal.h,
mfma_pair_error) has its interior summed as a rank-3 separable product per block and
pillar on v_mfma_f32_32x32x2_f32; its edge strips and every plane the bound refuses go to
the block kernels.  These tests select that path (forward block shape 16 x 32), check that
it ran (the launch's kernel name), and hold it to the FP32C contract:

* relative L2 of the complex field <= 1e-5 (|E|^2 <= 2e-5) against the FP64 kernel on the
  whole grid and against the FP64 oracle on sampled rows;
* per element, |E_mfma - E_fp64|_p <= 1e-5 sum_q |t_q| / R_pq (the pair budget 8e-6 plus
  the float sums), over mixed batches and random geometries near the admission boundary.
"""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import Plane, configs
from celeris_amd import rs as crs
from conftest import max_rel, rel_l2
from test_gpu_random_geometry import magnitude_sums, pixel_coords

pytestmark = pytest.mark.gpu

MFMA = (16, 32)
BOUND = 1e-5


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture
def mfma_ctx(rsctx):
    rsctx.set_block_shape(fwd=MFMA)
    yield rsctx
    rsctx.set_block_shape()


def forward(torch, ctx, px, py, t, planes, mode):
    print(f"[mfma forward] mode={mode} npil={np.size(px)} planes="
          + ", ".join(f"{P.nx}x{P.ny}" for P in planes), flush=True)
    f, inten = ctx.forward(dev(torch, px), dev(torch, py), dev(torch, np.atleast_2d(t)), planes,
                           mode=mode, field=True, intensity=True)
    torch.cuda.synchronize()
    return f.cpu().numpy(), inten.cpu().numpy()


@pytest.mark.parametrize("n,stride", [(161, 16), (321, 32), (1024, 128)])
def test_mfma_headline_grids(torch_cuda, mfma_ctx, n, stride):
    """91,809 pillars on the 161^2, 321^2 and 1024^2 grids: the matrix-core forward against
    the FP64 kernel (whole grid) and the FP64 oracle (every stride-th row)."""
    L = configs.config2(n=n)
    f32, i32 = forward(torch_cuda, mfma_ctx, L.px, L.py, L.t, L.planes, "fp32c")
    name, groups = mfma_ctx.last_launch(False)
    assert name == "crs::fwd_mfma" and groups > 0
    f64, _ = forward(torch_cuda, mfma_ctx, L.px, L.py, L.t, L.planes, "fp64")
    ref = oracle.forward(L.px, L.py, L.t, L.planes, row_stride=stride)
    mask = np.zeros((n, n), bool)
    mask[np.arange(0, n, stride)] = True
    mask = mask.ravel()
    efull = rel_l2(f32, f64)
    eint = rel_l2(i32, np.abs(f64) ** 2)
    eor = rel_l2(f32[mask], ref[mask])
    print(f"[mfma {n}^2] {groups} work-groups: vs fp64 {efull:.3e} (|E|^2 {eint:.3e}, maxrel "
          f"{max_rel(f32, f64):.3e}); vs oracle rows {eor:.3e}", flush=True)
    assert efull <= 1e-5 and eint <= 2e-5 and eor <= 1e-5
    # The block kernels' FP32C result on the same call agrees to the same contract.
    mfma_ctx.set_block_shape()
    fb, _ = forward(torch_cuda, mfma_ctx, L.px, L.py, L.t, L.planes, "fp32c")
    assert rel_l2(fb, f64) <= 1e-5 and rel_l2(f32, fb) <= 2e-5


def batch_planes(k, f):
    """Planes of one batch: windows the bound admits (whole blocks, ragged edges, a block
    narrower than the grid in one direction, an off-axis window, a zx slice, the second
    transmission set), and planes it refuses (too near the lens, too coarse, RS-I)."""
    lam = 2.0 * math.pi / k
    P = []
    P.append(Plane.psf(0.0, 0.0, f, k, 64, 2.0))                      # 4 x 2 blocks exactly
    P.append(Plane(k, -2.0, 0.07, -3.0, 0.065, f, 0.0, 50, 70))        # ragged: 3 x 2 + strips
    P.append(Plane(k, 1.0, 0.06, 0.5, 0.06, f * 1.3, 0.0, 17, 33))     # one block + 1-wide strips
    P.append(Plane(k, 12.0, -0.05, -8.0, 0.05, f, 0.0, 161, 40, 1))    # off axis, reversed x, set 1
    P.append(Plane(k, -1.5, 0.08, 0.0, 0.0, f - 4.0, 0.02, 40, 64))    # zx slice (depth per row)
    P.append(Plane(k, -1.5, 0.08, 0.0, 0.0, f - 4.0, 0.12, 40, 64))    # deeper zx slice
    P.append(Plane.psf(0.0, 0.0, 6.0, k, 40, 1.5))                     # near: refused
    P.append(Plane.psf(0.0, 0.0, f, k, 40, 40 * 4 * lam))              # coarse: refused
    P.append(Plane(k, -1.0, 0.06, -1.0, 0.06, f, 0.0, 48, 64, 0, -1, crs.KERNEL_RSI))  # RS-I
    return P


def test_mfma_mixed_batch_vs_oracle(torch_cuda, mfma_ctx):
    """One call over admitted and refused planes and two transmission sets: every element
    against the FP64 oracle within the per-element bound, every plane's relative L2."""
    torch = torch_cuda
    lam, p, n_side, f = 0.633, 0.35, 101, 100.0
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    k = 2.0 * math.pi / lam
    rng = np.random.default_rng(11)
    r = np.hypot(px, py)
    t0 = rng.uniform(0.7, 1.0, px.size) * np.exp(-1j * k * (np.sqrt(r * r + f * f) - f))
    t1 = rng.uniform(0.3, 1.0, px.size) * np.exp(1j * rng.uniform(-np.pi, np.pi, px.size))
    t = np.stack([t0, t1])
    planes = crs.resolve_offsets(batch_planes(k, f))
    E, _ = forward(torch, mfma_ctx, px, py, t, planes, "fp32c")
    assert mfma_ctx.last_launch(False)[0] == "crs::fwd_mfma"
    Eor = oracle.forward(px, py, t, planes)
    worst = 0.0
    for i, P in enumerate(planes):
        sl = slice(P.offset, P.offset + P.npix)
        fx, fy, fz = pixel_coords([P])
        S, _ = magnitude_sums(torch, px, py, np.abs(t[P.tset]), fx, fy, fz, np.zeros(P.npix))
        e = float(np.max(np.abs(E[sl] - Eor[sl]) / S))
        worst = max(worst, e)
        print(f"[mfma batch] plane {i} {P.nx}x{P.ny} z0={P.z0:.3g} set {P.tset}: max "
              f"|dE|/sum|t|/R {e:.2e}, relL2 {rel_l2(E[sl], Eor[sl]):.2e}", flush=True)
    assert worst <= BOUND


def sweep_geometry(seed):
    """Lenses and planes around the admission boundary: pixel pitches from 0.03 to 0.4
    wavelengths and planes from 15 to 400 um, psf windows anywhere in front of the lens and
    zx slices."""
    rng = np.random.default_rng(7000 + seed)
    lam = rng.uniform(0.45, 0.9)
    k = 2.0 * math.pi / lam
    p = rng.uniform(0.25, 0.5)
    n_side = int(rng.integers(20, 90))
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    t = rng.uniform(0.3, 1.0, px.size) * np.exp(1j * rng.uniform(-np.pi, np.pi, px.size))
    logu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))  # noqa: E731
    half = n_side * p / 2.0
    planes = []
    for _ in range(int(rng.integers(1, 4))):
        nx, ny = int(rng.integers(16, 90)), int(rng.integers(32, 100))
        pitch = lam * logu(0.03, 0.4)
        z = logu(15.0, 400.0)
        cx, cy = rng.uniform(-1.5 * half, 1.5 * half, 2)
        if rng.random() < 0.75:
            planes.append(Plane(k, float(cx) - 0.5 * (nx - 1) * pitch, pitch,
                                float(cy) - 0.5 * (ny - 1) * pitch, pitch, z, 0.0, nx, ny))
        else:
            planes.append(Plane(k, float(cx) - 0.5 * (nx - 1) * pitch, pitch, 0.0, 0.0, z,
                                pitch * logu(0.3, 3.0), nx, ny))
    return px, py, t, planes


@pytest.mark.parametrize("seed", list(range(24)))
def test_mfma_random_geometry(torch_cuda, mfma_ctx, seed):
    torch = torch_cuda
    px, py, t, planes = sweep_geometry(seed)
    planes = crs.resolve_offsets(planes)
    print(f"[mfma sweep {seed}] npil={px.size} planes="
          + ", ".join(f"{'zx' if P.sz else 'psf'} {P.nx}x{P.ny} z0={P.z0:.3g} s={P.sx:.3g}"
                      for P in planes), flush=True)
    dpx, dpy = dev(torch, px), dev(torch, py)
    dt = dev(torch, t[None, :])
    E32 = mfma_ctx.forward(dpx, dpy, dt, planes, mode="fp32c")[0].cpu().numpy()
    kernel = mfma_ctx.last_launch(False)[0]
    E64 = mfma_ctx.forward(dpx, dpy, dt, planes, mode="fp64")[0].cpu().numpy()
    fx, fy, fz = pixel_coords(planes)
    S, _ = magnitude_sums(torch, px, py, np.abs(t), fx, fy, fz, np.zeros(fx.size))
    fe = float(np.max(np.abs(E32 - E64) / S))
    print(f"[mfma sweep {seed}] {kernel}: max |dE|/sum|t|/R {fe:.2e}, relL2 "
          f"{rel_l2(E32, E64):.2e}", flush=True)
    assert fe <= BOUND
