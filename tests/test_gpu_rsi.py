"""GPU parity of the opt-in Rayleigh-Sommerfeld I kernel, exp(ikR)/R (z/R)(1/R - ik), which
the reference does not have (it computes exp(ikR)/R only, propagate.hpp:4-5): the FP64
and FP32 kernels against the oracle's RS-I restatement (itself pinned by the on-axis
closed form, -d/dz of the reference kernel and the long-double brute force in
tests/test_oracle.py), forward and adjoint, alone and mixed with reference-kernel
planes in one launch.

Tolerances: FP64 <= 1e-12 relative L2 of the complex field / gradient; FP32C (RS-I
planes in 4 x 4 blocks where the RS-I series bound admits them, else pair by pair at FP64
distances) <= 1e-5; FP32 (float R, k R, sincos, sums: the reference GPU kernel's
arithmetic) <= 2e-3.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

import oracle
from celeris_amd import KERNEL_RSI, Plane, configs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def planes_mixed():
    k = 2 * math.pi / 0.532
    return [
        replace(Plane.psf(0.0, 0.0, 50.0, k, 33, 5.0), kernel=KERNEL_RSI),
        Plane.psf(0.4, -0.3, 50.0, k, 17, 3.0),                             # reference kernel
        replace(Plane.zx(-6.0, 6.0, 25, 30.0, 70.0, 21, k), kernel=KERNEL_RSI),  # z per row
    ]


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5), ("fp32", 2e-3)])
def test_rsi_forward_and_adjoint_vs_oracle(torch_cuda, rsctx, mode, tol):
    torch = torch_cuda
    L = configs.config1()
    planes = planes_mixed()
    print(f"[rsi] mode={mode} npil={L.npil} planes={len(planes)}", flush=True)
    f, _ = rsctx.forward(dev(torch, L.px), dev(torch, L.py), dev(torch, L.t), planes, mode=mode,
                         field=True)
    torch.cuda.synchronize()
    f = f.cpu().numpy()
    ref = oracle.forward(L.px, L.py, L.t, planes)
    off = 0
    for p in planes:
        sl = slice(off, off + p.npix)
        assert rel_l2(f[sl], ref[sl]) <= tol, (p.kernel, rel_l2(f[sl], ref[sl]))
        off += p.npix
    # RS-I really differs from the reference kernel on the same plane
    plain = oracle.forward(L.px, L.py, L.t, [replace(planes[0], kernel=0)])
    assert rel_l2(plain, ref[: planes[0].npix]) > 0.5
    rng = np.random.default_rng(5)
    u = rng.standard_normal(f.size) + 1j * rng.standard_normal(f.size)
    g = rsctx.adjoint(dev(torch, L.px), dev(torch, L.py), dev(torch, u), planes, nsets=1, mode=mode)
    torch.cuda.synchronize()
    g = g.cpu().numpy().ravel()
    gref = oracle.adjoint(L.px, L.py, u, planes)[0]
    assert rel_l2(g, gref) <= tol


def test_rsi_fp32c_zero_depth_row_and_unknown_kernel(torch_cuda, rsctx):
    """A zx slice starting at z = 0: RS-I gives exactly 0 on that row (z/R = 0), which the
    FP32C exact path must keep; an unknown kernel id is refused."""
    torch = torch_cuda
    L = configs.small_lens(n_side=8, n=9)
    k = L.planes[0].k
    plane = replace(Plane.zx(-2.0, 2.0, 9, 0.0, 8.0, 5, k), kernel=KERNEL_RSI)
    f, _ = rsctx.forward(dev(torch, L.px), dev(torch, L.py), dev(torch, L.t), [plane], mode="fp32c")
    torch.cuda.synchronize()
    f = f.cpu().numpy().reshape(5, 9)
    ref = oracle.forward(L.px, L.py, L.t, [plane]).reshape(5, 9)
    assert np.all(f[0] == 0.0) and np.all(ref[0] == 0.0)
    assert rel_l2(f[1:], ref[1:]) <= 1e-5
    with pytest.raises(RuntimeError, match="unknown kernel"):
        rsctx.forward(dev(torch, L.px), dev(torch, L.py), dev(torch, L.t),
                      [replace(L.planes[0], kernel=7)], mode="fp64")


# ---- RS-I planes in FP32C blocks ---------------------------------------------------------
# Where the RS-I series bound admits 4 x 4 blocks (rs_internal.h, fp32c_pair_error with
# rsi), FP32C runs RS-I planes in crs::fwd_fp32c_rsi<4, 4> / crs::adj_fp32c_rsi<4, 4>: the
# reference kernel's pair loop with the RS-I phase shift, amplitude series and row depth
# ratio.  Held to the FP32C contract against the FP64 kernels: per pixel |dE| <= 1e-5
# sum_q |t_q| |G_pq| (|G| = z/R^2 sqrt(1/R^2 + k^2)), and the same per pillar for the
# adjoint.

RSI_FWD = "crs::fwd_fp32c_rsi<4, 4>"
RSI_ADJ = "crs::adj_fp32c_rsi<4, 4>"


def rsi_magnitude_sums(torch, px, py, at, fx, fy, fz, au, k):
    """S_p = sum_q at_q |G_pq| and T_q = sum_p au_p |G_pq| with the RS-I |G|, FP64 on the GPU."""
    d = torch.device("cuda:0")
    qx, qy = torch.from_numpy(px).to(d), torch.from_numpy(py).to(d)
    ta = torch.from_numpy(at).to(d)
    S = np.empty(fx.size)
    T = torch.zeros(px.size, dtype=torch.float64, device=d)
    step = max(1, (1 << 24) // max(1, px.size))
    for s in range(0, fx.size, step):
        e = min(fx.size, s + step)
        dx = torch.from_numpy(fx[s:e]).to(d)[:, None] - qx[None, :]
        dy = torch.from_numpy(fy[s:e]).to(d)[:, None] - qy[None, :]
        z = torch.from_numpy(fz[s:e]).to(d)[:, None]
        r2 = dx * dx + dy * dy + z * z
        g = z.abs() / r2 * torch.sqrt(1.0 / r2 + k * k)
        S[s:e] = (g @ ta).cpu().numpy()
        T += torch.from_numpy(au[s:e]).to(d) @ g
    return S, T.cpu().numpy()


def pixel_coords(planes):
    xs, ys, zs = [], [], []
    for P in planes:
        i = np.arange(P.nx, dtype=np.float64)
        j = np.arange(P.ny, dtype=np.float64)
        J, I = np.meshgrid(j, i, indexing="ij")
        xs.append((P.x0 + I * P.sx).ravel())
        ys.append((P.y0 + J * P.sy).ravel())
        zs.append((P.z0 + J * P.sz).ravel())
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)


def rsi_check(torch, ctx, px, py, t, u, planes, k, label):
    """FP32C against FP64 on RS-I planes: the per-pixel and per-pillar bounds; returns the
    kernels that ran and the worst ratios."""
    planes = crs_resolve(planes)
    E32, _ = ctx.forward(dev(torch, px), dev(torch, py), dev(torch, t), planes, mode="fp32c")
    kf = ctx.last_launch(False)[0]
    E64, _ = ctx.forward(dev(torch, px), dev(torch, py), dev(torch, t), planes, mode="fp64")
    g32 = ctx.adjoint(dev(torch, px), dev(torch, py), dev(torch, u), planes, mode="fp32c")
    ka = ctx.last_launch(True)[0]
    g64 = ctx.adjoint(dev(torch, px), dev(torch, py), dev(torch, u), planes, mode="fp64")
    torch.cuda.synchronize()
    E32, E64 = E32.cpu().numpy(), E64.cpu().numpy()
    g32, g64 = g32.cpu().numpy()[0], g64.cpu().numpy()[0]
    fx, fy, fz = pixel_coords(planes)
    S, T = rsi_magnitude_sums(torch, px, py, np.abs(t), fx, fy, fz, np.abs(u), k)
    ef = float(np.max(np.abs(E32 - E64) / S))
    eg = float(np.max(np.abs(g32 - g64) / T))
    print(f"[rsi blocks {label}] {kf} / {ka}: max |dE|/S {ef:.2e}, |dg|/T {eg:.2e}, "
          f"relL2 {rel_l2(E32, E64):.2e} / {rel_l2(g32, g64):.2e}", flush=True)
    assert ef <= 1e-5 and eg <= 1e-5
    return kf, ka


def crs_resolve(planes):
    from celeris_amd import rs as crs
    return crs.resolve_offsets(planes)


def test_rsi_blocks_headline(torch_cuda, rsctx):
    """The headline lens with its 161^2 focal plane and a 41 x 61 zx slice through the focus
    as RS-I planes: both run in the RS-I block kernels, within the FP32C contract."""
    L = configs.config2(n=161)
    k = L.planes[0].k
    planes = [replace(L.planes[0], kernel=KERNEL_RSI),
              replace(Plane.zx(-6.0, 6.0, 41, 80.0, 120.0, 61, k), kernel=KERNEL_RSI)]
    npix = sum(P.npix for P in planes)
    rng = np.random.default_rng(17)
    u = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    kf, ka = rsi_check(torch_cuda, rsctx, L.px, L.py, L.t.ravel(), u, planes, k, "headline")
    assert kf == RSI_FWD and ka == RSI_ADJ


def rsi_sweep_geometry(seed):
    """RS-I planes around the admission boundary: pitches 0.03 - 0.4 wavelengths, planes
    from 5 to 400 um (the nearest ones run pair by pair), psf windows and zx slices."""
    rng = np.random.default_rng(8100 + seed)
    lam = rng.uniform(0.45, 0.9)
    k = 2.0 * math.pi / lam
    p = rng.uniform(0.25, 0.5)
    n_side = int(rng.integers(16, 70))
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    logu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))  # noqa: E731
    half = n_side * p / 2.0
    planes = []
    for _ in range(int(rng.integers(1, 4))):
        nx, ny = int(rng.integers(5, 70)), int(rng.integers(5, 70))
        pitch = lam * logu(0.03, 0.4)
        z = logu(5.0, 400.0)
        cx, cy = rng.uniform(-1.5 * half, 1.5 * half, 2)
        if rng.random() < 0.7:
            planes.append(replace(Plane(k, float(cx) - 0.5 * (nx - 1) * pitch, pitch,
                                        float(cy) - 0.5 * (ny - 1) * pitch, pitch, z, 0.0, nx, ny),
                                  kernel=KERNEL_RSI))
        else:
            planes.append(replace(Plane(k, float(cx) - 0.5 * (nx - 1) * pitch, pitch, 0.0, 0.0, z,
                                        pitch * logu(0.3, 3.0), nx, ny), kernel=KERNEL_RSI))
    t = rng.uniform(0.2, 1.0, px.size) * np.exp(2j * math.pi * rng.uniform(0, 1, px.size))
    return px, py, t, planes, k


@pytest.mark.parametrize("seed", list(range(16)))
def test_rsi_blocks_random_geometry(torch_cuda, rsctx, seed):
    px, py, t, planes, k = rsi_sweep_geometry(seed)
    npix = sum(P.npix for P in planes)
    rng = np.random.default_rng(8200 + seed)
    u = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    rsi_check(torch_cuda, rsctx, px, py, t, u, planes, k, f"sweep {seed}")
