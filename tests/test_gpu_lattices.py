"""General pixel lattices through every mode: the plane descriptor is an affine lattice
fx = x0 + i sx, fy = y0 + j sy, z = z0 + j sz, and the psf / zx windows the analyses
build use only part of it.  These cases take the rest: reversed axes (negative steps),
tilted planes (sy and sz both non-zero), planes behind the lens plane (z < 0), a zx slice
walking towards the lens, and the RS-I kernel on them.

Checks, per output element with the bound of the random-geometry sweep:
  * FP64 against the oracle on the same inputs (FP64 roundoff of k R);
  * FP32C against FP64, |dE_p| <= 1e-5 sum_q |t_q| |G_pq|, and the adjoint likewise;
and invariants the mathematics fixes, so that a sign or index slip in the lattice
handling cannot pass as a consistent error of both paths:
  * a plane with both steps negated and the origin at the far corner holds the same
    pixels as the plain window, in reverse order;
  * a plane mirrored through the lens plane (z -> -z) gets the same field from
    exp(ikR)/R, which depends on z^2 only, and the negated field from RS-I, whose
    obliquity factor z/R is odd in z.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

import oracle
from celeris_amd import Plane
from celeris_amd import rs as crs
from conftest import rel_l2
from test_gpu_random_geometry import pixel_coords

pytestmark = pytest.mark.gpu

BOUND = 1e-5


def lens(n_side=40, p=0.35, lam=0.6, f=30.0, seed=3):
    rng = np.random.default_rng(seed)
    k = 2.0 * math.pi / lam
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    r = np.hypot(px, py)
    t = rng.uniform(0.6, 1.0, px.size) * np.exp(-1j * k * (np.sqrt(r * r + f * f) - f))
    return px, py, t, k, lam, f


def window(k, n, W, z, cx=0.3, cy=-0.2):
    return Plane.psf(cx, cy, z, k, n, W)


def reversed_window(P):
    """The pixels of P in reverse order along both axes."""
    return Plane(P.k, P.x0 + (P.nx - 1) * P.sx, -P.sx, P.y0 + (P.ny - 1) * P.sy, -P.sy,
                 P.z0, P.sz, P.nx, P.ny)


def mirrored(P):
    """P reflected through the lens plane z = 0."""
    return Plane(P.k, P.x0, P.sx, P.y0, P.sy, -P.z0, -P.sz, P.nx, P.ny, kernel=P.kernel)


def tilted(k, nx, ny, step, z, angle_deg):
    """A plane through (0, 0, z) tilted about the x axis: row j moves by step (cos a, sin a)."""
    a = math.radians(angle_deg)
    sy, sz = step * math.cos(a), step * math.sin(a)
    return Plane(k, -0.5 * (nx - 1) * step, step, -0.5 * (ny - 1) * sy, sy,
                 z - 0.5 * (ny - 1) * sz, sz, nx, ny)


def magnitude_sums(torch, px, py, at, planes, au):
    """S_p = sum_q at_q |G_pq| and T_q = sum_p au_p |G_pq| in FP64 on the GPU, with |G| =
    1/R for exp(ikR)/R and |z| / R^2 sqrt(1/R^2 + k^2) for RS-I."""
    dev = torch.device("cuda:0")
    qx = torch.from_numpy(px).to(dev)
    qy = torch.from_numpy(py).to(dev)
    ta = torch.from_numpy(at).to(dev)
    S, T = [], torch.zeros(px.size, dtype=torch.float64, device=dev)
    for P in crs.resolve_offsets(planes):
        fx, fy, fz = pixel_coords([P])
        dx = torch.from_numpy(fx).to(dev)[:, None] - qx[None, :]
        dy = torch.from_numpy(fy).to(dev)[:, None] - qy[None, :]
        z = torch.from_numpy(fz).to(dev)[:, None]
        inv = torch.rsqrt(dx * dx + dy * dy + z * z)
        if P.kernel == crs.KERNEL_RSI:
            inv = torch.abs(z) * inv * inv * torch.sqrt(inv * inv + P.k * P.k)
        S.append((inv @ ta).cpu().numpy())
        T += torch.from_numpy(au[P.offset:P.offset + fx.size]).to(dev) @ inv
    return np.concatenate(S), T.cpu().numpy()


def check_modes(rsctx, torch, px, py, t, planes, tag):
    """FP64 vs oracle and FP32C vs FP64 (forward and adjoint) on one call with every
    plane; returns the FP64 and FP32C fields."""
    dev = torch.device("cuda:0")
    dpx = torch.from_numpy(px).to(dev)
    dpy = torch.from_numpy(py).to(dev)
    dt = torch.from_numpy(np.ascontiguousarray(t[None, :])).to(dev)
    E64 = rsctx.forward(dpx, dpy, dt, planes, mode="fp64", field=True)[0].cpu().numpy()
    E32 = rsctx.forward(dpx, dpy, dt, planes, mode="fp32c", field=True)[0].cpu().numpy()
    rng = np.random.default_rng(11)
    u = rng.standard_normal(E64.size) + 1j * rng.standard_normal(E64.size)
    du = torch.from_numpy(u).to(dev)
    g64 = rsctx.adjoint(dpx, dpy, du, planes, nsets=1, mode="fp64").cpu().numpy().ravel()
    g32 = rsctx.adjoint(dpx, dpy, du, planes, nsets=1, mode="fp32c").cpu().numpy().ravel()

    fx, fy, fz = pixel_coords(crs.resolve_offsets(planes))
    S, T = magnitude_sums(torch, px, py, np.abs(t), planes, np.abs(u))
    kR = max(P.k for P in planes) * float(np.max(np.sqrt(
        (np.max(np.abs(fx)) + np.max(np.abs(px))) ** 2
        + (np.max(np.abs(fy)) + np.max(np.abs(py))) ** 2 + np.max(fz ** 2))))
    tol64 = 1e-12 + 1e-15 * kR
    Eor = oracle.forward(px, py, t, planes)
    gor = np.ravel(oracle.adjoint(px, py, u, planes))
    e64 = float(np.max(np.abs(E64 - Eor) / S))
    a64 = float(np.max(np.abs(g64 - gor) / T))
    fe = float(np.max(np.abs(E32 - E64) / S))
    ge = float(np.max(np.abs(g32 - g64) / T))
    print(f"[lattices {tag}] fp64 vs oracle {e64:.1e} / {a64:.1e}; fp32c vs fp64 {fe:.2e} / "
          f"{ge:.2e} ({rsctx.last_launch(False)[0]}); field relL2 {rel_l2(E32, E64):.2e}",
          flush=True)
    assert e64 <= tol64 and a64 <= tol64
    assert fe <= BOUND and ge <= BOUND
    return E64, E32


def test_reversed_axes(torch_cuda, rsctx):
    px, py, t, k, lam, f = lens()
    P = window(k, 37, 2.5, f)
    R = reversed_window(P)
    E64, E32 = check_modes(rsctx, torch_cuda, px, py, t, [P, R], "reversed")
    n = P.nx * P.ny
    a64 = E64[:n].reshape(P.ny, P.nx)
    b64 = E64[n:].reshape(P.ny, P.nx)[::-1, ::-1]
    a32 = E32[:n].reshape(P.ny, P.nx)
    b32 = E32[n:].reshape(P.ny, P.nx)[::-1, ::-1]
    # The reversed lattice's coordinates are x0' + i sx' = x0 + (n-1-i) sx up to a rounding.
    assert np.max(np.abs(b64 - a64)) <= 1e-10 * np.max(np.abs(a64))
    # FP32C blocks the two orders differently (37 = 4*9 + 1: the ragged edge blocks land
    # on opposite sides); each is within the bound of FP64, so of each other.
    assert rel_l2(b32, a32) <= 2e-6


def test_tilted_planes(torch_cuda, rsctx):
    px, py, t, k, lam, f = lens()
    planes = [tilted(k, 33, 29, 0.08, f, a) for a in (15.0, 45.0, -60.0)]
    planes.append(tilted(k, 21, 45, 0.5 * lam, 0.5 * f, 80.0))  # nearly along z
    check_modes(rsctx, torch_cuda, px, py, t, planes, "tilted")


def test_zx_towards_lens(torch_cuda, rsctx):
    px, py, t, k, lam, f = lens()
    up = Plane.zx(-3.0, 3.0, 41, 0.6 * f, 1.4 * f, 35, k)
    down = Plane(up.k, up.x0, up.sx, up.y0, up.sy, up.z0 + (up.ny - 1) * up.sz, -up.sz,
                 up.nx, up.ny)
    E64, E32 = check_modes(rsctx, torch_cuda, px, py, t, [up, down], "zx down")
    n = up.nx * up.ny
    a = E64[:n].reshape(up.ny, up.nx)
    b = E64[n:].reshape(up.ny, up.nx)[::-1, :]
    assert np.max(np.abs(b - a)) <= 1e-10 * np.max(np.abs(a))


@pytest.mark.parametrize("kernel", [crs.KERNEL_REF, crs.KERNEL_RSI])
def test_mirror_through_lens_plane(torch_cuda, rsctx, kernel):
    px, py, t, k, lam, f = lens()
    P = replace(window(k, 33, 2.0, f), kernel=kernel)
    Z = replace(tilted(k, 25, 23, 0.1, 0.8 * f, 30.0), kernel=kernel)
    planes = [P, mirrored(P), Z, mirrored(Z)]
    E64, E32 = check_modes(rsctx, torch_cuda, px, py, t, planes, f"mirror kernel {kernel}")
    sign = 1.0 if kernel == crs.KERNEL_REF else -1.0
    o = 0
    for A in (P, Z):
        n = A.nx * A.ny
        front64, back64 = E64[o:o + n], E64[o + n:o + 2 * n]
        front32, back32 = E32[o:o + n], E32[o + n:o + 2 * n]
        assert np.max(np.abs(back64 - sign * front64)) <= 1e-12 * np.max(np.abs(front64))
        assert rel_l2(back32, sign * front32) <= 2e-6
        o += 2 * n


def test_thin_planes(torch_cuda, rsctx):
    """Planes thinner than a block in one call: a row, a column, a single pixel and a few
    small rectangles (the forward's edge classes and the adjoint's padded strips alone)."""
    px, py, t, k, lam, f = lens()
    step = 0.07
    planes = [
        Plane(k, -2.0, step, 0.4, 0.0, f, 0.0, 57, 1),          # one row
        Plane(k, 0.3, 0.0, -2.0, step, f, 0.0, 1, 57),          # one column
        Plane(k, 0.1, step, -0.1, step, f, 0.0, 1, 1),          # one pixel
        Plane(k, -0.2, step, -0.2, step, f, 0.0, 2, 2),
        Plane(k, -0.5, step, 0.2, step, 0.9 * f, 0.0, 3, 7),
        Plane(k, -1.0, step, -1.0, step, 1.1 * f, 0.0, 9, 5),
        Plane(k, -1.0, step, 0.0, 0.0, 0.8 * f, 0.5, 5, 9),    # zx, 5 wide
    ]
    check_modes(rsctx, torch_cuda, px, py, t, planes, "thin")
