"""FP32C against FP64 over random geometries: forward and adjoint, per output element.

The FP32-compensated kernels choose, per work item, among series orders of the
expansion of R about a pixel block's centre (the cheapest whose worst-case pair error,
truncation plus float rounding plus the hardware sin/cos, stays within 8e-6 of |t|/R),
fall back to 4 x 4 blocks where the tuned shape is not admitted, and to every pair at
its own FP64 distance where no block is.  The fixed workloads exercise a few of these
choices; this sweep draws wavelength, pillar pitch and count, transmission (a focusing
lens or random complex values), plane kind (psf windows anywhere in front of the lens,
zx slices from near the lens plane outwards), grid sizes (ragged against the blocks),
window sizes and offsets from a seeded generator, so the choices and their boundaries
are met in many combinations.  Seeds 40-59 draw coarse pixel pitches (1 to 10
wavelengths) and far planes (up to 3 mm), where the blocks span several turns and the
float rounding of the block offsets is largest.

Bound, per output element: every pair's error is held to 8e-6 of |t_q|/R_pq by the
admission bound, and the float sums (at most 1,024 terms per wave before the FP64 fold)
add ~1e-6 of the summed magnitudes for a coherent pixel.  So
    |E_fp32c - E_fp64|_p  <=  1e-5 * sum_q |t_q| / R_pq
    |g_fp32c - g_fp64|_q  <=  1e-5 * sum_p |u_p| / R_pq
with the right-hand sides computed here in FP64 from the geometry; the FP64 kernels are
held to the oracle on the same inputs, per element, at (1e-12 + 1e-15 k R_max) of the
same scale (the FP64 rounding of k R).  (Relative L2 is
printed, not asserted: a window in a dark region, where the terms cancel, can have any
ratio of error to field.)
"""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import Plane
from celeris_amd import rs as crs
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SEEDS = list(range(60))
BOUND = 1e-5


def geometry(seed):
    rng = np.random.default_rng(1000 + seed)
    lam = rng.uniform(0.4, 1.0)
    k = 2.0 * math.pi / lam
    p = rng.uniform(0.2, 0.6)
    n_side = int(rng.integers(6, 72))
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    if rng.random() < 0.5:  # circular aperture
        keep = np.hypot(px, py) <= (n_side * p) / 2.0
        px, py = px[keep], py[keep]
    half = n_side * p / 2.0
    kind = "focus" if rng.random() < 0.5 else "random"
    if kind == "focus":
        f = float(np.exp(rng.uniform(np.log(2.0), np.log(300.0))))
        r = np.hypot(px, py)
        t = rng.uniform(0.5, 1.0, px.size) * np.exp(-1j * k * (np.sqrt(r * r + f * f) - f))
    else:
        f = None
        t = rng.standard_normal(px.size) + 1j * rng.standard_normal(px.size)
    # Pixel pitches from 0.02 to 1 wavelength (focal sampling is usually well below one),
    # so that first-order blocks, second-order blocks and exact pairs all occur; the
    # coarse seeds take 1 to 10 wavelengths and planes up to 3 mm away.
    logu = lambda lo, hi: float(np.exp(rng.uniform(np.log(lo), np.log(hi))))  # noqa: E731
    coarse = seed >= 40
    pitch = (1.0, 10.0) if coarse else (0.02, 1.0)
    zrange = (100.0, 3000.0) if coarse else (0.5, 300.0)
    planes = []
    for _ in range(int(rng.integers(1, 4))):
        if rng.random() < 0.7:
            z = f if (f is not None and rng.random() < 0.5 and not coarse) else logu(*zrange)
            n = int(rng.integers(2, 98))
            W = 0.5 * (n - 1) * lam * logu(*pitch)
            cx, cy = rng.uniform(-1.5 * half, 1.5 * half, 2)
            planes.append(Plane.psf(float(cx), float(cy), z, k, n, W))
        else:
            nx, nz = int(rng.integers(2, 98)), int(rng.integers(2, 98))
            z1 = logu(zrange[0] * 0.6, zrange[1] * 0.5)
            z2 = z1 + (nz - 1) * lam * logu(*pitch)
            xw = 0.5 * (nx - 1) * lam * logu(*pitch)
            x0 = float(rng.uniform(-half, half))
            planes.append(Plane.zx(x0 - xw, x0 + xw, nx, z1, z2, nz, k))
    return dict(lam=lam, p=p, px=px, py=py, t=t, kind=kind, f=f, planes=planes)


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


def magnitude_sums(torch, px, py, at, fx, fy, fz, au):
    """S_p = sum_q at_q / R_pq and T_q = sum_p au_p / R_pq, in FP64 on the GPU."""
    dev = torch.device("cuda:0")
    qx = torch.from_numpy(px).to(dev)
    qy = torch.from_numpy(py).to(dev)
    ta = torch.from_numpy(at).to(dev)
    S = np.empty(fx.size)
    T = torch.zeros(px.size, dtype=torch.float64, device=dev)
    step = max(1, (1 << 24) // max(1, px.size))
    for s in range(0, fx.size, step):
        e = min(fx.size, s + step)
        dx = torch.from_numpy(fx[s:e]).to(dev)[:, None] - qx[None, :]
        dy = torch.from_numpy(fy[s:e]).to(dev)[:, None] - qy[None, :]
        z = torch.from_numpy(fz[s:e]).to(dev)[:, None]
        inv = torch.rsqrt(dx * dx + dy * dy + z * z)
        S[s:e] = (inv @ ta).cpu().numpy()
        T += torch.from_numpy(au[s:e]).to(dev) @ inv
    return S, T.cpu().numpy()


@pytest.mark.parametrize("seed", SEEDS)
def test_fp32c_vs_fp64_random_geometry(torch_cuda, rsctx, seed):
    torch = torch_cuda
    g = geometry(seed)
    px, py, t, planes = g["px"], g["py"], g["t"], g["planes"]
    desc = ", ".join(f"{'zx' if P.sz else 'psf'} {P.nx}x{P.ny} z0={P.z0:.3g}"
                     for P in planes)
    print(f"[random geometry {seed}] lambda={g['lam']:.3f} p={g['p']:.3f} npil={px.size} "
          f"{g['kind']} planes: {desc}", flush=True)
    dev = torch.device("cuda:0")
    dpx = torch.from_numpy(px).to(dev)
    dpy = torch.from_numpy(py).to(dev)
    dt = torch.from_numpy(np.ascontiguousarray(t[None, :])).to(dev)

    E64, _ = rsctx.forward(dpx, dpy, dt, planes, mode="fp64", field=True)
    E32, _ = rsctx.forward(dpx, dpy, dt, planes, mode="fp32c", field=True)
    kernel = rsctx.last_launch(False)[0]
    E64 = E64.cpu().numpy()
    E32 = E32.cpu().numpy()
    rng = np.random.default_rng(7 + seed)
    u = rng.standard_normal(E64.size) + 1j * rng.standard_normal(E64.size)
    du = torch.from_numpy(u).to(dev)
    g64 = rsctx.adjoint(dpx, dpy, du, planes, nsets=1, mode="fp64").cpu().numpy()
    g32 = rsctx.adjoint(dpx, dpy, du, planes, nsets=1, mode="fp32c").cpu().numpy()

    fx, fy, fz = pixel_coords(crs.resolve_offsets(planes))
    S, T = magnitude_sums(torch, px, py, np.abs(t), fx, fy, fz, np.abs(u))
    # The FP64 kernels against the oracle on the same inputs, to FP64 roundoff: k R is
    # rounded once in each (relative 2^-53, so ~1.1e-16 k R of phase per pair, and the two
    # sincos implementations differ by an ulp of it), hence 1e-12 + 1e-15 k R_max.
    kR = max(P.k for P in planes) * float(np.max(np.sqrt(
        (np.max(np.abs(fx)) + np.max(np.abs(px))) ** 2
        + (np.max(np.abs(fy)) + np.max(np.abs(py))) ** 2 + np.max(fz ** 2))))
    Eor = oracle.forward(px, py, t, planes)
    gor = oracle.adjoint(px, py, u, planes)
    tol64 = 1e-12 + 1e-15 * kR
    assert float(np.max(np.abs(E64 - Eor) / S)) <= tol64
    assert float(np.max(np.abs(g64.ravel() - np.ravel(gor)) / T)) <= tol64
    fe = float(np.max(np.abs(E32 - E64) / S))
    ge = float(np.max(np.abs(g32 - g64) / T))
    print(f"[random geometry {seed}] {kernel}: forward max |dE|/sum|t|/R {fe:.2e}, "
          f"adjoint max |dg|/sum|u|/R {ge:.2e}", flush=True)
    assert fe <= BOUND
    assert ge <= BOUND
    print(f"[random geometry {seed}] field relL2 {rel_l2(E32, E64):.2e}, gradient relL2 "
          f"{rel_l2(g32, g64):.2e}")
    # Plain FP32 (the reference kernel's arithmetic: k R formed and reduced in float):
    # each pair's phase is off by a few float roundings of k R (~2^-24 k R each) and its
    # 1/R by ~1e-7, so |dE_p| <= (4e-7 k R_max + 1e-6) sum_q |t_q| / R_pq.
    Ef = rsctx.forward(dpx, dpy, dt, planes, mode="fp32", field=True)[0].cpu().numpy()
    e32 = float(np.max(np.abs(Ef - E64) / S))
    print(f"[random geometry {seed}] plain FP32 (k R_max {kR:.0f}): max {e32:.2e}")
    assert e32 <= 4e-7 * kR + 1e-6


def batched_geometry(seed):
    """Several transmission sets and one wavelength per plane, as the batched analyses
    use them (chromatic scans, tolerance trials, field points)."""
    rng = np.random.default_rng(5000 + seed)
    p = rng.uniform(0.25, 0.5)
    n_side = int(rng.integers(8, 48))
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * p
    X, Y = np.meshgrid(c, c)
    px, py = X.ravel(), Y.ravel()
    nsets = int(rng.integers(1, 4))
    t = rng.uniform(0.3, 1.0, (nsets, px.size)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (nsets, px.size)))
    half = n_side * p / 2.0
    planes = []
    for _ in range(int(rng.integers(2, 6))):
        lam = rng.uniform(0.4, 0.9)
        k = 2.0 * math.pi / lam
        tset = int(rng.integers(0, nsets))
        n = int(rng.integers(2, 60))
        z = float(np.exp(rng.uniform(np.log(1.0), np.log(200.0))))
        W = 0.5 * (n - 1) * lam * float(np.exp(rng.uniform(np.log(0.03), np.log(0.8))))
        cx, cy = rng.uniform(-half, half, 2)
        if rng.random() < 0.8:
            planes.append(Plane.psf(float(cx), float(cy), z, k, n, W, tset=tset))
        else:
            planes.append(Plane.zx(float(cx) - W, float(cx) + W, n, z, z + (n - 1) * 0.3 * lam, n,
                                   k, tset=tset))
    return px, py, t, planes


@pytest.mark.parametrize("seed", list(range(16)))
def test_fp32c_vs_fp64_random_batches(torch_cuda, rsctx, seed):
    """Batched calls: per-plane wavelength and transmission set; forward per element
    against FP64 with the bound above (each plane against its own set's |t|), adjoint per
    set and pillar against sum_p |u_p| / R_pq over that set's planes."""
    torch = torch_cuda
    px, py, t, planes = batched_geometry(seed)
    nsets = t.shape[0]
    print(f"[random batch {seed}] npil={px.size} sets={nsets} planes="
          + ", ".join(f"{'zx' if P.sz else 'psf'} {P.nx}x{P.ny} set {P.tset}" for P in planes),
          flush=True)
    dev = torch.device("cuda:0")
    dpx = torch.from_numpy(px).to(dev)
    dpy = torch.from_numpy(py).to(dev)
    dt = torch.from_numpy(np.ascontiguousarray(t)).to(dev)
    E64 = rsctx.forward(dpx, dpy, dt, planes, mode="fp64")[0].cpu().numpy()
    E32 = rsctx.forward(dpx, dpy, dt, planes, mode="fp32c")[0].cpu().numpy()
    rng = np.random.default_rng(99 + seed)
    u = rng.standard_normal(E64.size) + 1j * rng.standard_normal(E64.size)
    du = torch.from_numpy(u).to(dev)
    g64 = rsctx.adjoint(dpx, dpy, du, planes, nsets=nsets, mode="fp64").cpu().numpy()
    g32 = rsctx.adjoint(dpx, dpy, du, planes, nsets=nsets, mode="fp32c").cpu().numpy()
    Eor = oracle.forward(px, py, t, planes)
    gor = oracle.adjoint(px, py, u, planes, nsets=nsets)
    fe, ge = 0.0, 0.0
    T = np.zeros((nsets, px.size))
    for P in crs.resolve_offsets(planes):
        sl = slice(P.offset, P.offset + P.npix)
        fx, fy, fz = pixel_coords([P])
        S, Tp = magnitude_sums(torch, px, py, np.abs(t[P.tset]), fx, fy, fz, np.abs(u[sl]))
        assert float(np.max(np.abs(E64[sl] - Eor[sl]) / S)) <= 1e-12
        fe = max(fe, float(np.max(np.abs(E32[sl] - E64[sl]) / S)))
        T[P.tset] += Tp
    used = sorted({P.tset for P in planes})
    assert float(np.max(np.abs(g64[used] - gor[used]) / T[used])) <= 1e-12
    ge = float(np.max(np.abs(g32[used] - g64[used]) / T[used]))
    print(f"[random batch {seed}] {rsctx.last_launch(False)[0]}: forward max {fe:.2e}, "
          f"adjoint max {ge:.2e}", flush=True)
    assert fe <= BOUND
    assert ge <= BOUND
    unused = [s for s in range(nsets) if s not in used]
    assert np.all(g32[unused] == 0) and np.all(g64[unused] == 0)
