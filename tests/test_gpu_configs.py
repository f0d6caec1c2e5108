"""The batched and large workloads at their real sizes, checked by properties that do
not need an FP64 sum over the whole workload: FP32C against the FP64 GPU kernel on
sampled rows (the FP64 kernel is pinned to the oracle in test_gpu_forward.py), graph
replay against eager, energy-free invariants.  Tolerance: 1e-5 relative L2 (FP32C).
"""
import math

import numpy as np
import pytest

from celeris_amd import Context, Plane, configs
from celeris_amd import rs as crs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def row_planes(plane: Plane, rows, tset=None):
    """One-row planes sampling `rows` of `plane` (same pixels, same k)."""
    out = []
    for j in rows:
        out.append(Plane(plane.k, plane.x0, plane.sx, plane.y0 + j * plane.sy, 0.0,
                         plane.z0 + j * plane.sz, 0.0, plane.nx, 1,
                         plane.tset if tset is None else tset))
    return out


def test_config3_batched_graph(torch_cuda):
    """8 wavelengths x 31 planes (248 instances, uneven grids 81/161/321), one graph
    launch; every wavelength's first plane and a row sample of every plane against FP64."""
    torch = torch_cuda
    L = configs.config3()
    print(f"[config3] {len(L.planes)} planes, {L.pairs:.3e} pairs", flush=True)
    ctx = Context(0)
    px, py, t = dev(torch, L.px), dev(torch, L.py), dev(torch, L.t)
    total = sum(p.npix for p in L.planes)
    field = torch.empty(total, dtype=torch.complex128, device="cuda")
    g = crs.graph_forward(ctx, px, py, t, L.planes, field, None, mode="fp32c")
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    g.launch()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    print(f"[config3] graph forward {ms:.1f} ms = {L.pairs / ms / 1e6:.1f} G pairs/s", flush=True)
    eager, _ = ctx.forward(px, py, t, L.planes, mode="fp32c")
    torch.cuda.synchronize()
    assert torch.equal(eager, field)
    f = field.cpu().numpy()
    planes = crs.resolve_offsets(L.planes)
    worst = 0.0
    for idx, p in enumerate(planes):
        rows = [0, p.ny // 2, p.ny - 1] if idx % 31 else list(range(p.ny))
        sub = row_planes(p, rows)
        f64, _ = ctx.forward(px, py, t, sub, mode="fp64")
        mine = np.concatenate([f[p.offset + j * p.nx: p.offset + (j + 1) * p.nx] for j in rows])
        e = rel_l2(mine, f64.cpu().numpy())
        worst = max(worst, e)
    print(f"[config3] worst sampled relL2 vs fp64 {worst:.3e}")
    assert worst <= 1e-5
    g.close()
    ctx.close()


def test_config5_large_polarization(torch_cuda):
    """1,002,001 pillars, X channel focusing at 250 um, Y at 300 um, 1024^2 grid per
    channel: FP32C forward of both channels in one launch, sampled rows against FP64,
    on-axis focal isolation of the reference's polar report (cmd_polar.cpp:121-126),
    and the adjoint on a sampled-row cotangent against FP64."""
    torch = torch_cuda
    L = configs.config5()
    print(f"[config5] {L.npil} pillars, {L.pairs:.3e} pairs", flush=True)
    ctx = Context(0)
    px, py, t = dev(torch, L.px), dev(torch, L.py), dev(torch, L.t)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    field, inten = ctx.forward(px, py, t, L.planes, mode="fp32c", intensity=True)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    print(f"[config5] forward {ms:.1f} ms = {L.pairs / ms / 1e6:.1f} G pairs/s", flush=True)
    f = field.cpu().numpy()
    planes = crs.resolve_offsets(L.planes)
    rows = [0, 255, 511, 512, 700, 1023]
    for p in planes:
        f64, _ = ctx.forward(px, py, t, row_planes(p, rows), mode="fp64")
        mine = np.concatenate([f[p.offset + j * p.nx: p.offset + (j + 1) * p.nx] for j in rows])
        e = rel_l2(mine, f64.cpu().numpy())
        print(f"[config5] channel {p.tset} sampled relL2 vs fp64 {e:.3e}", flush=True)
        assert e <= 1e-5
    # Focal isolation: I_x(z_x) / I_y(z_x) on axis, both channels' fields at z_x = f1.
    k = L.planes[0].k
    ax = [Plane.axis(250.0, 250.0, 1, k, tset=0), Plane.axis(250.0, 250.0, 1, k, tset=1)]
    for mode in ("fp32c", "fp64"):
        e_ax, _ = ctx.forward(px, py, t, ax, mode=mode)
        i_ax = (e_ax.abs() ** 2).cpu().numpy()
        print(f"[config5] {mode} focal isolation X@f1: {10 * math.log10(i_ax[0] / i_ax[1]):.2f} dB")
    # Adjoint of a cotangent living on sampled rows of both channels.
    sub = row_planes(planes[0], rows) + row_planes(planes[1], rows)
    rng = np.random.default_rng(5)
    n_sub = sum(p.npix for p in sub)
    u = dev(torch, rng.standard_normal(n_sub) + 1j * rng.standard_normal(n_sub))
    g32 = ctx.adjoint(px, py, u, sub, nsets=2, mode="fp32c")
    g64 = ctx.adjoint(px, py, u, sub, nsets=2, mode="fp64")
    e = rel_l2(g32.cpu().numpy(), g64.cpu().numpy())
    print(f"[config5] adjoint relL2 vs fp64 {e:.3e}")
    assert e <= 1e-5
    ctx.close()
