"""The batched and large workloads at their real sizes, checked by properties that do
not need an FP64 sum over the whole workload: FP32C against the FP64 GPU kernel on
sampled rows (the FP64 kernel is pinned to the oracle in test_gpu_forward.py), graph
replay against eager, energy-free invariants.  Tolerance: 1e-5 relative L2 (FP32C).
"""
import math

import numpy as np
import pytest

import oracle
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
    on-axis focal isolation of the reference's polar report (cmd_polar.cpp:121-126) in
    FP32C and FP64, and the adjoint of a whole-grid cotangent against FP64 and the
    oracle on pillar samples."""
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
    # Focal isolation (cmd_polar.cpp:121-126): 10 log10(I_own(z) / max(I_other(z), 1e-30))
    # on axis at each channel's focal length; FP32C and FP64 agree within 0.01 dB.
    k = L.planes[0].k
    iso = {}
    for mode in ("fp32c", "fp64"):
        ax = [Plane.axis(f, f, 1, k, tset=s) for f in (250.0, 300.0) for s in (0, 1)]
        e_ax, _ = ctx.forward(px, py, t, ax, mode=mode)
        i = (e_ax.abs() ** 2).cpu().numpy()  # [x@f1, y@f1, x@f2, y@f2]
        iso[mode] = (10 * math.log10(i[0] / max(i[1], 1e-30)), 10 * math.log10(i[3] / max(i[2], 1e-30)))
        print(f"[config5] {mode} focal isolation X@f1 {iso[mode][0]:.3f} dB, Y@f2 {iso[mode][1]:.3f} dB")
    assert abs(iso["fp32c"][0] - iso["fp64"][0]) <= 0.01 and abs(iso["fp32c"][1] - iso["fp64"][1]) <= 0.01
    assert iso["fp64"][0] > 10 and iso["fp64"][1] > 10  # each channel focuses its own light
    # Adjoint of a whole-grid cotangent on both channels (2 x 1024^2 pixels): FP32C on all
    # pillars; FP64 (GPU) on a 4,096-pillar sample; the FP64 oracle on 32 of them.
    rng = np.random.default_rng(5)
    n_all = sum(p.npix for p in L.planes)
    u_np = rng.standard_normal(n_all) + 1j * rng.standard_normal(n_all)
    u = dev(torch, u_np)
    g32 = ctx.adjoint(px, py, u, L.planes, nsets=2, mode="fp32c").cpu().numpy()
    pick = np.arange(0, L.npil, L.npil // 4096)[:4096]
    spx, spy = dev(torch, L.px[pick]), dev(torch, L.py[pick])
    g64 = ctx.adjoint(spx, spy, u, L.planes, nsets=2, mode="fp64").cpu().numpy()
    e = rel_l2(g32[:, pick], g64)
    few = pick[::128]
    gref = oracle.adjoint(L.px[few], L.py[few], u_np, L.planes, nsets=2)
    e64 = rel_l2(g64[:, ::128], gref)
    print(f"[config5] full-grid adjoint relL2 vs fp64 on {pick.size} pillars {e:.3e}; "
          f"fp64 vs oracle on {few.size} {e64:.3e}")
    assert e <= 1e-5 and e64 <= 1e-12
    ctx.close()
