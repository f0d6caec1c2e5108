"""HIP-graph capture of the batched forward/adjoint, and the failure paths.

Graphs: each replay is checked against the eager result, with eager work of the
same context and torch ops between replays, and after the bound inputs change in
place.  Failures are forced from the host (CELERIS_RS_FAIL=alloc|launch) so every
kernel still runs normally; the reference's contract is that a failed call returns
false and leaves the caller's output untouched (propagate.cu:164-171, :184-192).
"""
import math
import os

import numpy as np
import pytest

import oracle
from celeris_amd import Context, Plane, configs
from celeris_amd import rs as crs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def batch():
    px, py = configs.square_grid(30, 0.35)
    rng = np.random.default_rng(21)
    t = rng.uniform(0.5, 1, (2, px.size)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (2, px.size)))
    planes = [Plane.psf(0, 0, 30.0, 2 * math.pi / 0.5, 17, 2.0, tset=0),
              Plane.psf(0, 0, 32.0, 2 * math.pi / 0.6, 9, 2.0, tset=1),
              Plane.zx(-2, 2, 11, 20, 40, 5, 2 * math.pi / 0.5, tset=0)]
    return px, py, t, planes


@pytest.mark.parametrize("mode", ["fp32c", "fp64"])
def test_graph_forward_replays_match_eager(torch_cuda, mode):
    torch = torch_cuda
    px, py, t, planes = batch()
    ctx = Context(0)
    dpx, dpy, dt = dev(torch, px), dev(torch, py), dev(torch, t)
    total = sum(p.npix for p in planes)
    field = torch.empty(total, dtype=torch.complex128, device="cuda")
    inten = torch.empty(total, dtype=torch.float64, device="cuda")
    print(f"[graph] capture forward {mode}", flush=True)
    g = crs.graph_forward(ctx, dpx, dpy, dt, planes, field, inten, mode=mode)
    for rep in range(3):
        ref_f, ref_i = ctx.forward(dpx, dpy, dt, planes, mode=mode, intensity=True)
        field.zero_()
        inten.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(field, ref_f), rep
        assert torch.equal(inten, ref_i), rep
        # eager work in between, then change the bound input in place
        _ = (dt * 2).sum().item()
        ctx.forward(dpx, dpy, dt * 0.5, planes[:1], mode=mode)
        dt.mul_(torch.exp(1j * torch.tensor(0.3, dtype=torch.float64, device="cuda")))
    g.close()
    ctx.close()


def test_graph_adjoint_replays_match_eager(torch_cuda):
    torch = torch_cuda
    px, py, t, planes = batch()
    ctx = Context(0)
    dpx, dpy = dev(torch, px), dev(torch, py)
    total = sum(p.npix for p in planes)
    rng = np.random.default_rng(8)
    du = dev(torch, rng.standard_normal(total) + 1j * rng.standard_normal(total))
    g_out = torch.empty((2, px.size), dtype=torch.complex128, device="cuda")
    print("[graph] capture adjoint", flush=True)
    gr = crs.graph_adjoint(ctx, dpx, dpy, du, planes, g_out, nsets=2)
    for rep in range(3):
        ref = ctx.adjoint(dpx, dpy, du, planes, nsets=2)
        g_out.zero_()
        gr.launch()
        torch.cuda.synchronize()
        assert torch.equal(g_out, ref), rep
        du.mul_(1.5 - 0.5j)
    gr.close()
    ctx.close()


def test_injected_failures(torch_cuda, monkeypatch):
    torch = torch_cuda
    px, py, t, planes = batch()
    dpx, dpy, dt = dev(torch, px), dev(torch, py), dev(torch, t)
    # Drop-in: false with out untouched.
    L = configs.small_lens(n_side=10, n=5)
    out = np.full((5, 5), 7.0)
    monkeypatch.setenv("CELERIS_RS_FAIL", "launch")
    import ctypes

    from celeris_amd import _lib

    lib = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)
    t0 = np.ascontiguousarray(L.t[0])
    ok = lib.crs_propagate_psf(L.px.ctypes.data_as(dp), L.py.ctypes.data_as(dp),
                               t0.ctypes.data_as(dp), L.npil, 0.0, 0.0, L.focal,
                               L.planes[0].k, 5, 1.0, out.ctypes.data_as(dp))
    assert ok == 0 and np.all(out == 7.0)
    # Batched: status codes, output untouched.
    ctx = Context(0)
    field = torch.full((sum(p.npix for p in planes),), 3.0 + 0j, dtype=torch.complex128,
                       device="cuda")
    with pytest.raises(RuntimeError, match="CRS_ERR_LAUNCH"):
        ctx.forward(dpx, dpy, dt, planes, out_field=field)
    monkeypatch.setenv("CELERIS_RS_FAIL", "alloc")
    fresh = Context(0)
    with pytest.raises(RuntimeError, match="CRS_ERR_ALLOC"):
        fresh.forward(dpx, dpy, dt, planes, out_field=field)
    torch.cuda.synchronize()
    assert torch.all(field == 3.0)
    monkeypatch.delenv("CELERIS_RS_FAIL")
    f, _ = fresh.forward(dpx, dpy, dt, planes)  # recovers once the fault is gone
    torch.cuda.synchronize()
    assert torch.isfinite(f.real).all()
    fresh.close()
    ctx.close()


def test_invalid_arguments(torch_cuda, rsctx):
    torch = torch_cuda
    px, py, t, planes = batch()
    dpx, dpy, dt = dev(torch, px), dev(torch, py), dev(torch, t)
    bad = [Plane.psf(0, 0, 30.0, 10.0, 9, 2.0, tset=5)]
    with pytest.raises(RuntimeError, match="CRS_ERR_INVALID"):
        rsctx.forward(dpx, dpy, dt, bad)
    with pytest.raises(RuntimeError, match="CRS_ERR_INVALID"):
        rsctx.forward(dpx, dpy, dt, [Plane(10.0, 0, 0.1, 0, 0.1, 5, 0, 0, 3)])
    # Two planes writing the same output elements: refused (one writer per element).
    overlap = [Plane.psf(0, 0, 30.0, 10.0, 9, 2.0, offset=0), Plane.psf(0, 0, 31.0, 10.0, 5, 2.0, offset=80)]
    with pytest.raises(RuntimeError, match="overlap"):
        rsctx.forward(dpx, dpy, dt, overlap)
    # Caller-provided buffers are checked before any launch writes through them.
    total = sum(p.npix for p in planes)
    with pytest.raises(ValueError, match="elements"):
        rsctx.forward(dpx, dpy, dt, planes,
                      out_field=torch.empty(total - 1, dtype=torch.complex128, device="cuda:0"))
    with pytest.raises(ValueError, match="float64"):
        rsctx.forward(dpx, dpy, dt, planes, intensity=True,
                      out_intensity=torch.empty(total, dtype=torch.float32, device="cuda:0"))
    u = torch.zeros(total, dtype=torch.complex128, device="cuda:0")
    with pytest.raises(ValueError, match="elements"):
        rsctx.adjoint(dpx, dpy, u, planes, nsets=2,
                      out=torch.empty((1, px.size), dtype=torch.complex128, device="cuda:0"))
    with pytest.raises(ValueError, match="complex128"):
        crs.graph_forward(rsctx, dpx, dpy, dt.to(torch.complex64), planes,
                          torch.empty(total, dtype=torch.complex128, device="cuda:0"))
    with pytest.raises(ValueError, match="elements"):
        crs.graph_adjoint(rsctx, dpx, dpy, u, planes,
                          torch.empty((1, px.size - 1), dtype=torch.complex128, device="cuda:0"))


def test_tuning_timing_and_launch_info(torch_cuda, tmp_path, monkeypatch):
    """crs_set_tuning changes the split (work-group count) but not the result beyond
    roundoff; crs_last_launch names the pair kernel; the autotuner installs one of its
    candidates and caches it by shape."""
    torch = torch_cuda
    _, _, _, planes = batch()
    px, py = configs.square_grid(64, 0.35)  # 4,096 pillars: up to 16 pillar chunks
    rng = np.random.default_rng(4)
    t = rng.uniform(0.5, 1, (2, px.size)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (2, px.size)))
    ctx = Context(0)
    dpx, dpy, dt = dev(torch, px), dev(torch, py), dev(torch, t)
    with pytest.raises(RuntimeError):
        ctx.last_launch(False)  # nothing launched yet
    outs, groups = [], []
    for target in (3, 4096):
        ctx.set_tuning("fp32c", target, target)
        f, _ = ctx.forward(dpx, dpy, dt, planes, mode="fp32c")
        outs.append(f.cpu().numpy())
        name, g = ctx.last_launch(False)
        assert name.startswith("crs::fwd_fp32c<") and g > 0  # 1 x 1 blocks here (coarse zx plane)
        groups.append(g)
    assert groups[1] > groups[0]
    assert np.linalg.norm(outs[0] - outs[1]) <= 1e-6 * np.linalg.norm(outs[1])
    ctx.set_timing(True)
    ctx.forward(dpx, dpy, dt, planes, mode="fp64")
    assert ctx.last_kernel_ms(False) > 0 and ctx.last_launch(False)[0] == "crs::fwd_fp64"
    ctx.set_timing(False)

    from celeris_amd import tune

    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    u = dev(torch, np.ones(sum(p.npix for p in planes), np.complex128))
    best = tune.autotune(ctx, dpx, dpy, dt, planes, u, mode="fp32c", targets=(512, 2048), reps=1)
    assert best["fwd"] in (512, 2048) and best["adj"] in (512, 2048)
    assert tuple(best["fwd_shape"]) in tune.FWD_SHAPES and tuple(best["adj_shape"]) in tune.ADJ_SHAPES
    # every candidate was timed, and names the kernel that ran
    assert len(best["fwd_ms"]) == 2 * len(tune.FWD_SHAPES) and len(best["adj_ms"]) == 2 * len(tune.ADJ_SHAPES)
    assert all(r["kernel"].startswith("crs::fwd_fp32c<") for r in best["fwd_ms"].values())
    again = tune.autotune(ctx, dpx, dpy, dt, planes, u, mode="fp32c", targets=(512, 2048), reps=1)
    assert again["fwd"] == best["fwd"] and again["adj"] == best["adj"]  # from the cache
    assert again["fwd_shape"] == best["fwd_shape"] and again["adj_shape"] == best["adj_shape"]
    # the installed choice gives the same field as the built-in one, to roundoff
    f_tuned, _ = ctx.forward(dpx, dpy, dt, planes, mode="fp32c")
    ctx.set_block_shape()
    ctx.set_tuning("fp32c", 0, 0)
    f_def, _ = ctx.forward(dpx, dpy, dt, planes, mode="fp32c")
    assert torch.linalg.vector_norm(f_tuned - f_def) <= 2e-6 * torch.linalg.vector_norm(f_def)
    ctx.close()


def test_one_context_alternating_streams(torch_cuda):
    """One context, forward and adjoint calls alternating between two streams with no
    host synchronisation between them, on shapes that change the plan and grow the
    scratch from call to call (the context orders each call after the previous one and
    frees scratch only once that work is done): every result against the oracle."""
    torch = torch_cuda
    L = configs.config1()
    k = L.planes[0].k
    ctx = Context(0)
    dpx, dpy, dt = dev(torch, L.px), dev(torch, L.py), dev(torch, L.t)
    shapes = [[Plane.psf(0.0, 0.0, 50.0, k, 21, 2.0)],
              [Plane.psf(0.0, 0.0, 50.0, k, 81, 5.0), Plane.psf(1.0, -1.0, 48.0, k, 41, 3.0)],
              [Plane.psf(0.5, 0.5, 52.0, k, 121, 6.0)]]
    rng = np.random.default_rng(21)
    us = [rng.standard_normal(sum(p.npix for p in pl)) + 1j * rng.standard_normal(sum(p.npix for p in pl))
          for pl in shapes]
    dus = [dev(torch, u) for u in us]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    out = []
    # two created streams, then the null stream (whose ordering event the context
    # records only when the next call on another stream needs it) against a created one
    for i in range(18):
        pl = shapes[i % 3]
        sa, sb = (s1, s2) if i < 9 else (None, s1)
        if i % 2 == 0:
            f, _ = ctx.forward(dpx, dpy, dt, pl, mode="fp32c", stream=sa)
            out.append(("fwd", i % 3, f))
        else:
            g = ctx.adjoint(dpx, dpy, dus[i % 3], pl, nsets=1, mode="fp32c", stream=sb)
            out.append(("adj", i % 3, g))
    torch.cuda.synchronize()
    for kind, si, res in out:
        pl = shapes[si]
        if kind == "fwd":
            ref = oracle.forward(L.px, L.py, L.t, pl)
        else:
            ref = oracle.adjoint(L.px, L.py, us[si], pl)
        e = rel_l2(res.cpu().numpy(), ref)
        print(f"{kind} shape {si}: relL2 {e:.2e}")
        assert e <= 1e-5
    ctx.close()


def test_one_context_shared_by_threads(torch_cuda):
    """Four host threads share one Python Context (its lock serialises the library
    calls), each running forwards and adjoints of its own plane list on its own stream;
    every result equals, bit for bit, the same call made alone on one thread (the same
    plan, the same cut, the same fixed-order sums)."""
    import threading

    torch = torch_cuda
    L = configs.config1()
    k = L.planes[0].k
    dpx, dpy, dt = dev(torch, L.px), dev(torch, L.py), dev(torch, L.t)
    shapes = [[Plane.psf(0.0, 0.0, 50.0, k, 33, 2.0)],
              [Plane.psf(0.2, -0.1, 49.0, k, 57, 4.0)],
              [Plane.psf(-0.5, 0.5, 51.0, k, 81, 5.0), Plane.psf(1.0, -1.0, 48.0, k, 41, 3.0)],
              [Plane.zx(-3.0, 3.0, 45, 40.0, 60.0, 37, k)]]
    rng = np.random.default_rng(33)
    dus = [dev(torch, rng.standard_normal(sum(p.npix for p in pl))
               + 1j * rng.standard_normal(sum(p.npix for p in pl))) for pl in shapes]
    solo = Context(0)
    want = [(solo.forward(dpx, dpy, dt, pl, mode="fp32c")[0].cpu(),
             solo.adjoint(dpx, dpy, du, pl, nsets=1, mode="fp32c").cpu())
            for pl, du in zip(shapes, dus)]
    solo.close()
    ctx = Context(0)
    got = [[] for _ in shapes]
    errors = []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(6):
                    f, _ = ctx.forward(dpx, dpy, dt, shapes[i], mode="fp32c")
                    g = ctx.adjoint(dpx, dpy, dus[i], shapes[i], nsets=1, mode="fp32c")
                    got[i].append((f, g))
            s.synchronize()
        except Exception as e:  # reported on the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(shapes))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for i, runs in enumerate(got):
        assert len(runs) == 6
        for f, g in runs:
            assert torch.equal(f.cpu(), want[i][0])
            assert torch.equal(g.cpu(), want[i][1])
    ctx.close()
