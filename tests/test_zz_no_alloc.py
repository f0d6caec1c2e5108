"""crs_forward_noalloc — the forward with no device allocation (scratch carved from a
static device array) — against crs_forward, the entry point it is a variant of, and the
oracle: two calls in a row, the second smaller; a call whose partial sums must be cut to
fit the array; a call too large for it (refused with CRS_ERR_INVALID, outputs untouched)."""
import ctypes

import numpy as np
import pytest

import oracle
from celeris_amd import Context, Plane, configs
from celeris_amd import _lib
from celeris_amd.rs import pack_planes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def noalloc(torch, px, py, t, planes, mode):
    npil = px.numel()
    arr, total = pack_planes(planes)
    f = torch.full((total,), complex("nan"), dtype=torch.complex128, device="cuda")
    it = torch.full((total,), float("nan"), dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = _lib.load().crs_forward_noalloc(
        stream, {"fp32c": 0, "fp64": 1, "fp32": 2}[mode], ctypes.c_void_p(px.data_ptr()),
        ctypes.c_void_p(py.data_ptr()), ctypes.c_void_p(t.data_ptr()), npil, t.numel() // npil,
        arr, len(planes), ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(it.data_ptr()))
    return st, f, it


@pytest.mark.parametrize("mode", ["fp32c", "fp64"])
def test_noalloc_matches_forward_two_calls(torch_cuda, mode):
    """Two calls in a row on the same device (the second smaller: fewer pillars, pixels
    and planes), each against crs_forward with the same built-in plan (bit for bit) and
    the oracle."""
    torch = torch_cuda
    ctx = Context(0)
    L = configs.config1()
    k = L.planes[0].k
    cases = [(L.px, L.py, L.t, [L.planes[0], Plane.psf(1.0, -0.5, 48.0, k, 37, 3.0)]),
             (L.px[:1500], L.py[:1500], L.t[:, :1500], [Plane.psf(0.2, 0.1, 50.0, k, 23, 2.0)])]
    for i, (px, py, t, planes) in enumerate(cases):
        dpx, dpy, dt = dev(torch, px), dev(torch, py), dev(torch, t)
        print(f"[noalloc] {mode} call {i}: {px.size} pillars, {len(planes)} planes", flush=True)
        st, f, it = noalloc(torch, dpx, dpy, dt, planes, mode)
        assert st == 0, _lib.load().crs_last_error()
        ref_f, ref_i = ctx.forward(dpx, dpy, dt, planes, mode=mode, intensity=True)
        torch.cuda.synchronize()
        assert torch.equal(f, ref_f) and torch.equal(it, ref_i)
        orc = oracle.forward(px, py, t, planes)
        e = rel_l2(f.cpu().numpy(), orc)
        print(f"[noalloc] {mode} call {i}: vs oracle {e:.2e}")
        assert e <= (1e-5 if mode == "fp32c" else 1e-12)
    ctx.close()


def test_noalloc_headline_cut_to_fit_and_refusal(torch_cuda):
    """The headline workload (91,809 pillars x 161^2): its built-in forward plan needs
    more partial-sum space than the array holds, so the allocation-free call cuts the
    work into fewer pieces; the field still meets crs_forward's to 1e-6 and the FP64
    kernel's to 1e-5.  Then a call of 3,000,000 pillars, whose staged pillars alone
    exceed the array, is refused with CRS_ERR_INVALID and leaves the outputs untouched."""
    torch = torch_cuda
    ctx = Context(0)
    L = configs.config2(n=161)
    dpx, dpy, dt = dev(torch, L.px), dev(torch, L.py), dev(torch, L.t)
    st, f, _ = noalloc(torch, dpx, dpy, dt, L.planes, "fp32c")
    assert st == 0, _lib.load().crs_last_error()
    ref, _ = ctx.forward(dpx, dpy, dt, L.planes, mode="fp32c")
    f64, _ = ctx.forward(dpx, dpy, dt, L.planes, mode="fp64")
    torch.cuda.synchronize()
    e1 = float(torch.linalg.vector_norm(f - ref) / torch.linalg.vector_norm(ref))
    e2 = float(torch.linalg.vector_norm(f - f64) / torch.linalg.vector_norm(f64))
    print(f"[noalloc] headline: vs crs_forward {e1:.2e}, vs fp64 {e2:.2e}")
    assert e1 <= 1e-6 and e2 <= 1e-5
    n = 3_000_000
    big = torch.zeros(n, dtype=torch.float64, device="cuda")
    tb = torch.ones(n, dtype=torch.complex128, device="cuda")
    st, f, it = noalloc(torch, big, big, tb, [Plane.psf(0.0, 0.0, 50.0, 10.0, 5, 1.0)], "fp32c")
    torch.cuda.synchronize()
    assert st == 2 and "allocation-free maximum" in _lib.load().crs_last_error().decode()
    assert torch.isnan(f.real).all() and torch.isnan(it).all()
    ctx.close()
