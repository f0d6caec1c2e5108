"""GPU parity of the opt-in Rayleigh-Sommerfeld I kernel, exp(ikR)/R (z/R)(1/R - ik), which
the reference does not have (it computes exp(ikR)/R only, propagate.hpp:4-5): the FP64
and FP32 kernels against the oracle's RS-I restatement (itself pinned by the on-axis
closed form, -d/dz of the reference kernel and the long-double brute force in
tests/test_oracle.py), forward and adjoint, alone and mixed with reference-kernel
planes in one launch.

Tolerances: FP64 <= 1e-12 relative L2 of the complex field / gradient; FP32C (RS-I
planes pair by pair at FP64 distances) <= 1e-5; FP32 (float R, k R, sincos, sums: the
reference GPU kernel's arithmetic) <= 2e-3.
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
