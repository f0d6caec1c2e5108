"""GPU parity of the forward sum: HIP kernels (through the C ABI) vs the FP64 oracle.

Tolerances (written per test): FP64 mode <= 1e-12 relative L2 of the complex field;
FP32C mode <= 1e-5 relative L2 (the contract for the compensated mode, on the
complex field with the exp(ikz) factor restored and on |E|^2); plain FP32 mode,
which is the reference GPU kernel's arithmetic, <= 1e-3 on |E|^2.
"""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import Plane, configs
from celeris_amd import rs as crs
from conftest import max_rel, rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_forward(torch, ctx, px, py, t, planes, mode):
    print(f"[gpu_forward] mode={mode} npil={np.size(px)} planes={len(planes)}", flush=True)
    f, inten = ctx.forward(dev(torch, px), dev(torch, py), dev(torch, np.atleast_2d(t)), planes,
                           mode=mode, field=True, intensity=True)
    torch.cuda.synchronize()
    return f.cpu().numpy(), inten.cpu().numpy()


SMALL = [
    dict(n_side=20, n=17),   # 400 pillars, 17x17 (ragged vs the 4x4 thread tile)
    dict(n_side=23, n=9),    # 529 pillars (not a multiple of the LDS tile)
    dict(n_side=7, n=33),    # 49 pillars, 33x33
]


@pytest.mark.parametrize("case", SMALL)
@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_small_field_vs_oracle(torch_cuda, rsctx, case, mode, tol):
    L = configs.small_lens(n_side=case["n_side"], n=case["n"])
    ref = oracle.forward(L.px, L.py, L.t, L.planes)
    f, inten = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, mode)
    e = rel_l2(f, ref)
    ei = rel_l2(inten, np.abs(ref) ** 2)
    print(f"{mode} {case}: field relL2 {e:.3e} intensity relL2 {ei:.3e} maxrel {max_rel(f, ref):.3e}")
    assert e <= tol
    assert ei <= 2 * tol


def test_small_fp32_plain_vs_oracle(torch_cuda, rsctx):
    L = configs.small_lens(n_side=20, n=17)
    ref = oracle.forward(L.px, L.py, L.t, L.planes)
    _, inten = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp32")
    e = rel_l2(inten, np.abs(ref) ** 2)
    print(f"fp32 plain intensity relL2 {e:.3e}")
    assert e <= 1e-3


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_config1_vs_oracle(torch_cuda, rsctx, mode, tol):
    L = configs.config1()
    ref = oracle.forward(L.px, L.py, L.t, L.planes)
    f, inten = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, mode)
    e, ei = rel_l2(f, ref), rel_l2(inten, np.abs(ref) ** 2)
    print(f"config1 {mode}: field relL2 {e:.3e} I relL2 {ei:.3e} maxrel {max_rel(f, ref):.3e}")
    assert e <= tol and ei <= 2 * tol


def test_single_pillar_closed_form(torch_cuda, rsctx):
    """One pillar: E = a exp(i(phi + kR)) / R exactly, so |E|^2 = a^2 / R^2."""
    a, phi, x, y = 0.83, 1.234, 3.7, -2.2
    k = 2 * math.pi / 0.6
    plane = Plane.psf(0.5, -0.25, 40.0, k, 21, 3.0)
    i = np.arange(21)
    fx = plane.x0 + i * plane.sx
    fy = plane.y0 + i * plane.sy
    R = np.sqrt((fx[None, :] - x) ** 2 + (fy[:, None] - y) ** 2 + 40.0 ** 2)
    exact = (a * np.exp(1j * (phi + k * R)) / R).ravel()
    t = np.array([[a * np.exp(1j * phi)]])
    for mode, tol in (("fp64", 1e-13), ("fp32c", 2e-6)):
        f, inten = gpu_forward(torch_cuda, rsctx, np.array([x]), np.array([y]), t, [plane], mode)
        print(mode, rel_l2(f, exact), rel_l2(inten, np.abs(exact) ** 2))
        assert rel_l2(f, exact) <= tol
        assert rel_l2(inten, np.abs(exact) ** 2) <= tol


def test_inphase_focus_identity(torch_cuda, rsctx):
    """Hyperbolic phase: at (0, 0, f) every pillar arrives in phase, E = e^{ikf} sum a/R
    (the closed form behind analyze_focus's peak_same_amp, focal.cpp:108-115)."""
    L = configs.config1()
    a = np.abs(L.t[0])
    k = 2 * math.pi / L.wavelengths[0]
    R = np.sqrt(L.px ** 2 + L.py ** 2 + L.focal ** 2)
    exact = np.exp(1j * k * L.focal) * np.sum(a / R)
    plane = Plane.psf(0.0, 0.0, L.focal, k, 3, 1.0)  # centre pixel is (0, 0)
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-6)):
        f, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, [plane], mode)
        err = abs(f[4] - exact) / abs(exact)
        print(mode, "focus identity rel err", err)
        assert err <= tol


def test_batched_uneven_planes(torch_cuda, rsctx):
    """Several planes, wavelengths, transmission sets and uneven grids in one launch,
    with a gap in the output layout; each plane must equal its own oracle sum."""
    px, py = configs.square_grid(18, 0.35)
    rng = np.random.default_rng(7)
    t = rng.uniform(0.5, 1, (3, px.size)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (3, px.size)))
    planes = [
        Plane.psf(0.0, 0.0, 30.0, 2 * math.pi / 0.45, 13, 2.0, tset=0, offset=0),
        Plane.psf(0.3, -0.2, 28.5, 2 * math.pi / 0.55, 7, 1.5, tset=2, offset=13 * 13 + 5),
        Plane.zx(-2.0, 2.0, 11, 20.0, 40.0, 6, 2 * math.pi / 0.65, tset=1),
        Plane.axis(10.0, 50.0, 9, 2 * math.pi / 0.5, tset=1),
        Plane(2 * math.pi / 0.6, -1.0, 0.13, 0.7, 0.0, 25.0, 0.0, 5, 1, 0, -1),
    ]
    ref = oracle.forward(px, py, t, planes)
    total = ref.size
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-5)):
        f, _ = gpu_forward(torch_cuda, rsctx, px, py, t, planes, mode)
        assert f.size == total
        for p in crs.resolve_offsets(planes):
            sl = slice(p.offset, p.offset + p.npix)
            e = rel_l2(f[sl], ref[sl])
            print(mode, p.nx, p.ny, p.tset, e)
            assert e <= tol


def test_deterministic(torch_cuda, rsctx):
    L = configs.small_lens(n_side=30, n=21)
    a, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp32c")
    b, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp32c")
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_linearity(torch_cuda, rsctx):
    """E(a t1 + b t2) = a E(t1) + b E(t2) at the config-2 pillar count (91,809)."""
    L = configs.config2(n=33)
    rng = np.random.default_rng(3)
    t1 = L.t
    t2 = (rng.uniform(0.5, 1.0, L.npil) * np.exp(1j * rng.uniform(-3, 3, L.npil)))[None]
    a, b = 0.6 - 0.2j, -0.3 + 0.9j
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-5)):
        e1, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, t1, L.planes, mode)
        e2, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, t2, L.planes, mode)
        e12, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, a * t1 + b * t2, L.planes, mode)
        err = rel_l2(e12, a * e1 + b * e2)
        print(mode, "linearity", err)
        assert err <= tol


def test_config2_headline_rows_vs_oracle(torch_cuda, rsctx):
    """Headline shape (91,809 pillars, 161^2): FP32C and FP64 against the FP64 oracle
    on every 16th row, and FP32C against FP64 on the whole grid."""
    L = configs.config2(n=161)
    stride = 16
    ref = oracle.forward(L.px, L.py, L.t, L.planes, row_stride=stride)
    rows = np.arange(0, 161, stride)
    mask = np.zeros((161, 161), bool)
    mask[rows] = True
    mask = mask.ravel()
    f64, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp64")
    f32c, i32c = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp32c")
    e64 = rel_l2(f64[mask], ref[mask])
    e32 = rel_l2(f32c[mask], ref[mask])
    efull = rel_l2(f32c, f64)
    eint = rel_l2(i32c, np.abs(f64) ** 2)
    print(f"config2 161: fp64 vs oracle {e64:.3e}; fp32c vs oracle {e32:.3e}; "
          f"fp32c vs fp64 full {efull:.3e}, |E|^2 {eint:.3e}, maxrel {max_rel(f32c, f64):.3e}")
    assert e64 <= 1e-12
    assert e32 <= 1e-5
    assert efull <= 1e-5 and eint <= 2e-5


@pytest.mark.parametrize("n,stride", [(321, 32), (1024, 128)])
def test_config2_large_grids_vs_oracle(torch_cuda, rsctx, n, stride):
    """The headline lens (91,809 pillars) on its other two grids, 321^2 and 1024^2: FP32C
    and FP64 against the FP64 oracle on every stride-th row, FP32C against FP64 on the
    whole grid (field and |E|^2), and the FP32C adjoint of a whole-grid cotangent
    against the FP64 adjoint."""
    L = configs.config2(n=n)
    ref = oracle.forward(L.px, L.py, L.t, L.planes, row_stride=stride)
    rows = np.arange(0, n, stride)
    mask = np.zeros((n, n), bool)
    mask[rows] = True
    mask = mask.ravel()
    f64, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp64")
    f32c, i32c = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, L.planes, "fp32c")
    e64 = rel_l2(f64[mask], ref[mask])
    e32 = rel_l2(f32c[mask], ref[mask])
    efull = rel_l2(f32c, f64)
    eint = rel_l2(i32c, np.abs(f64) ** 2)
    print(f"config2 {n}: fp64 vs oracle {e64:.3e}; fp32c vs oracle {e32:.3e}; "
          f"fp32c vs fp64 full {efull:.3e}, |E|^2 {eint:.3e}", flush=True)
    assert e64 <= 1e-12
    assert e32 <= 1e-5
    assert efull <= 1e-5 and eint <= 2e-5
    rng = np.random.default_rng(n)
    u = rng.standard_normal(n * n) + 1j * rng.standard_normal(n * n)
    du = dev(torch_cuda, u)
    dpx, dpy = dev(torch_cuda, L.px), dev(torch_cuda, L.py)
    g32 = rsctx.adjoint(dpx, dpy, du, L.planes, nsets=1, mode="fp32c").cpu().numpy()
    g64 = rsctx.adjoint(dpx, dpy, du, L.planes, nsets=1, mode="fp64").cpu().numpy()
    eg = rel_l2(g32, g64)
    # the FP64 adjoint against the oracle on a pillar sample
    pick = np.arange(0, L.npil, L.npil // 48)
    gref = oracle.adjoint(L.px[pick], L.py[pick], u, L.planes)
    eg64 = rel_l2(g64[0, pick], gref[0])
    print(f"config2 {n}: adjoint fp32c vs fp64 {eg:.3e}; fp64 vs oracle (sample) {eg64:.3e}")
    assert eg <= 1e-5 and eg64 <= 1e-12


def test_dropin_psf_matches_oracle(torch_cuda):
    L = configs.config1()
    p = L.planes[0]
    k = p.k
    out = crs.propagate_psf(L.px, L.py, L.t[0], 0.0, 0.0, L.focal, k, 81, 5.0)
    assert out is not None and out.shape == (81, 81)
    ref = np.abs(oracle.forward(L.px, L.py, L.t, L.planes)) ** 2
    e = rel_l2(out.ravel(), ref)
    print("drop-in psf relL2", e)
    assert e <= 2e-5


def test_dropin_zx_matches_oracle(torch_cuda):
    L = configs.small_lens(n_side=24, n=9)
    k = L.planes[0].k
    out = crs.propagate_zx(L.px, L.py, L.t[0], -3.0, 3.0, 31, 10.0, 30.0, 21, k)
    assert out is not None and out.shape == (21, 31)
    ref = np.abs(oracle.forward(L.px, L.py, L.t, [Plane.zx(-3.0, 3.0, 31, 10.0, 30.0, 21, k)])) ** 2
    assert rel_l2(out.ravel(), ref) <= 2e-5


def test_dropin_rejections(torch_cuda):
    """The reference's false returns (propagate.cu:139, :201)."""
    px = np.array([0.0])
    t = np.array([1.0 + 0j])
    assert crs.propagate_psf(px[:0], px[:0], t[:0], 0, 0, 10, 10, 5, 1.0) is None
    assert crs.propagate_psf(px, px, t, 0, 0, 10, 10, 0, 1.0) is None
    assert crs.propagate_zx(px, px, t, -1, 1, 1, 5, 10, 5, 10) is None
    assert crs.propagate_zx(px, px, t, -1, 1, 5, 5, 10, 1, 10) is None


def test_dropin_single_pixel_is_nan(torch_cuda):
    """n == 1: step = 2W/0 = inf, coordinates 0*inf = NaN, |E|^2 = NaN, as in both
    reference paths (propagate.cu:143,51; focal.cpp:223-227) and the oracle."""
    px, py = configs.square_grid(4, 0.5)
    t = np.ones(px.size, complex)
    out = crs.propagate_psf(px, py, t, 0, 0, 10, 10, 1, 1.0)
    assert out is not None and out.shape == (1, 1) and np.isnan(out[0, 0])
    ref = oracle.forward(px, py, t[None], [Plane.psf(0, 0, 10, 10, 1, 1.0)])
    assert np.isnan(ref[0])


def test_dropin_two_pixel_grid(torch_cuda):
    px, py = configs.square_grid(5, 0.5)
    t = np.exp(1j * np.linspace(0, 1, px.size))
    out = crs.propagate_psf(px, py, t, 0.2, 0.1, 12, 9, 2, 0.7)
    ref = np.abs(oracle.forward(px, py, t[None], [Plane.psf(0.2, 0.1, 12, 9, 2, 0.7)])) ** 2
    assert rel_l2(out.ravel(), ref) <= 2e-5


def test_default_mode_switch(torch_cuda):
    L = configs.small_lens(n_side=16, n=9)
    k = L.planes[0].k
    W = -L.planes[0].x0
    prev = crs.set_default_mode("fp64")
    try:
        out = crs.propagate_psf(L.px, L.py, L.t[0], 0, 0, L.focal, k, 9, W)
    finally:
        crs.set_default_mode(prev)
    ref = np.abs(oracle.forward(L.px, L.py, L.t, L.planes)) ** 2
    assert rel_l2(out.ravel(), ref) <= 1e-12


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_forward_general_and_first_order_planes(torch_cuda, rsctx, mode, tol):
    """One launch whose work-groups take different series orders (see the adjoint test
    of the same name): each plane against the oracle."""
    L = configs.config1()
    k = L.planes[0].k
    planes = [L.planes[0], Plane.psf(0.0, 0.0, 50.0, k, 41, 1.0)]
    ref = oracle.forward(L.px, L.py, L.t, planes)
    f, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, planes, mode)
    for p in crs.resolve_offsets(planes):
        sl = slice(p.offset, p.offset + p.npix)
        e = rel_l2(f[sl], ref[sl])
        print(f"forward mixed orders {mode} plane {p.nx}x{p.ny} relL2 {e:.3e}")
        assert e <= tol


def test_mixed_exact_and_blocked_planes(torch_cuda, rsctx):
    """A call mixing a plane the block expansion cannot serve (0.5 um from the lens,
    1 um pixels: every pair at its own distance) with planes it can: the 4 x 4 kernels
    run, take the flagged plane pair by pair, and every plane, forward and adjoint,
    matches the oracle to 1e-5."""
    L = configs.config1()
    k = L.planes[0].k
    planes = [Plane.psf(0.0, 0.0, 50.0, k, 41, 1.0), Plane.psf(0.3, -0.2, 0.5, k, 9, 4.0),
              L.planes[0]]
    ref = oracle.forward(L.px, L.py, L.t, planes)
    f, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, planes, "fp32c")
    assert rsctx.last_launch(False)[0] == "crs::fwd_fp32c<4, 4>"  # first of its launches
    for p in crs.resolve_offsets(planes):
        sl = slice(p.offset, p.offset + p.npix)
        e = rel_l2(f[sl], ref[sl])
        print(f"mixed exact/blocked forward plane z={p.z0} {p.nx}x{p.ny} relL2 {e:.3e}")
        assert e <= 1e-5
    rng = np.random.default_rng(11)
    u = rng.standard_normal(ref.size) + 1j * rng.standard_normal(ref.size)
    gref = oracle.adjoint(L.px, L.py, u, planes)
    g = rsctx.adjoint(dev(torch_cuda, L.px), dev(torch_cuda, L.py), dev(torch_cuda, u), planes,
                      nsets=1, mode="fp32c")
    assert rsctx.last_launch(True)[0] == "crs::adj_fp32c<8, 8>"  # the built-in adjoint shape
    eg = rel_l2(g.cpu().numpy(), gref)
    print(f"mixed exact/blocked adjoint relL2 {eg:.3e}")
    assert eg <= 1e-5
    # The near plane alone still takes the 1 x 1 kernels.
    gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, planes[1:2], "fp32c")
    assert rsctx.last_launch(False)[0] == "crs::fwd_fp32c<1, 1>"


@pytest.mark.parametrize("mode,tol,shape,ashape", [
    ("fp64", 1e-12, None, None), ("fp32c", 1e-5, (4, 4), (4, 4)), ("fp32c", 1e-5, (4, 8), (4, 8)),
    ("fp32c", 1e-5, (8, 4), (8, 4)), ("fp32c", 1e-5, (4, 4), (8, 8))])
def test_forward_ragged_edges_blocked(torch_cuda, rsctx, mode, tol, shape, ashape):
    """Blocks of every forward shape with every kind of leftover: nx, ny = 4k + 1, 2, 3 and
    8k + 1 ... 7 (the 1 x BY, BX x 1 and 1 x 1 edge tiles), rectangular planes, planes
    narrower or shorter than one block."""
    L = configs.config1()
    k = L.planes[0].k
    planes = [
        Plane.psf(0.0, 0.0, 50.0, k, 42, 1.0),                           # 42 = 4*10 + 2 = 8*5 + 2
        Plane(k, -1.1, 0.05, -0.9, 0.05, 50.0, 0.0, 43, 41),             # 43 = 8*5 + 3, 41 = 8*5 + 1
        Plane(k, 0.3, 0.05, -0.5, 0.05, 49.0, 0.0, 3, 22),               # narrower than a block
        Plane(k, -0.4, 0.04, 0.2, 0.04, 51.0, 0.0, 17, 2),               # shorter than a block
        Plane(k, -0.6, 0.03, -0.3, 0.03, 50.5, 0.0, 15, 23),             # 15 = 8 + 7, 23 = 16 + 7
    ]
    ref = oracle.forward(L.px, L.py, L.t, planes)
    if shape:
        rsctx.set_block_shape(fwd=shape, adj=ashape)
    try:
        f, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, planes, mode)
        for p in crs.resolve_offsets(planes):
            sl = slice(p.offset, p.offset + p.npix)
            e = rel_l2(f[sl], ref[sl])
            print(f"ragged {mode} {shape} {p.nx}x{p.ny} relL2 {e:.3e}")
            assert e <= tol
        if mode == "fp32c":
            assert rsctx.last_launch(False)[0] == f"crs::fwd_fp32c<{shape[0]}, {shape[1]}>"
            rng = np.random.default_rng(5)
            u = rng.standard_normal(ref.size) + 1j * rng.standard_normal(ref.size)
            gref = oracle.adjoint(L.px, L.py, u, planes)
            g = rsctx.adjoint(dev(torch_cuda, L.px), dev(torch_cuda, L.py), dev(torch_cuda, u),
                              planes, nsets=1, mode="fp32c")
            assert rsctx.last_launch(True)[0] == f"crs::adj_fp32c<{ashape[0]}, {ashape[1]}>"
            eg = rel_l2(g.cpu().numpy(), gref)
            print(f"ragged adjoint {ashape} relL2 {eg:.3e}")
            assert eg <= tol
    finally:
        rsctx.set_block_shape()


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_zx_slice_through_lens_plane(torch_cuda, rsctx, mode, tol):
    """A zx slice whose rows run from z = -6 to +6 um (none at 0): 4-row pixel blocks
    straddle the lens plane, where no distance bound holds, so the series expansion must
    not be used (plane_min_abs_z = 0 sends the call to the exact path).  Forward and
    adjoint against the oracle; the reference evaluates R with z^2, symmetric in z."""
    L = configs.small_lens(n_side=16, n=9)
    k = L.planes[0].k
    planes = [Plane.zx(-3.0, 3.0, 33, -6.0, 6.0, 12, k)]
    ref = oracle.forward(L.px, L.py, L.t, planes)
    f, _ = gpu_forward(torch_cuda, rsctx, L.px, L.py, L.t, planes, mode)
    e = rel_l2(f, ref)
    print(f"zx through z = 0 {mode}: forward relL2 {e:.3e}")
    assert e <= tol
    rng = np.random.default_rng(3)
    u = rng.standard_normal(ref.size) + 1j * rng.standard_normal(ref.size)
    gref = oracle.adjoint(L.px, L.py, u, planes)
    g = rsctx.adjoint(torch_cuda.from_numpy(L.px).cuda(), torch_cuda.from_numpy(L.py).cuda(),
                      torch_cuda.from_numpy(u).cuda(), planes, nsets=1, mode=mode)
    eg = rel_l2(g.cpu().numpy(), gref)
    print(f"zx through z = 0 {mode}: adjoint relL2 {eg:.3e}")
    assert eg <= tol
