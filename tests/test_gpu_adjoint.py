"""GPU parity of the adjoint kernel and the autograd bridge.

The adjoint has no reference counterpart; it is pinned by the FP64 oracle's adjoint
(itself pinned by the dot-product identity), by the dot-product identity on the GPU,
and by central finite differences of the encircled-energy loss.
Tolerances: FP64 <= 1e-12, FP32C <= 1e-5 relative L2 of the gradient vector.
"""
import math

import numpy as np
import pytest

import oracle
from celeris_amd import Plane, configs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_adjoint(torch, ctx, px, py, u, planes, nsets, mode):
    print(f"[gpu_adjoint] mode={mode} npil={np.size(px)} planes={len(planes)}", flush=True)
    g = ctx.adjoint(dev(torch, px), dev(torch, py), dev(torch, u), planes, nsets=nsets, mode=mode)
    torch.cuda.synchronize()
    return g.cpu().numpy()


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_adjoint_small_vs_oracle(torch_cuda, rsctx, mode, tol):
    L = configs.small_lens(n_side=21, n=15)
    rng = np.random.default_rng(11)
    u = rng.standard_normal(225) + 1j * rng.standard_normal(225)
    ref = oracle.adjoint(L.px, L.py, u, L.planes)
    g = gpu_adjoint(torch_cuda, rsctx, L.px, L.py, u, L.planes, 1, mode)
    e = rel_l2(g, ref)
    print(f"adjoint {mode} relL2 {e:.3e}")
    assert e <= tol


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_adjoint_multiset_vs_oracle(torch_cuda, rsctx, mode, tol):
    px, py = configs.square_grid(40, 0.3)  # 1,600 pillars
    planes = [
        Plane.psf(0.0, 0.0, 25.0, 2 * math.pi / 0.5, 21, 2.0, tset=0),
        Plane.psf(0.1, 0.0, 27.0, 2 * math.pi / 0.6, 9, 1.0, tset=2),
        Plane.zx(-2.0, 2.0, 13, 15.0, 35.0, 7, 2 * math.pi / 0.6, tset=2),
        Plane.axis(10.0, 40.0, 17, 2 * math.pi / 0.5, tset=0),
    ]
    total = sum(p.npix for p in planes)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    ref = oracle.adjoint(px, py, u, planes, nsets=3)
    g = gpu_adjoint(torch_cuda, rsctx, px, py, u, planes, 3, mode)
    assert np.all(g[1] == 0)  # no plane uses set 1
    for s in (0, 2):
        e = rel_l2(g[s], ref[s])
        print(mode, "set", s, e)
        assert e <= tol


def test_dot_product_identity(torch_cuda, rsctx):
    """Re<G t, u> = Re<t, G^H u> with the GPU forward and adjoint (config 1 size)."""
    L = configs.config1()
    rng = np.random.default_rng(2)
    u = rng.standard_normal(81 * 81) + 1j * rng.standard_normal(81 * 81)
    for mode, tol in (("fp64", 1e-12), ("fp32c", 1e-5)):
        f, _ = rsctx.forward(dev(torch_cuda, L.px), dev(torch_cuda, L.py), dev(torch_cuda, L.t),
                             L.planes, mode=mode)
        g = rsctx.adjoint(dev(torch_cuda, L.px), dev(torch_cuda, L.py), dev(torch_cuda, u),
                          L.planes, mode=mode)
        lhs = np.vdot(u, f.cpu().numpy())
        rhs = np.vdot(g.cpu().numpy()[0], L.t[0])
        err = abs(lhs - rhs) / abs(lhs)
        print(mode, "dot identity", err)
        assert err <= tol


def _ee_loss(torch, field, plane, r0):
    from celeris_amd.autograd import encircled_energy

    inten = field.real ** 2 + field.imag ** 2
    return -encircled_energy(inten, plane, r0)


def test_autograd_fd_encircled_energy(torch_cuda):
    """d(-EE)/d(phase) and d(-EE)/d(amplitude) through the adjoint kernel vs central
    finite differences (h = 1e-4 rad / 1e-4) computed with the FP64 forward."""
    torch = torch_cuda
    from celeris_amd.autograd import rs_field

    L = configs.small_lens(n_side=24, n=31, focal=25.0)
    plane = L.planes[0]
    r0 = 1.22 * L.wavelengths[0] / L.na
    px, py = dev(torch, L.px), dev(torch, L.py)
    amp0 = np.abs(L.t[0])
    phi0 = np.angle(L.t[0])

    def loss_np(phi, amp, mode="fp64"):
        t = torch.polar(dev(torch, amp), dev(torch, phi))[None]
        return float(_ee_loss(torch, rs_field(t, px, py, [plane], mode=mode), plane, r0))

    phi = dev(torch, phi0).requires_grad_(True)
    amp = dev(torch, amp0).requires_grad_(True)
    for mode, tol in (("fp64", 1e-7), ("fp32c", 1e-4)):
        phi.grad = None
        amp.grad = None
        t = torch.polar(amp, phi)[None]
        loss = _ee_loss(torch, rs_field(t, px, py, [plane], mode=mode), plane, r0)
        loss.backward()
        gphi = phi.grad.cpu().numpy()
        gamp = amp.grad.cpu().numpy()
        pick = np.random.default_rng(1).choice(L.npil, 8, replace=False)
        h = 1e-4
        for q in pick:
            e = np.zeros(L.npil)
            e[q] = h
            fd_phi = (loss_np(phi0 + e, amp0) - loss_np(phi0 - e, amp0)) / (2 * h)
            fd_amp = (loss_np(phi0, amp0 + e) - loss_np(phi0, amp0 - e)) / (2 * h)
            scale_p = np.max(np.abs(gphi))
            scale_a = np.max(np.abs(gamp))
            print(mode, q, gphi[q], fd_phi, gamp[q], fd_amp)
            assert abs(gphi[q] - fd_phi) <= tol * scale_p
            assert abs(gamp[q] - fd_amp) <= tol * scale_a


def test_config4_gradient_fp32c_vs_fp64(torch_cuda):
    """Config 4: -EE within 1.22 lambda/NA at z = f, 91,809 pillars, 161^2 grid.  The
    FP32C gradient with respect to every pillar's phase and amplitude (adjoint kernel
    through the autograd Function) against
      * the FP64 GPU gradient, over all pillars;
      * central finite differences, h = 1e-4 rad (phase) and 1e-4 (amplitude), on 64
        seeded pillars, for both (FD exploits linearity: E(t_q + dt) = E + dt G_q with
        G_q the FP64 field of pillar q alone);
      * the FP64 CPU gradient of the oracle's adjoint on those 64 pillars,
        dL/dphi_q = -Im(conj(g_q) t_q), dL/da_q = Re(conj(g_q) e^{i phi_q}),
        g_q = sum_p conj(G_pq) u_p with u = dL/dRe E + i dL/dIm E at the FP64 field."""
    torch = torch_cuda
    import oracle
    from celeris_amd import Context
    from celeris_amd.autograd import rs_field

    L = configs.config2(n=161)
    plane = L.planes[0]
    r0 = 1.22 * L.wavelengths[0] / L.na
    px, py = dev(torch, L.px), dev(torch, L.py)
    grads = {}
    for mode in ("fp64", "fp32c"):
        phi = dev(torch, L.meta["phase"]).requires_grad_(True)
        amp = dev(torch, L.meta["amp"]).requires_grad_(True)
        t = torch.polar(amp, phi)[None]
        loss = _ee_loss(torch, rs_field(t, px, py, [plane], mode=mode), plane, r0)
        loss.backward()
        grads[mode] = (phi.grad.cpu().numpy(), amp.grad.cpu().numpy(), float(loss.detach()))
    ep = rel_l2(grads["fp32c"][0], grads["fp64"][0])
    ea = rel_l2(grads["fp32c"][1], grads["fp64"][1])
    print(f"config4 grad relL2 phase {ep:.3e} amp {ea:.3e} loss {grads['fp64'][2]:.6f}")
    assert ep <= 1e-5 and ea <= 1e-5

    ctx = Context(0)
    E, _ = ctx.forward(px, py, dev(torch, L.t), [plane], mode="fp64")
    pick = np.unique((configs.uniform(configs.SEED, configs.STREAM_FD_PILLARS, 64) * L.npil).astype(int))
    assert pick.size >= 60
    h = 1e-4
    gphi32, gamp32 = grads["fp32c"][0], grads["fp32c"][1]
    sp, sa = np.max(np.abs(grads["fp64"][0])), np.max(np.abs(grads["fp64"][1]))
    one = torch.ones((1, 1), dtype=torch.complex128, device="cuda")
    worst_p = worst_a = 0.0
    for q in pick:
        Gq, _ = ctx.forward(px[q:q + 1].contiguous(), py[q:q + 1].contiguous(), one, [plane],
                            mode="fp64")
        tq = complex(L.t[0, q])
        eq = tq / abs(tq)
        lp = float(_ee_loss(torch, E + tq * (np.exp(1j * h) - 1) * Gq, plane, r0))
        lm = float(_ee_loss(torch, E + tq * (np.exp(-1j * h) - 1) * Gq, plane, r0))
        fd_phi = (lp - lm) / (2 * h)
        lp = float(_ee_loss(torch, E + h * eq * Gq, plane, r0))
        lm = float(_ee_loss(torch, E - h * eq * Gq, plane, r0))
        fd_amp = (lp - lm) / (2 * h)
        worst_p = max(worst_p, abs(gphi32[q] - fd_phi) / sp)
        worst_a = max(worst_a, abs(gamp32[q] - fd_amp) / sa)
    print(f"config4 FD on {pick.size} pillars: worst |fp32c - FD| / max|g|: phase {worst_p:.2e}, "
          f"amplitude {worst_a:.2e}")
    assert worst_p <= 1e-5 and worst_a <= 1e-5

    # FP64 CPU gradient (the oracle's adjoint) on the same pillars.
    Ev = E.clone().requires_grad_(True)
    _ee_loss(torch, Ev, plane, r0).backward()
    u = Ev.grad.detach().cpu().numpy()
    g = oracle.adjoint(L.px[pick], L.py[pick], u, [plane])[0]
    tq = L.t[0, pick]
    cpu_phi = -np.imag(np.conj(g) * tq)
    cpu_amp = np.real(np.conj(g) * tq / np.abs(tq))
    dp = np.max(np.abs(gphi32[pick] - cpu_phi)) / sp
    da = np.max(np.abs(gamp32[pick] - cpu_amp)) / sa
    dp64 = np.max(np.abs(grads["fp64"][0][pick] - cpu_phi)) / sp
    print(f"config4 vs FP64 CPU gradient on {pick.size} pillars: fp32c phase {dp:.2e} amp {da:.2e}; "
          f"fp64 GPU phase {dp64:.2e}")
    assert dp <= 1e-5 and da <= 1e-5 and dp64 <= 1e-10
    ctx.close()


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_adjoint_general_and_first_order_planes(torch_cuda, rsctx, mode, tol):
    """One launch whose work-groups take different series orders: the config-1 window
    (81 x 81 over +-5 um at f = 50: 4 x 4 blocks need the second-order terms) and a
    finer 41 x 41 window over +-1 um (first order suffices), both against the oracle."""
    L = configs.config1()
    k = L.planes[0].k
    planes = [L.planes[0], Plane.psf(0.0, 0.0, 50.0, k, 41, 1.0)]
    total = sum(p.npix for p in planes)
    rng = np.random.default_rng(17)
    u = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    ref = oracle.adjoint(L.px, L.py, u, planes)
    g = gpu_adjoint(torch_cuda, rsctx, L.px, L.py, u, planes, 1, mode)
    e = rel_l2(g, ref)
    print(f"adjoint mixed orders {mode} relL2 {e:.3e}")
    assert e <= tol
    # each plane alone, so that an error confined to one order's path cannot hide
    off = 0
    for p in planes:
        up = u[off:off + p.npix]
        off += p.npix
        gp = gpu_adjoint(torch_cuda, rsctx, L.px, L.py, up, [Plane(**{**p.__dict__, "offset": -1})], 1, mode)
        ep = rel_l2(gp, oracle.adjoint(L.px, L.py, up, [p]))
        print(f"  plane {p.nx}x{p.ny} relL2 {ep:.3e}")
        assert ep <= tol


def test_inverse_design_loop_raises_encircled_energy(torch_cuda):
    """The design loop the adjoint feeds: Adam on the phases of a 24 x 24 lens from
    seeded random phases.  Encircled energy within 1.22 lambda f / D of the centre of a
    41 x 41 window at f must rise several-fold, and every step's gradient comes from the
    adjoint kernel (FP32C)."""
    from celeris_amd import configs
    from celeris_amd.inverse_design import optimize_phase

    px, py = configs.square_grid(24, 0.4)
    lam, f = 0.532, 15.0
    k = 2 * math.pi / lam
    D = 24 * 0.4
    W = 3.0
    plane = Plane.psf(0.0, 0.0, f, k, 41, W)
    rng = np.random.default_rng(5)
    phi0 = rng.uniform(-math.pi, math.pi, px.size)
    amp = np.full(px.size, 0.9)
    print("[adjoint] inverse design loop", flush=True)
    res = optimize_phase(px, py, amp, phi0, plane, radius=1.22 * lam * f / D, steps=40, lr=0.15)
    ee = res.encircled_energy
    print(f"inverse design EE {ee[0]:.4f} -> {ee[-1]:.4f}")
    assert len(ee) == 41 and res.phase.shape == phi0.shape
    assert ee[-1] > 3.0 * ee[0] and ee[-1] > 0.3


def test_small_problem_work_splits(torch_cuda, rsctx):
    """Small problems: at a high work-group target the forward splits the pillar sum into
    64-pillar chunks (whole 256-pillar tiles at a low one) and the adjoint into more quad
    runs.  Both splits against the oracle, and against each other."""
    L = configs.config1()
    planes = L.planes
    rng = np.random.default_rng(11)
    npix = planes[0].npix
    u = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    ref_g = oracle.adjoint(L.px, L.py, u, planes)
    ref_f = oracle.forward(L.px, L.py, L.t, planes)
    dpx, dpy = torch_cuda.from_numpy(L.px).cuda(), torch_cuda.from_numpy(L.py).cuda()
    dt, du = torch_cuda.from_numpy(L.t).cuda(), torch_cuda.from_numpy(u).cuda()
    out = {}
    for target, pq in ((4096, 1), (64, 4)):  # pq: key of the result only
        rsctx.set_tuning("fp32c", target, target)
        print(f"[adjoint] small problem, work-group target {target}", flush=True)
        f, _ = rsctx.forward(dpx, dpy, dt, planes, mode="fp32c")
        g = rsctx.adjoint(dpx, dpy, du, planes, nsets=1, mode="fp32c")
        torch_cuda.cuda.synchronize()
        name, groups = rsctx.last_launch(True)
        fname, fgroups = rsctx.last_launch(False)
        print(f"  {fname} x {fgroups}, {name} x {groups}")
        assert name.startswith("crs::adj_fp32c<") and fname.startswith("crs::fwd_fp32c<")
        assert "<1, 1>" not in name and "<1, 1>" not in fname
        out[pq] = (f.cpu().numpy(), g.cpu().numpy(), fgroups)
        assert rel_l2(out[pq][0], ref_f) <= 1e-5 and rel_l2(out[pq][1], ref_g) <= 1e-5
    rsctx.set_tuning("fp32c", 0, 0)
    assert out[1][2] > out[4][2]  # finer forward chunks at the high target
    assert rel_l2(out[1][1], out[4][1]) <= 1e-6
