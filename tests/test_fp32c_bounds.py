"""The FP32C series-order bound, checked against the exact dropped terms (CPU).

The 4 x 4 kernels keep R - R0 = x1 (1 - y) and R0/R = 1 - 2y (first order, with
x1 = D h, y = D h^2, h = 1/(2 R0), R^2 = R0^2 + D) when, for the plane and the pillars'
extent, the dropped terms are bounded by 1.6e-7 turns and 4e-6 (fp32c_first_order in
celeris_amd/csrc/rs_kernels.hip).  The bound is restated here line by line and checked
against the exact remainders (R - R0) - x1 (1 - y) and R0/R - (1 - 2y), evaluated
without cancellation, over random planes, pillars and pixel blocks: no pair may exceed
the bound its plane was given.  A mistake in the derivation (a dropped factor, the
wrong root s_k, the obliquity taken at the wrong point) shows up as a violation.

The second order is taken wherever that check fails, on any plane the host lets use
blocks at all (fp32c_plane_order in celeris_amd/csrc/rs_api.cpp, a cruder bound for
pillars at any angle); its exact remainders are checked against the budgets over random
planes and pillars at every obliquity.
"""
import math

import numpy as np


def plane_min_abs_z(z0, sz, ny):
    zl = z0 + (ny - 1) * sz
    if (z0 > 0 and zl > 0) or (z0 < 0 and zl < 0):
        return min(abs(z0), abs(zl))
    return 0.0


def first_order_bounds(P, ext, bx, by):
    """(phase bound in turns, amplitude bound) as fp32c_first_order computes them."""
    kt = P["k"] / (2 * math.pi)
    xl = P["x0"] + (P["nx"] - 1) * P["sx"]
    yl = P["y0"] + (P["ny"] - 1) * P["sy"]
    rx = max(abs(P["x0"]), abs(xl)) + ext[0]
    ry = max(abs(P["y0"]), abs(yl)) + ext[1]
    zmin = plane_min_abs_z(P["z0"], P["sz"], P["ny"])
    rho2 = rx * rx + ry * ry
    sn = math.sqrt(rho2 / (rho2 + zmin * zmin))
    hx = 0.5 * (bx - 1) * abs(P["sx"]) * kt
    hy = 0.5 * (by - 1) * abs(P["sy"]) * kt
    hz = 0.5 * (by - 1) * abs(P["sz"]) * kt
    e = math.sqrt(hx * hx + hy * hy)
    z = zmin * kt
    a, c = 2 * e, 2 * hz + (e * e + hz * hz) / z
    s3 = min(sn, (math.sqrt(c * c + 15 * a * a) - c) / (5 * a)) if a > 0 else 0.0
    s2 = min(sn, (math.sqrt(c * c + 8 * a * a) - c) / (4 * a)) if a > 0 else 0.0
    B3, B2 = a * s3 + c, a * s2 + c
    return B3 ** 3 * (1 - s3 * s3) / (16 * z * z), 3 * B2 * B2 * (1 - s2 * s2) / (8 * z * z)


def dropped_terms(P, px, py, ox, oy):
    """Exact remainders of the first-order series for every pixel of the 4 x 4 block
    with origin (ox, oy) against pillars (px, py), in turns / relative."""
    kt = P["k"] / (2 * math.pi)
    ic, jc = ox + 1.5, oy + 1.5
    FXc, FYc = kt * (P["x0"] + ic * P["sx"]), kt * (P["y0"] + jc * P["sy"])
    zc = kt * (P["z0"] + jc * P["sz"])
    X, Y = kt * px, kt * py
    dX0, dY0 = FXc - X, FYc - Y
    R0 = np.sqrt(dX0 * dX0 + dY0 * dY0 + zc * zc)
    h = 0.5 / R0
    ph, am = np.zeros_like(R0), np.zeros_like(R0)
    for r in range(4):
        ey = kt * P["sy"] * (r - 1.5)
        zr = kt * (P["z0"] + (oy + r) * P["sz"])
        for c in range(4):
            ex = kt * P["sx"] * (c - 1.5)
            D = ex * (2 * dX0 + ex) + ey * (2 * dY0 + ey) + (zr * zr - zc * zc)
            R = np.sqrt(R0 * R0 + D)
            x1, y = D * h, D * h * h
            dR = D / (R + R0)                      # R - R0 without cancellation
            ph = np.maximum(ph, np.abs(dR - x1 * (1 - y)))
            am = np.maximum(am, np.abs(R0 / R - (1 - 2 * y)))
    return float(ph.max()), float(am.max())


def test_first_order_bound_covers_every_pair():
    rng = np.random.default_rng(20261203)
    checked = 0
    for _ in range(300):
        lam = rng.uniform(0.4, 1.0)
        k = 2 * math.pi / lam
        half = rng.uniform(2.0, 80.0)             # pillar extent (um)
        n = int(rng.integers(4, 200))
        if rng.random() < 0.75:                   # psf window
            z0, sz = float(np.exp(rng.uniform(np.log(2.0), np.log(400.0)))), 0.0
            step = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3))))
            sy = step
        else:                                     # zx slice: rows step in z
            z0 = float(np.exp(rng.uniform(np.log(5.0), np.log(400.0))))
            sz = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3))))
            step, sy = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3)))), 0.0
        x0 = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * step
        y0 = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * sy
        P = dict(k=k, x0=x0, sx=step, y0=y0, sy=sy, z0=z0, sz=sz, nx=n, ny=n)
        bphase, bamp = first_order_bounds(P, (half, half), 4, 4)
        # Pillars anywhere within the extent, including its corners, and blocks anywhere
        # in the plane (the corners are where the bound is tightest).
        px = np.concatenate([rng.uniform(-half, half, 400), [half, -half, half, -half]])
        py = np.concatenate([rng.uniform(-half, half, 400), [half, half, -half, -half]])
        for ox, oy in [(0, 0), (n - 4, n - 4), (0, n - 4), (n - 4, 0),
                       (int(rng.integers(0, n - 3)), int(rng.integers(0, n - 3)))]:
            ph, am = dropped_terms(P, px, py, ox, oy)
            assert ph <= bphase * (1 + 1e-9) + 1e-15, (P, ph, bphase)
            assert am <= bamp * (1 + 1e-9) + 1e-15, (P, am, bamp)
            checked += 1
    assert checked == 1500


def test_bound_is_not_vacuous():
    """The bound stays within a modest factor of the worst observed remainder on the
    headline workload's plane (161 x 161 over +-5.4 um at z = 100 um, 91,809 pillars
    within +-52.85 um): it would not pass a geometry whose real error is far larger."""
    lam, f = 0.633, 100.0
    k = 2 * math.pi / lam
    W = 5.404873920333813
    step = 2 * W / 160
    P = dict(k=k, x0=-W, sx=step, y0=-W, sy=step, z0=f, sz=0.0, nx=161, ny=161)
    half = 151 * 0.35
    bphase, bamp = first_order_bounds(P, (half, half), 4, 4)
    g = np.linspace(-half, half, 61)
    X, Y = np.meshgrid(g, g)
    ph, am = dropped_terms(P, X.ravel(), Y.ravel(), 0, 0)
    assert bphase <= 1.6e-7 and bamp <= 4e-6       # the headline runs first order
    assert ph <= bphase and am <= bamp
    assert bphase < 20 * ph and bamp < 20 * am


def host_plane_order(P):
    """fp32c_plane_order (celeris_amd/csrc/rs_api.cpp): 4 = pair by pair, else blocks."""
    kt = P["k"] / (2 * math.pi)
    ex, ey, ez = (1.5 * abs(P[s]) * kt for s in ("sx", "sy", "sz"))
    e = math.sqrt(ex * ex + ey * ey) + ez
    zmin = plane_min_abs_z(P["z0"], P["sz"], P["ny"]) * kt
    if not zmin > 0:
        return 4
    r, po, ao = e / zmin, 1, 1
    if 0.5 * e * r * r > 1.6e-7:
        if 0.625 * e * r ** 3 > 1.6e-7:
            return 4
        po = 2
    if 1.5 * r * r > 4e-6:
        if 2.5 * r ** 3 > 4e-6:
            return 4
        ao = 2
    return 2 * (po - 1) + (ao - 1)


def second_order_terms(P, px, py, ox, oy):
    """Exact remainders of the second-order series (R - R0 = x1 (1 - y (1 - 2y)),
    R0/R = 1 + y (6y - 2)) over the 4 x 4 block at (ox, oy)."""
    kt = P["k"] / (2 * math.pi)
    ic, jc = ox + 1.5, oy + 1.5
    FXc, FYc = kt * (P["x0"] + ic * P["sx"]), kt * (P["y0"] + jc * P["sy"])
    zc = kt * (P["z0"] + jc * P["sz"])
    dX0, dY0 = FXc - kt * px, FYc - kt * py
    R0 = np.sqrt(dX0 * dX0 + dY0 * dY0 + zc * zc)
    h = 0.5 / R0
    ph, am = np.zeros_like(R0), np.zeros_like(R0)
    for r in range(4):
        ey = kt * P["sy"] * (r - 1.5)
        zr = kt * (P["z0"] + (oy + r) * P["sz"])
        for c in range(4):
            ex = kt * P["sx"] * (c - 1.5)
            D = ex * (2 * dX0 + ex) + ey * (2 * dY0 + ey) + (zr * zr - zc * zc)
            R = np.sqrt(R0 * R0 + D)
            x1, y = D * h, D * h * h
            ph = np.maximum(ph, np.abs(D / (R + R0) - x1 * (1 - y * (1 - 2 * y))))
            am = np.maximum(am, np.abs(R0 / R - (1 + y * (6 * y - 2))))
    return float(ph.max()), float(am.max())


def test_host_order_keeps_second_order_within_budget():
    """Where the host lets a plane use 4 x 4 blocks at all, the kernels may take the
    second order for any pillars, so its remainders must meet the budgets (1.6e-7
    turns, 4e-6) for pillars at every obliquity: here from 1e-3 to 1e3 times the plane
    distance, in all directions, against blocks at the plane's corners."""
    rng = np.random.default_rng(7)
    allowed, worst_p, worst_a = 0, 0.0, 0.0
    for _ in range(1000):
        lam = rng.uniform(0.4, 1.0)
        n = int(rng.integers(4, 120))
        z0 = float(np.exp(rng.uniform(np.log(0.5), np.log(300.0))))
        pitch = lambda: lam * float(np.exp(rng.uniform(np.log(0.005), np.log(1.0))))  # noqa: E731
        if rng.random() < 0.7:
            step = pitch()
            sy, sz = step, 0.0
        else:
            step, sy, sz = pitch(), 0.0, pitch()
        P = dict(k=2 * math.pi / lam, x0=-0.5 * (n - 1) * step, sx=step,
                 y0=-0.5 * (n - 1) * sy, sy=sy, z0=z0, sz=sz, nx=n, ny=n)
        if host_plane_order(P) == 4:
            continue
        allowed += 1
        rho = np.geomspace(1e-3, 1e3, 200) * z0
        ang = rng.uniform(0, 2 * math.pi, rho.size)
        px, py = rho * np.cos(ang), rho * np.sin(ang)
        for ox, oy in [(0, 0), (n - 4, n - 4), (0, n - 4), (n - 4, 0)]:
            ph, am = second_order_terms(P, px, py, ox, oy)
            worst_p, worst_a = max(worst_p, ph / 1.6e-7), max(worst_a, am / 4e-6)
    print(f"{allowed} planes allowed; worst second-order remainder {worst_p:.2f} of the "
          f"phase budget, {worst_a:.2f} of the amplitude budget")
    assert allowed > 300
    assert worst_p <= 1.0 and worst_a <= 1.0
