"""The FP32C per-pair error bound, checked against the exact dropped terms (CPU).

The kernels expand each pixel's distance about its block centre: with R^2 = R0^2 + D,
h = 1/(2 R0), x1 = D h, y = D h^2,
    R - R0 = x1 (1 - y + 2 y^2 - ...),    R0/R = 1 - 2 y + 6 y^2 - ...
and keep the phase to order 1 or 2 and the amplitude to order 1 or 2.  Which orders a
work item may use is decided from a bound on the dropped terms (celeris_amd/csrc/
rs_internal.h: series_geom, series_peak, fp32c_pair_error, fp32c_order_obliquity): for a
block of half diagonal e', half depth e_z', nearest depth z' (turns) and pairs of
obliquity s <= sn, |D|/R0 <= B(s) = 2 e' s + 2 e_z' + (e'^2 + e_z'^2)/z' and
R0 >= z' / sqrt(1 - s^2), so
    phase order 1:      2 x1 y^2 <= max_s B^3 (1 - s^2) / (16 z'^2)
    phase order 2:      5 x1 y^3 <= max_s 5 B^4 (1 - s^2)^1.5 / (128 z'^3)
    amplitude order 1:  6 y^2    <= max_s 3 B^2 (1 - s^2) / (8 z'^2)
    amplitude order 2:  20 y^3   <= max_s 5 B^3 (1 - s^2)^1.5 / (16 z'^3).
The bound is restated here line by line and checked against the exact remainders,
formed without cancellation, over random planes, block shapes, pillars filling the
extent (corners included) and blocks across the plane: no pair may exceed it.  A mistake
in the derivation (a dropped factor, the wrong root of the derivative, the obliquity taken
at the wrong point, a block offset measured from the wrong centre) shows up as a
violation, and the bound is shown not to be vacuous (the worst pairs come close to it).
The leading dropped term bounds the whole tail up to a factor 1 + 4|y| (|y| < 1e-3
wherever blocks are admitted), which the checks allow.
"""
import math

import numpy as np
import pytest

PAIR_BUDGET = 8e-6
SHAPES = [(4, 4), (4, 8), (8, 4), (8, 8)]


def plane_min_abs_z(z0, sz, ny):
    zl = z0 + (ny - 1) * sz
    if (z0 > 0 and zl > 0) or (z0 < 0 and zl < 0):
        return min(abs(z0), abs(zl))
    return 0.0


def series_geom(P, bx, by):
    kt = P["k"] / (2 * math.pi)
    hx = 0.5 * (bx - 1) * abs(P["sx"]) * kt
    hy = 0.5 * (by - 1) * abs(P["sy"]) * kt
    hz = 0.5 * (by - 1) * abs(P["sz"]) * kt
    e = math.sqrt(hx * hx + hy * hy)
    z = kt * plane_min_abs_z(P["z0"], P["sz"], P["ny"])
    return dict(a=2 * e, c=2 * hz + (e * e + hz * hz) / z, z=z, e=e + hz)


def series_peak(a, c, sn, k, twom):
    """max over 0 <= s <= sn of (a s + c)^k (1 - s^2)^(twom/2), at the root of the derivative."""
    s = 0.0
    if a > 0:
        A, Bq, C = (k + twom) * a, twom * c, -k * a
        s = (-Bq + math.sqrt(Bq * Bq - 4 * A * C)) / (2 * A)
    s = min(max(s, 0.0), sn)
    om = 1 - s * s
    return (a * s + c) ** k * (om if twom == 2 else om * math.sqrt(om))


def truncation_bounds(g, sn, phase_order, amp_order):
    z2, z3 = g["z"] ** 2, g["z"] ** 3
    if phase_order == 1:
        ph = series_peak(g["a"], g["c"], sn, 3, 2) / (16 * z2)
    else:
        ph = 5 * series_peak(g["a"], g["c"], sn, 4, 3) / (128 * z3)
    if amp_order == 1:
        am = 3 * series_peak(g["a"], g["c"], sn, 2, 2) / (8 * z2)
    else:
        am = 5 * series_peak(g["a"], g["c"], sn, 3, 3) / (16 * z3)
    return ph, am


def pair_error(g, sn, order):
    """fp32c_pair_error: order 0 = P1A1, 1 = P1A2, 2 = P2A2."""
    po, ao = {0: (1, 1), 1: (1, 2), 2: (2, 2)}[order]
    ph, am = truncation_bounds(g, sn, po, ao)
    round_phase = 2.0 ** -24 * (2 + 6 * g["e"])
    return 2 * math.pi * (ph + round_phase) + am + 2.0 ** -21 + 2.5e-7


def obliquity(P, ex, ey):
    xl = P["x0"] + (P["nx"] - 1) * P["sx"]
    yl = P["y0"] + (P["ny"] - 1) * P["sy"]
    rx = max(abs(P["x0"]), abs(xl)) + ex
    ry = max(abs(P["y0"]), abs(yl)) + ey
    zmin = plane_min_abs_z(P["z0"], P["sz"], P["ny"])
    rho2 = rx * rx + ry * ry
    return math.sqrt(rho2 / (rho2 + zmin * zmin)) if rho2 > 0 else 0.0


def remainders(P, px, py, ox, oy, bx, by):
    """Exact remainders (phase in turns, amplitude relative) of orders 1 and 2 for every
    pixel of the bx x by block at (ox, oy) against pillars (px, py): max over pairs of
    (phase1, phase2, amp1, amp2)."""
    kt = P["k"] / (2 * math.pi)
    ic, jc = ox + 0.5 * (bx - 1), oy + 0.5 * (by - 1)
    FXc, FYc = kt * (P["x0"] + ic * P["sx"]), kt * (P["y0"] + jc * P["sy"])
    zc = kt * (P["z0"] + jc * P["sz"])
    dX0, dY0 = FXc - kt * px, FYc - kt * py
    R0 = np.sqrt(dX0 * dX0 + dY0 * dY0 + zc * zc)
    h = 0.5 / R0
    out = np.zeros(4)
    for r in range(by):
        ey = kt * P["sy"] * (r - 0.5 * (by - 1))
        zr = kt * (P["z0"] + (oy + r) * P["sz"])
        for c in range(bx):
            ex = kt * P["sx"] * (c - 0.5 * (bx - 1))
            D = ex * (2 * dX0 + ex) + ey * (2 * dY0 + ey) + (zr * zr - zc * zc)
            R = np.sqrt(R0 * R0 + D)
            x1, y = D * h, D * h * h
            dR = D / (R + R0)                      # R - R0 without cancellation
            rr = R0 / R
            out = np.maximum(out, [np.abs(dR - x1 * (1 - y)).max(),
                                   np.abs(dR - x1 * (1 - y * (1 - 2 * y))).max(),
                                   np.abs(rr - (1 - 2 * y)).max(),
                                   np.abs(rr - (1 + y * (6 * y - 2))).max()])
    return out


def random_plane(rng, bx, by):
    lam = rng.uniform(0.4, 1.0)
    n = int(rng.integers(max(bx, by), 200))
    if rng.random() < 0.75:                   # psf window
        z0, sz = float(np.exp(rng.uniform(np.log(2.0), np.log(400.0)))), 0.0
        step = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3))))
        sy = step
    else:                                     # zx slice: rows step in z
        z0 = float(np.exp(rng.uniform(np.log(5.0), np.log(400.0))))
        sz = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3))))
        step, sy = lam * float(np.exp(rng.uniform(np.log(0.01), np.log(0.3)))), 0.0
    return lam, n, dict(k=2 * math.pi / lam, sx=step, sy=sy, z0=z0, sz=sz, nx=n, ny=n)


@pytest.mark.parametrize("bx,by", SHAPES)
def test_truncation_bounds_cover_every_pair(bx, by):
    rng = np.random.default_rng(20261203 + 10 * bx + by)
    ratios = np.zeros(4)
    for _ in range(120):
        lam, n, P = random_plane(rng, bx, by)
        half = rng.uniform(2.0, 80.0)             # pillar extent (um)
        P["x0"] = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * P["sx"]
        P["y0"] = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * P["sy"]
        g = series_geom(P, bx, by)
        sn = obliquity(P, half, half)
        b = [truncation_bounds(g, sn, 1, 1)[0], truncation_bounds(g, sn, 2, 2)[0],
             truncation_bounds(g, sn, 1, 1)[1], truncation_bounds(g, sn, 2, 2)[1]]
        # |y| <= B / (4 R0) <= (a sn + c) / (4 z'): for D < 0 the dropped tail terms share
        # the leading term's sign and add at most a factor 1 + 4 |y| to it; 1e-14 is the
        # FP64 floor of the remainders' own evaluation.
        ymax = (g["a"] * sn + g["c"]) / (4 * g["z"])
        px = np.concatenate([rng.uniform(-half, half, 300), [half, -half, half, -half]])
        py = np.concatenate([rng.uniform(-half, half, 300), [half, half, -half, -half]])
        for ox, oy in [(0, 0), (n - bx, n - by), (0, n - by), (n - bx, 0),
                       (int(rng.integers(0, n - bx + 1)), int(rng.integers(0, n - by + 1)))]:
            rem = remainders(P, px, py, ox, oy, bx, by)
            for i in range(4):
                assert rem[i] <= b[i] * (1 + 4 * ymax) + 1e-14, (bx, by, i, P, rem[i], b[i])
                ratios[i] = max(ratios[i], rem[i] / b[i] if b[i] > 0 else 0.0)
    print(f"{bx}x{by}: worst remainder / bound: phase1 {ratios[0]:.2f} phase2 {ratios[1]:.2f} "
          f"amp1 {ratios[2]:.2f} amp2 {ratios[3]:.2f}")
    # the worst pairs come close to the bound (a 20% looser constant would show)
    assert ratios[0] > 0.8 and ratios[2] > 0.8


def test_headline_orders_and_bound_not_vacuous():
    """The headline plane (161^2 over +-5.4 um at z = 100 um, 91,809 pillars within
    +-52.85 um): 4 x 4 and 4 x 8 blocks run the first order (6 ops per pair) under the
    budget; the bound lies within a modest factor of the worst real remainder."""
    lam, f = 0.633, 100.0
    k = 2 * math.pi / lam
    W = 5.404873920333813
    step = 2 * W / 160
    P = dict(k=k, x0=-W, sx=step, y0=-W, sy=step, z0=f, sz=0.0, nx=161, ny=161)
    half = 151 * 0.35
    sn = obliquity(P, half, half)
    gx = np.linspace(-half, half, 61)
    X, Y = np.meshgrid(gx, gx)
    for bx, by in [(4, 4), (4, 8)]:
        g = series_geom(P, bx, by)
        assert pair_error(g, sn, 0) <= PAIR_BUDGET
        ph1, am1 = truncation_bounds(g, sn, 1, 1)
        rem = remainders(P, X.ravel(), Y.ravel(), 0, 0, bx, by)
        assert rem[0] <= ph1 and rem[2] <= am1
        assert ph1 < 20 * rem[0] and am1 < 20 * rem[2]


def order_obliquity(P, bx, by, order):
    """fp32c_order_obliquity: largest sn whose pair error meets the budget, by bisection."""
    g = series_geom(P, bx, by)
    if not pair_error(g, 0.0, order) <= PAIR_BUDGET:
        return -1.0
    if pair_error(g, 1.0, order) <= PAIR_BUDGET:
        return 1.0
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if pair_error(g, mid, order) <= PAIR_BUDGET:
            lo = mid
        else:
            hi = mid
    return lo


def test_order_limits_are_the_budget_crossings():
    """The obliquity limits the host hands the kernels are where each order's bound
    crosses the budget: every obliquity at or below a limit meets it, one just above does
    not (the bound grows with the obliquity), and P1A2 is admitted at least as far as P1A1."""
    rng = np.random.default_rng(3)
    crossings = 0
    for _ in range(300):
        bx, by = SHAPES[int(rng.integers(0, 4))]
        lam, n, P = random_plane(rng, bx, by)
        P["x0"], P["y0"] = -0.5 * (n - 1) * P["sx"], -0.5 * (n - 1) * P["sy"]
        g = series_geom(P, bx, by)
        o1, o2 = order_obliquity(P, bx, by, 0), order_obliquity(P, bx, by, 1)
        assert o2 >= o1
        for o, order in ((o1, 0), (o2, 1)):
            if 0 < o < 1:
                crossings += 1
                for s in np.linspace(0, o, 7):
                    assert pair_error(g, s, order) <= PAIR_BUDGET
                assert pair_error(g, min(1.0, o + 1e-9), order) > PAIR_BUDGET * (1 - 1e-12)
    assert crossings > 20


def test_host_admission_keeps_highest_order_within_budget():
    """Where the host lets a plane use bx x by blocks (P2A2 within the budget at any
    obliquity), the kernels may take P2A2 for any pillars, so its real remainders plus
    the rounding and sin/cos terms must meet the budget for pillars at every obliquity:
    here from 1e-3 to 1e3 times the plane distance, in all directions, against blocks at
    the plane's corners."""
    rng = np.random.default_rng(7)
    allowed, worst = 0, 0.0
    for _ in range(600):
        bx, by = SHAPES[int(rng.integers(0, 4))]
        lam = rng.uniform(0.4, 1.0)
        n = int(rng.integers(max(bx, by), 120))
        z0 = float(np.exp(rng.uniform(np.log(0.5), np.log(300.0))))
        pitch = lambda: lam * float(np.exp(rng.uniform(np.log(0.005), np.log(1.0))))  # noqa: E731
        if rng.random() < 0.7:
            step = pitch()
            sy, sz = step, 0.0
        else:
            step, sy, sz = pitch(), 0.0, pitch()
        P = dict(k=2 * math.pi / lam, x0=-0.5 * (n - 1) * step, sx=step,
                 y0=-0.5 * (n - 1) * sy, sy=sy, z0=z0, sz=sz, nx=n, ny=n)
        g = series_geom(P, bx, by)
        if not pair_error(g, 1.0, 2) <= PAIR_BUDGET:
            continue
        allowed += 1
        rho = np.geomspace(1e-3, 1e3, 200) * z0
        ang = rng.uniform(0, 2 * math.pi, rho.size)
        px, py = rho * np.cos(ang), rho * np.sin(ang)
        fixed = 2 * math.pi * 2.0 ** -24 * (2 + 6 * g["e"]) + 2.0 ** -21 + 2.5e-7
        for ox, oy in [(0, 0), (n - bx, n - by), (0, n - by), (n - bx, 0)]:
            rem = remainders(P, px, py, ox, oy, bx, by)
            worst = max(worst, (2 * math.pi * rem[1] + rem[3] + fixed) / PAIR_BUDGET)
    print(f"{allowed} planes admitted; worst P2A2 pair error {worst:.2f} of the budget")
    assert allowed > 150
    assert worst <= 1.0


# ---- RS-I planes ----------------------------------------------------------------------
# The RS-I pair t exp(ikR)/R (z/R)(1/R - ik) is, relative to its block-centre magnitude
# c1 2 pi zk rf and phase th0 - 1/4,
#     Z = (1 + 4y)^-1 (1 + i e (1 + 4y)^-1/2) exp(2 pi i (R' - R0')),   e = h / pi,
# and the kernels evaluate it as (1 + e^2/2) A(y) exp(2 pi i (e / 2 pi + P(D))) with
#     P1: P = D h (1 - h^2/pi^2) - D^2 h^3,   P2: P = x1 (1 - h^2/pi^2 - y (1 - 2y)),
#     A1: A = 1 - 4y,                         A2: A = 1 - 4y + 16y^2.
# rs_internal.h (fp32c_pair_error, rsi) bounds |Z - series| by its truncation terms; the
# rounding and sin/cos terms are separate.  Restated here and checked against the exact Z.


def rsi_truncation_bound(g, sn, order):
    """2 pi (phase) + amplitude truncation terms of fp32c_pair_error for an RS-I plane."""
    z2, z3 = g["z"] ** 2, g["z"] ** 3
    if order == 2:
        ph = 5 * series_peak(g["a"], g["c"], sn, 4, 3) / (128 * z3)
    else:
        ph = series_peak(g["a"], g["c"], sn, 3, 2) / (16 * z2)
    if order == 0:
        am = series_peak(g["a"], g["c"], sn, 2, 2) / z2
    else:
        am = series_peak(g["a"], g["c"], sn, 3, 3) / z3
    e = 1 / (2 * math.pi * g["z"])
    ymax = (g["a"] * sn + g["c"]) / (4 * g["z"])
    ph += (e ** 3 / 3 + 6 * e * ymax * ymax) * (1 + 4 * ymax) / (2 * math.pi)
    am += e ** 4 / 8 + 2.5 * e * e * ymax
    return 2 * math.pi * ph + am, ymax


def rsi_remainders(P, px, py, ox, oy, bx, by):
    """max over the block's pairs of |Z - series| for orders P1A1, P1A2, P2A2."""
    kt = P["k"] / (2 * math.pi)
    ic, jc = ox + 0.5 * (bx - 1), oy + 0.5 * (by - 1)
    FXc, FYc = kt * (P["x0"] + ic * P["sx"]), kt * (P["y0"] + jc * P["sy"])
    zc = kt * (P["z0"] + jc * P["sz"])
    dX0, dY0 = FXc - kt * px, FYc - kt * py
    R0 = np.sqrt(dX0 * dX0 + dY0 * dY0 + zc * zc)
    h = 0.5 / R0
    e = h / math.pi
    out = np.zeros(3)
    for r in range(by):
        ey = kt * P["sy"] * (r - 0.5 * (by - 1))
        zr = kt * (P["z0"] + (oy + r) * P["sz"])
        for c in range(bx):
            ex = kt * P["sx"] * (c - 0.5 * (bx - 1))
            D = ex * (2 * dX0 + ex) + ey * (2 * dY0 + ey) + (zr * zr - zc * zc)
            R = np.sqrt(R0 * R0 + D)
            x1, y = D * h, D * h * h
            dR = D / (R + R0)
            exact = (1 + 1j * e * R0 / R) * (R0 / R) ** 2 * np.exp(2j * math.pi * dR)
            w1 = D * h * (1 - h * h / math.pi ** 2) - D * D * h ** 3
            w2 = x1 * (1 - h * h / math.pi ** 2 - y * (1 - 2 * y))
            a1, a2 = 1 - 4 * y, 1 + y * (16 * y - 4)
            f = (1 + e * e / 2) * np.exp(1j * e)
            for i, (ph, am) in enumerate(((w1, a1), (w1, a2), (w2, a2))):
                ser = f * am * np.exp(2j * math.pi * ph)
                out[i] = max(out[i], float(np.abs(exact - ser).max()))
    return out


@pytest.mark.parametrize("bx,by", [(4, 4), (8, 8)])
def test_rsi_series_bound_covers_every_pair(bx, by):
    rng = np.random.default_rng(20261204 + 10 * bx + by)
    worst = np.zeros(3)
    for _ in range(120):
        lam, n, P = random_plane(rng, bx, by)
        half = rng.uniform(2.0, 80.0)
        P["x0"] = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * P["sx"]
        P["y0"] = rng.uniform(-1.5 * half, 1.5 * half) - 0.5 * (n - 1) * P["sy"]
        g = series_geom(P, bx, by)
        sn = obliquity(P, half, half)
        px = np.concatenate([rng.uniform(-half, half, 200), [half, -half, half, -half]])
        py = np.concatenate([rng.uniform(-half, half, 200), [half, half, -half, -half]])
        bounds = [rsi_truncation_bound(g, sn, o) for o in range(3)]
        for ox, oy in [(0, 0), (n - bx, n - by), (int(rng.integers(0, n - bx + 1)), 0)]:
            rem = rsi_remainders(P, px, py, ox, oy, bx, by)
            for o in range(3):
                b, ymax = bounds[o]
                assert rem[o] <= b * (1 + 4 * ymax) + 1e-13, (bx, by, o, P, rem[o], b)
                worst[o] = max(worst[o], rem[o] / b)
    print(f"RS-I {bx}x{by}: worst |Z - series| / bound per order {np.round(worst, 3)}")
    assert worst[0] > 0.3  # not vacuous


def test_rsi_bound_in_the_library_matches():
    """fp32c_pair_error(rsi) = the truncation terms above + rounding + sin/cos, and RS-I
    planes admit 4 x 4 blocks at the headline geometry."""
    lam, f = 0.633, 100.0
    k = 2 * math.pi / lam
    W = 5.404873920333813
    step = 2 * W / 160
    P = dict(k=k, x0=-W, sx=step, y0=-W, sy=step, z0=f, sz=0.0, nx=161, ny=161)
    g = series_geom(P, 4, 4)
    sn = obliquity(P, 151 * 0.35, 151 * 0.35)
    b, _ = rsi_truncation_bound(g, sn, 0)
    full = b + 2 * math.pi * 2.0 ** -24 * (3 + 6 * g["e"]) + 2.0 ** -20 + 2.5e-7
    assert full <= PAIR_BUDGET
    b2, _ = rsi_truncation_bound(g, 1.0, 2)
    assert b2 + 2 * math.pi * 2.0 ** -24 * (3 + 6 * g["e"]) + 2.0 ** -20 + 2.5e-7 <= PAIR_BUDGET
