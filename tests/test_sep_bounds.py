"""The separable block form's pair-error bound, checked against exact errors (CPU).

The FP32C adjoint takes a pair (c, r) of an 8 x 8 block (pixel offsets a_c + b_r on
R^2 = R0^2 + a + b) as U_c (V_r + a_c W_r): with F(D) = (R0/R(D)) exp(2 pi i (R(D) - R0)),
h = 1/(2 R0) and y = D h^2,
    F2(D) = (1 - 2y + 6y^2) exp(2 pi i D h (1 - y + 2y^2)),
    l1(b) = b (8 h^4 (1 - 4y) - 4 pi i h^3 (1 - 3y)),
    F(a + b) ~ F2(a) F2(b) (1 + a l1(b))
(celeris_amd/csrc/rs_internal.h: sep_pair_error, sep_plane_error).  The bound is restated
here line by line; its truncation part (rank, series of l1 and of F, over the obliquity s
and azimuth phi of the pillar from the block centre, with R0 = z' / sqrt(1 - s^2)) is
checked against |F(a + b) - F2(a) F2(b) (1 + a l1(b))| formed in FP64 without
cancellation (R - R0 = D / (R + R0)) over random planes (psf windows and zx slices,
pitches and depths around the admission boundary), blocks across each plane and pillars
at random positions: no pair may exceed its bound, the worst pairs come within a factor of
a few of it (the bound is not vacuous), and dropping the rank-2 term (a l1) leaves errors
the bound does not cover (the check has teeth).  The float rounding and sin / cos terms
of the bound are the kernels' own and are checked on the GPU (tests/test_gpu_sep.py).
"""
import math

import numpy as np

PAIR_BUDGET = 8e-6
TWO_PI = 2.0 * math.pi
FP64_FLOOR = 1e-14


def plane_min_abs_z(z0, sz, ny):
    zl = z0 + (ny - 1) * sz
    if (z0 > 0 and zl > 0) or (z0 < 0 and zl < 0):
        return min(abs(z0), abs(zl))
    return 0.0


def sep_terms(P, s, cphi, sphi, bx, by):
    """(truncation part, rounding and hardware part) of sep_pair_error."""
    kt = P["k"] / TWO_PI
    ex = 0.5 * (bx - 1) * abs(P["sx"]) * kt
    ey = 0.5 * (by - 1) * abs(P["sy"]) * kt
    ez = 0.5 * (by - 1) * abs(P["sz"]) * kt
    z = kt * plane_min_abs_z(P["z0"], P["sz"], P["ny"])
    if not (z > 0 and 0 <= s < 1):
        return math.inf, math.inf
    cz = math.sqrt(1 - s * s)
    R0 = z / cz
    h = 0.5 / R0
    rho = R0 * s
    A = 2 * ex * rho * cphi + ex * ex
    B = 2 * ey * rho * sphi + ey * ey + 2 * R0 * cz * ez + ez * ez
    h2, h3 = h * h, h * h * h
    ya, yb = A * h2, B * h2
    if not (ya < 0.02 and yb < 0.02):
        return math.inf, math.inf
    pi = math.pi
    l1 = yb * (8 * h2 + 4 * pi * h) / (1 - 4 * yb)
    l2 = yb * (32 * h2 * h2 + 12 * pi * h3) / (1 - 6 * yb)
    l3 = yb * (128 * h3 * h3 + 40 * pi * h3 * h2) / (1 - 8 * yb)
    t1, t2, t3 = A * l1, A * A * l2, A ** 3 * l3 * (1 + 16 * ya)
    ell = t1 + t2 + t3
    rank = ell * ell * math.exp(ell) / 2 + t2 + t3
    dl1 = A * (128 * h2 + 40 * pi * h) * yb ** 3 / (1 - 4 * yb)
    side_phase = 5 * (A * h * ya ** 3 / (1 - 4 * ya) + B * h * yb ** 3 / (1 - 4 * yb))
    side_amp = 20 * (ya ** 3 / (1 - 4 * ya) + yb ** 3 / (1 - 4 * yb))
    trunc = TWO_PI * side_phase + side_amp + 1.1 * (rank + dl1)
    other = TWO_PI * 2.0 ** -24 * (3 + 8 * (A + B) * h) + 1.2 * 2.0 ** -21 + 2.5e-7
    return trunc, other


def sep_pair_error(P, s, cphi, sphi, bx, by):
    t, o = sep_terms(P, s, cphi, sphi, bx, by)
    return t + o


def sep_plane_error(P, bx, by):
    worst = 0.0
    for j in range(33):
        phi = 0.25 * TWO_PI * j / 32
        for i in range(256):
            e = sep_pair_error(P, 0.999 * i / 255, math.cos(phi), math.sin(phi), bx, by)
            if not e <= worst:
                worst = e
    return worst


def F_exact(R0, D):
    R = np.sqrt(R0 * R0 + D)
    return (R0 / R) * np.exp(2j * np.pi * (D / (R + R0)))


def F2(h, D):
    y = D * h * h
    return (1 - 2 * y + 6 * y * y) * np.exp(2j * np.pi * D * h * (1 - y + 2 * y * y))


def lam1(h, b):
    y = b * h * h
    return b * (8 * h ** 4 * (1 - 4 * y) - 4j * np.pi * h ** 3 * (1 - 3 * y))


def random_plane(rng, lam):
    k = TWO_PI / lam
    pitch = lam * float(np.exp(rng.uniform(np.log(0.03), np.log(0.3))))
    z = float(np.exp(rng.uniform(np.log(20.0), np.log(400.0))))
    nx, ny = int(rng.integers(16, 80)), int(rng.integers(16, 80))
    if rng.random() < 0.75:
        return dict(k=k, x0=float(rng.uniform(-20, 20)), sx=pitch, y0=float(rng.uniform(-20, 20)),
                    sy=pitch, z0=z, sz=0.0, nx=nx, ny=ny)
    return dict(k=k, x0=float(rng.uniform(-20, 20)), sx=pitch, y0=0.0, sy=0.0, z0=z,
                sz=pitch * float(np.exp(rng.uniform(np.log(0.3), np.log(2.0)))), nx=nx, ny=ny)


def pair_errors(P, bx, by, px, py, rank2=True):
    """Exact truncation errors and the bound's truncation part, every pair of every
    bx x by block of P (block origins on the plane's bx x by lattice) for pillars px, py."""
    kt = P["k"] / TWO_PI
    errs, bnds = [], []
    for oy in range(0, P["ny"] - by + 1, by):
        for ox in range(0, P["nx"] - bx + 1, bx):
            ic, jc = ox + (bx - 1) / 2, oy + (by - 1) / 2
            FX = kt * (P["x0"] + ic * P["sx"])
            FY = kt * (P["y0"] + jc * P["sy"])
            zc = kt * (P["z0"] + jc * P["sz"])
            ex = (np.arange(bx) - (bx - 1) / 2) * P["sx"] * kt
            ey = (np.arange(by) - (by - 1) / 2) * P["sy"] * kt
            ez = (np.arange(by) - (by - 1) / 2) * P["sz"] * kt
            dX = FX - kt * px
            dY = FY - kt * py
            R0 = np.sqrt(dX * dX + dY * dY + zc * zc)
            h = 0.5 / R0
            a = 2 * ex[None, :] * dX[:, None] + ex[None, :] ** 2                     # [q, c]
            b = (2 * ey[None, :] * dY[:, None] + ey[None, :] ** 2
                 + 2 * zc * ez[None, :] + ez[None, :] ** 2)                           # [q, r]
            A3, B3 = a[:, :, None], b[:, None, :]
            H = h[:, None, None]
            approx = F2(H, A3) * F2(H, B3)
            if rank2:
                approx = approx * (1 + A3 * lam1(H, B3))
            e = np.abs(approx - F_exact(R0[:, None, None], A3 + B3)).max(axis=(1, 2))
            rho = np.hypot(dX, dY)
            s = rho / R0
            for q in range(px.size):
                cphi = abs(dX[q]) / rho[q] if rho[q] > 0 else 1.0
                sphi = abs(dY[q]) / rho[q] if rho[q] > 0 else 0.0
                t, _ = sep_terms(P, float(s[q]), cphi, sphi, bx, by)
                errs.append(float(e[q]))
                bnds.append(t)
    return np.array(errs), np.array(bnds)


def test_sep_bound_covers_exact_errors():
    rng = np.random.default_rng(2024)
    worst_ratio, tested = 0.0, 0
    for _ in range(40):
        lam = float(rng.uniform(0.45, 0.9))
        P = random_plane(rng, lam)
        if not sep_plane_error(P, 8, 8) <= PAIR_BUDGET:
            continue  # the kernels do not take this plane in the separable form
        half = float(rng.uniform(5, 80))
        px = rng.uniform(-half, half, 24)
        py = rng.uniform(-half, half, 24)
        e, bnd = pair_errors(P, 8, 8, px, py)
        # FP64 forms both sides to a few ulps of values near 1 (~1e-15): the floor.
        assert np.all(e <= bnd + FP64_FLOOR), (P, float(np.max(e - bnd)))
        big = bnd > 1e-12
        if np.any(big):
            worst_ratio = max(worst_ratio, float(np.max(e[big] / bnd[big])))
        tested += 1
    print(f"admitted planes {tested}, worst exact / bound {worst_ratio:.3f}")
    assert tested >= 8
    assert worst_ratio >= 0.05  # the bound is not vacuous


def test_sep_rank2_term_matters():
    """Without the rank-2 term a l1 the truncation exceeds the bound on admitted planes:
    a kernel that dropped it would be caught."""
    rng = np.random.default_rng(7)
    broke = 0
    for _ in range(20):
        P = random_plane(rng, float(rng.uniform(0.45, 0.9)))
        if not sep_plane_error(P, 8, 8) <= PAIR_BUDGET:
            continue
        px = rng.uniform(-60, 60, 16)
        py = rng.uniform(-60, 60, 16)
        e, bnd = pair_errors(P, 8, 8, px, py, rank2=False)
        broke += int(np.any(e > bnd))
    assert broke >= 3


def test_sep_headline_admission():
    """The headline lens's 161^2, 321^2 and 1024^2 windows (+-4 lambda/NA at f = 100 um,
    lambda 0.633) take the separable 8 x 8 adjoint; the 81^2 window over +-5 um at
    f = 50 um, lambda 0.532 (4,096 pillars) does not; the worst pair bound is the
    admission's."""
    lam, na = 0.633, 0.4681
    for n, ok in ((161, True), (321, True), (1024, True)):
        W = 4 * lam / na
        s = 2 * W / (n - 1)
        P = dict(k=TWO_PI / lam, x0=-W, sx=s, y0=-W, sy=s, z0=100.0, sz=0.0, nx=n, ny=n)
        assert (sep_plane_error(P, 8, 8) <= PAIR_BUDGET) == ok
    P = dict(k=TWO_PI / 0.532, x0=-5.0, sx=0.125, y0=-5.0, sy=0.125, z0=50.0, sz=0.0, nx=81, ny=81)
    assert not sep_plane_error(P, 8, 8) <= PAIR_BUDGET
