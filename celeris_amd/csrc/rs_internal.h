// rs_internal.h — device-side descriptors shared by the HIP kernels (rs_kernels.hip)
// and the host planner (rs_api.cpp), and the FP32C per-pair error bound both sides
// evaluate.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace crs {

// Work-group size of every pair kernel: 4 wave64s, one per SIMD of a CU.
constexpr int kThreads = 256;
// Waves (wave64) per work-group.
constexpr int kWaves = kThreads / 64;
// Outputs per work-group of the FP32 / FP64 partial-sum reductions (each wave adds a
// quarter of the slots for kReduceCols outputs).
constexpr int kReduceCols = 64;
// FP32 / FP64 adjoint: rows per strip ("quad") of a plane.
constexpr int kQuad = 4;
// FP32 / FP64 forward: shortest pillar chunk of a split.
constexpr int kMinChunk = 64;

// Plane kernel selector (crs_kernel): RS-I, exp(ikR)/R (z/R)(1/R - ik).
constexpr int32_t kKernelRsi = 1;

// ---- FP32C ------------------------------------------------------------------------
// Pixels are handled in blocks of BX columns x BY rows (BX, BY in {4, 8}) that share
// one FP64 reference distance per pillar, or 1 x 1 where a plane is too close or too
// coarsely sampled for the block expansion ("exact" planes).  The forward gives each
// lane one block ("thread tile") and a work-group 64 tiles of one plane, its
// kFwdWaves waves splitting the pillars; the adjoint gives each thread one pillar and walks a plane in
// strips of kStrip rows, each strip in BX-column blocks.
constexpr int kFwdTiles = 64;
// FP32C forward work-group: waves that split the pillars of its 64 thread tiles.
constexpr int kFwdWaves = 4;
constexpr int kFwdThreads = 64 * kFwdWaves;
// FP32C adjoint strip height (rows); BY = 4 kernels take a strip as two halves.
constexpr int kStrip = 8;
// FP32C adjoint: columns per LDS tile of a strip (kAdjCols x kStrip pixel records).
constexpr int kAdjCols = 64;
// FP32C staged-pixel columns are padded to a multiple of this (the widest block).
constexpr int kColAlign = 8;
// FP32C forward: pillars one wave sums in float before the work-group folds its waves'
// sums into FP64 running sums (so a float sum never runs over more than this).
constexpr int kWaveFold = 1024;
// Pillars per FP32C extent record (one per LDS tile of staged pillars).
constexpr int kExtTile = kThreads;

// Bytes of the file-scope device array the allocation-free forward entry point
// (crs_forward_noalloc) carves its scratch from: staged pillars, partial sums and
// descriptor tables of one call.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

// Series orders of the FP32C pair evaluation: phase order / amplitude order.
//   kP1A1: R - R0 = x1 (1 - y),            R0/R = 1 - 2y             6 FP32 ops per pair
//   kP1A2: R - R0 = x1 (1 - y),            R0/R = 1 - 2y + 6y^2      7
//   kP2A2: R - R0 = x1 (1 - y + 2y^2),     R0/R = 1 - 2y + 6y^2      9
// with x1 = D / (2 R0), y = D / (4 R0^2), D = R^2 - R0^2 (the adjoint has kP1A1 and kP2A2).
constexpr int kP1A1 = 0;
constexpr int kP1A2 = 1;
constexpr int kP2A2 = 2;
// The adjoint's separable block form (8 x 8 blocks of planes the host marks `sep`).
constexpr int kSep = 3;

// Lower bound on |z| of every row z0 + j sz (0 <= j < ny) of a plane and of every
// pixel block's centre: the nearer end when all rows lie on one side of the lens
// plane, else 0 (a block of rows straddling z = 0 can have its centre there).  The
// series-order bounds need the nearest distance a pixel block can have to the pillars'
// plane; 0 sends the plane to the exact (1 x 1) path.
__host__ __device__ inline double plane_min_abs_z(double z0, double sz, int ny) {
    const double zl = z0 + (ny - 1) * sz;
    if ((z0 > 0.0 && zl > 0.0) || (z0 < 0.0 && zl < 0.0)) return fmin(fabs(z0), fabs(zl));
    return 0.0;
}

// FP32C staged pillar buffer (forward): positions (double2, npil), {|t|, arg t / 2 pi}
// (float2, nsets x npil) and, per kExtTile pillars, their extent {max |x|, max |y|}.
struct StagedLayout {
    size_t pos, ap, ext, bytes;
};
inline StagedLayout fp32c_staged_layout(int64_t npil, int64_t nsets) {
    StagedLayout L;
    L.pos = 0;
    L.ap = static_cast<size_t>(npil) * 16;
    L.ext = L.ap + ((static_cast<size_t>(npil * nsets) * 8 + 255) & ~size_t(255));
    L.bytes = L.ext + static_cast<size_t>((npil + kExtTile - 1) / kExtTile) * 16;
    return L;
}

// One plane as the kernels see it (host computes the derived fields).
struct DevPlane {
    double k;       // rad/um
    double kturn;   // k / (2 pi), turns per um
    double x0, sx;  // fx = x0 + i*sx
    double y0, sy;  // fy = y0 + j*sy
    double z0, sz;  // z  = z0 + j*sz
    double zref;    // kturn * z0: FP32C phase reference, restored as exp(2 pi i zref)
    int32_t nx, ny;
    int32_t tx;     // FP32C forward: whole BX-wide blocks across a row (0: exact plane)
    int32_t ty;     // FP32C forward: whole BY-high blocks down a column (0: exact plane)
    int32_t tset;
    int32_t nslots; // FP32 / FP64 forward: pillar chunks the sum is split into
    int64_t offset; // first output element of the plane
    int64_t part;   // FP32 / FP64 forward: first partial element of slot 0
    int64_t sbase;  // FP32C adjoint: first staged pixel record of the plane
    int32_t nxp;    // FP32C adjoint: nx rounded up to kColAlign; FP32/FP64: unused
    int32_t nquads; // FP32 / FP64 adjoint: ceil(ny / 4); FP32C: strips ceil(ny / kStrip)
    int32_t bx, by; // FP32C forward: block shape of the plane's launch (1 x 1 if exact)
    // FP32C: steepest pair obliquity (sin of the angle from the z axis) at which the
    // series orders kP1A1 / kP1A2 meet the pair budget for the plane's block shape (-1:
    // never); beyond obl2, kP2A2, which the host admitted at any obliquity.
    double obl1, obl2;
    int32_t kernel;  // crs_kernel: 0 exp(ikR)/R, 1 RS-I (FP32C: pair by pair, 1 x 1 blocks)
    // FP32C adjoint: 1 if the plane's blocks take the separable form (its staged pixel
    // records then hold u (k/2pi) exp(-2 pi i z'_ref) as {Re, Im, -Re}).
    int32_t sep;
};

// ---- FP32 / FP64 work decomposition (split sums, uniform slots) --------------------

// Forward work-group: pixels [tile0, tile0 + kThreads) of `plane`, summed over
// pillars [q0, q1); results go to partial[part + pixel].
struct FwdBlock {
    int32_t plane, tile0, q0, q1;
    int64_t part;
    int32_t shape;
    int32_t pad;
};

// A run of consecutive elements of one plane: kThreads (staging) or kReduceCols
// (forward reduction).
struct PixBlock {
    int32_t plane, pix0;
};

// Four consecutive rows j0..j0+3 of one plane (rows >= ny are padding).
struct AdjQuad {
    int32_t plane, j0;
};

// Adjoint work-group: pillars q0 + tid of transmission set `tset`, summed over quads
// [u0, u1) of the quad list; results go to partial[part + q].
struct AdjBlock {
    int32_t tset, q0, u0, u1;
    int64_t part;
};

// ---- FP32C work decomposition (stream-K: a cost-cut work line) ---------------------
// Forward: a "tile" is 64 thread tiles of one class of one plane (class 0 = BX x BY
// blocks, 1 = 1 x BY, 2 = BX x 1, 3 = 1 x 1).  A "segment" is one tile against a pillar
// range; each segment writes its own partial slot (64 x BX*BY float pairs).  Work-group
// w runs segments [wg_seg[w], wg_seg[w+1]), a contiguous run of the tile-major
// (tile, pillar) work line cut into pieces of falling cost.
struct FwdTile {
    int32_t plane, shape, tile0;
    int32_t nel;    // pixels per lane of this class (BX*BY, BY, BX or 1)
    int32_t seg0, nseg;  // its segments = partial slots seg0 .. seg0+nseg-1
};
struct FwdSeg {
    int32_t tile, q0, q1, pad;
};
// Forward reduction work-group: elements [el0, el0 + 64) (e*64 + lane) of a tile.
struct FwdRed {
    int32_t tile, el0;
};

// Adjoint: a segment is one group of kThreads pillars (q0 ...) of set `tset` against a
// run of its set's strip list, from column c0 of strip u0 to column c1 (exclusive) of
// strip u1 (columns in blocks of the launch's BX); it writes its own partial slot
// (kThreads float pairs).  Each (set, group) lists its slots (CSR over all launches).
struct AdjStrip {
    int32_t plane, j0;
};
struct AdjSeg {
    int32_t tset, q0, u0, c0;
    int32_t u1, c1, slot, pad;
};

// FP32C pixel tiles.  A plane blocked KBX x KBY has four classes of thread tiles, so
// no lane computes a pixel outside the plane: class 0 = KBX x KBY blocks over the
// largest multiple-of-block corner (tx x ty of them), 1 = 1 x KBY tiles of the
// rx = nx - KBX tx leftover columns, 2 = KBX x 1 tiles of the ry = ny - KBY ty leftover
// rows, 3 = 1 x 1 tiles of the leftover corner.  An exact plane (tx = ty = 0) is all
// class 3.
__host__ __device__ inline int tile_count(const DevPlane& P, int shape, int kbx, int kby) {
    const int rx = P.nx - kbx * P.tx, ry = P.ny - kby * P.ty;
    switch (shape) {
        case 0: return P.tx * P.ty;
        case 1: return rx * P.ty;
        case 2: return P.tx * ry;
        default: return rx * ry;
    }
}

__host__ __device__ inline void tile_origin(const DevPlane& P, int shape, int t, int kbx, int kby,
                                            int& ox, int& oy) {
    const int rx = P.nx - kbx * P.tx;
    switch (shape) {
        case 0: ox = kbx * (t % P.tx); oy = kby * (t / P.tx); break;
        case 1: ox = kbx * P.tx + t % rx; oy = kby * (t / rx); break;
        case 2: ox = kbx * (t % P.tx); oy = kby * P.ty + t / P.tx; break;
        default: ox = kbx * P.tx + t % rx; oy = kby * P.ty + t / rx; break;
    }
}

// Device arguments of one FP32C forward / adjoint pair-kernel launch.
struct Fp32cFwdArgs {
    const double2* pos;
    const float2* ap;
    const double2* ext;
    int npil;
    const DevPlane* planes;
    const FwdTile* tiles;
    const FwdSeg* segs;
    const int32_t* wg_seg;  // nwg + 1 entries
    int nwg;
    float2* partial;
    int64_t slot_elems;
};
struct Fp32cAdjArgs {
    const double* px;
    const double* py;
    int npil;
    const DevPlane* planes;
    const AdjStrip* strips;
    const AdjSeg* segs;
    const int32_t* wg_seg;  // nwg + 1 entries
    int nwg;
    const float4* pix;
    float2* partial;
};

// ---- FP32C per-pair error bound ------------------------------------------------------
// A pair's computed term differs from t_q exp(i k R) / R by at most eps |t_q| / R_pq,
// with eps the sum of:
//   * the dropped series terms (turns of phase x 2 pi, plus relative amplitude):
//       order 1 phase 2 x1 y^2 <= B^3 / (16 R0^2),  order 2 phase 5 x1 y^3 <= 5 B^4 / (128 R0^3)
//       order 1 amplitude 6 y^2 <= 3 B^2 / (8 R0^2),  order 2 amplitude 20 y^3 <= 5 B^3 / (16 R0^3)
//     where |D| / R0 <= B(s) = a s + c (a = 2 e', c = 2 e_z' + (e'^2 + e_z'^2) / z', e' the
//     block's half diagonal, e_z' its half depth, z' the nearest depth, all in turns) for
//     a pair of obliquity s = rho / R0, and R0 >= z' / sqrt(1 - s^2);
//   * float rounding of the phase, <= 2^-24 (2 + 6 (e' + e_z')) turns (the final sum at
//     magnitude < 1 + |x1|, the staged arg t and th0, D's terms and the pixel offsets);
//   * float rounding of the amplitude (v_rsq_f32 1.6 ulp, the staged |t|, two products),
//     <= 2^-21;
//   * v_sin_f32 / v_cos_f32, <= 1.25e-7 absolute each on gfx950 over [-2, 2] turns
//     (measured), taken as 2.5e-7.
// A series order is admitted where eps <= kPairBudget.  The float accumulation of the
// admitted terms (at most kWaveFold per float sum, then FP64) has its own error, ~1e-6
// of sum |t|/R at worst for a coherent pixel, so the total stays below 1e-5 of it.
//
// RS-I planes (G = exp(ikR)/R (z/R)(1/R - ik)): in turns the pair is
//     t (k/2pi)^2 z' exp(2 pi i R') (1/R' - 2 pi i) / R'^2
//   = c1 2 pi zk rf (1 + 4y)^-1 sqrt(1 + e'^2) exp(2 pi i (R' - 1/4 + atan(e') / 2 pi)),
// e = 1/(2 pi R0') = h / pi, e' = e (1 + 4y)^-1/2, zk = (k/2pi) z' of the pixel's row;
// so the same pair loop serves with th0 - 1/4 + e / 2 pi, the phase's D-linear
// coefficient h (1 - h^2/pi^2) (the -2 e y of atan(e')), the amplitude series of
// (1 + 4y)^-1 (1 - 4y, then + 16y^2) times (1 + e^2/2) and the row's z'_r / z'_c.  The
// bound's amplitude terms become 16 y^2 <= B^2 / R0^2 and 64 y^3 <= B^3 / R0^3; the
// dropped e terms add e^3/3 + 6 e y^2 radians of phase and e^4/8 + 2.5 e^2 |y| of
// amplitude (e <= 1/(2 pi z')); the extra products double the amplitude rounding, and
// th0's shift adds 2^-24 of phase rounding.
constexpr double kPairBudget = 8e-6;
constexpr double kTwoPiD = 6.28318530717958647693;

struct SeriesGeom {
    double a, c, z, e;  // B(s) = a s + c; z: nearest depth (turns); e: e' + e_z'
};

// Geometry of bx x by blocks of plane P (lengths in turns).
__host__ __device__ inline SeriesGeom series_geom(const DevPlane& P, int bx, int by) {
    const double hx = 0.5 * (bx - 1) * fabs(P.sx) * P.kturn;
    const double hy = 0.5 * (by - 1) * fabs(P.sy) * P.kturn;
    const double hz = 0.5 * (by - 1) * fabs(P.sz) * P.kturn;
    const double e = sqrt(hx * hx + hy * hy);
    SeriesGeom g;
    g.z = P.kturn * plane_min_abs_z(P.z0, P.sz, P.ny);
    g.a = 2.0 * e;
    g.c = 2.0 * hz + (e * e + hz * hz) / g.z;
    g.e = e + hz;
    return g;
}

// max over 0 <= s <= sn of (a s + c)^k (1 - s^2)^m (m = 1 or 3/2, given as 2m): the
// function rises to the positive root of its derivative, ((k + 2m) a s^2 + 2m c s - k a
// = 0), and falls after it.
__host__ __device__ inline double series_peak(double a, double c, double sn, int k, int twom) {
    double s = 0.0;
    if (a > 0.0) {
        const double A = (k + twom) * a, Bq = twom * c, C = -k * a;
        s = (-Bq + sqrt(Bq * Bq - 4.0 * A * C)) / (2.0 * A);
    }
    s = fmin(fmax(s, 0.0), sn);
    const double B = a * s + c, om = 1.0 - s * s;
    const double f = twom == 2 ? om : om * sqrt(om);
    return pow(B, k) * f;
}

// eps (fraction of the pair's magnitude) of series order `order` for pairs of obliquity
// <= sn; rsi: the plane's pairs take the RS-I kernel.
__host__ __device__ inline double fp32c_pair_error(const SeriesGeom& g, double sn, int order,
                                                   bool rsi = false) {
    if (!(g.z > 0.0)) return INFINITY;
    const double z2 = g.z * g.z, z3 = z2 * g.z;
    double phase, amp;
    if (order == kP2A2) {
        phase = 5.0 * series_peak(g.a, g.c, sn, 4, 3) / (128.0 * z3);
    } else {
        phase = series_peak(g.a, g.c, sn, 3, 2) / (16.0 * z2);
    }
    const double amp_c1 = rsi ? 1.0 : 3.0 / 8.0, amp_c2 = rsi ? 1.0 : 5.0 / 16.0;
    if (order == kP1A1) {
        amp = amp_c1 * series_peak(g.a, g.c, sn, 2, 2) / z2;
    } else {
        amp = amp_c2 * series_peak(g.a, g.c, sn, 3, 3) / z3;
    }
    double round_phase = 5.9604644775390625e-8 * (2.0 + 6.0 * g.e);
    double round_amp = 4.76837158203125e-7;
    if (rsi) {
        const double e = 1.0 / (kTwoPiD * g.z);             // >= 1 / (2 pi R0')
        const double ymax = (g.a * sn + g.c) / (4.0 * g.z);  // >= |y|
        phase += (e * e * e / 3.0 + 6.0 * e * ymax * ymax) * (1.0 + 4.0 * ymax) / kTwoPiD;
        amp += e * e * e * e / 8.0 + 2.5 * e * e * ymax;
        round_phase += 5.9604644775390625e-8;
        round_amp *= 2.0;
    }
    const double trig = 2.5e-7;
    return kTwoPiD * (phase + round_phase) + amp + round_amp + trig;
}

// Steepest pillar-to-pixel obliquity of plane P against pillars within |x| <= ex,
// |y| <= ey (um): sin of the angle from the z axis.
__host__ __device__ inline double plane_obliquity(const DevPlane& P, double ex, double ey) {
    const double xl = P.x0 + (P.nx - 1) * P.sx, yl = P.y0 + (P.ny - 1) * P.sy;
    const double rx = fmax(fabs(P.x0), fabs(xl)) + ex;
    const double ry = fmax(fabs(P.y0), fabs(yl)) + ey;
    const double zmin = plane_min_abs_z(P.z0, P.sz, P.ny);
    const double rho2 = rx * rx + ry * ry;
    if (!(rho2 > 0.0)) return 0.0;
    return sqrt(rho2 / (rho2 + zmin * zmin));
}

// Largest obliquity sn in [0, 1] at which series order `order` of bx x by blocks of P
// meets the pair budget (the error grows with sn), or -1 if it never does.
inline double fp32c_order_obliquity(const DevPlane& P, int bx, int by, int order) {
    const SeriesGeom g = series_geom(P, bx, by);
    const bool rsi = P.kernel == kKernelRsi;
    if (!(fp32c_pair_error(g, 0.0, order, rsi) <= kPairBudget)) return -1.0;
    if (fp32c_pair_error(g, 1.0, order, rsi) <= kPairBudget) return 1.0;
    double lo = 0.0, hi = 1.0;
    for (int i = 0; i < 60; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (fp32c_pair_error(g, mid, order, rsi) <= kPairBudget) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Cheapest series order admitted for P's blocks against pillars within the extent
// (ex, ey), from the plane's precomputed obliquity limits.
__host__ __device__ inline int fp32c_series_order(const DevPlane& P, double ex, double ey) {
    const double sn = plane_obliquity(P, ex, ey);
    if (sn <= P.obl1) return kP1A1;
    if (sn <= P.obl2) return kP1A2;
    return kP2A2;
}

// ---- Separable block evaluation ----------------------------------------------------------
// For a pillar and a block with centre distance R0 (turns), pixel (c, r) has
// R^2 = R0^2 + a_c + b_r, and with F(D) = (R0 / R(D)) exp(2 pi i (R(D) - R0)),
//     t exp(2 pi i R') / R = c1 exp(2 pi i th0) F(a_c + b_r) = U_c V_r K(a_c, b_r),
//     U_c = c1 exp(2 pi i th0) F(a_c),  V_r = F(b_r),  K(a, b) = F(a + b) / (F(a) F(b)).
// log K = sum_{j>=1} a^j l_j(b), l_j(b) = (psi^(j)(b) - psi^(j)(0)) / j!, psi = log F; with
// h = 1/(2 R0) and y = b h^2 the leading terms are
//     l1 = b (8 h^4 (1 - 4y + ...) - 4 pi i h^3 (1 - 3y + ...)),   l2 = y (-32 h^4 - ... + 12 pi i h^3 ...).
// The separable form keeps rank 2: pair (c, r) = U_c (V_r + a_c W_r), W_r = V_r l1(b_r),
// with F to second order (phase D h (1 - y + 2y^2), amplitude 1 - 2y + 6y^2, y = D h^2) and
// l1 to first order in y.  Per pair that is 6 FP32 FMAs and no transcendental; per pillar
// and block one sin / cos per column and per row.
//
// Pair error, as a fraction of |t| / R, for a pillar at obliquity s and azimuth phi from
// the block centre (|dX0| = rho |cos phi|, |dY0| = rho |sin phi|, rho = R0 s, block depth
// zc = R0 sqrt(1 - s^2), R0 >= z' / sqrt(1 - s^2)): with |a| <= A = 2 e_x |dX0| + e_x^2,
// |b| <= B = 2 e_y |dY0| + e_y^2 + 2 zc e_z + e_z^2 (e the block's half extents in turns),
// y_a = A h^2, y_b = B h^2, the series in y give
//     |l1| <= y_b (8 h^2 + 4 pi h) / (1 - 4 y_b),   |l2| <= y_b (32 h^4 + 12 pi h^3) / (1 - 6 y_b),
//     |l3| <= y_b (128 h^6 + 40 pi h^5) / (1 - 8 y_b),
// so |log K| <= ell = t1 + t2 + t3 (t1 = A|l1|, t2 = A^2|l2|, t3 = A^3|l3|(1 + 16 y_a), the
// tail j >= 4 included) and eps is the sum of
//   * the dropped rank terms, |K - 1 - a l1| <= ell^2 e^ell / 2 + t2 + t3;
//   * the series remainders: F, phase 5 y^3 |D h| and amplitude 20 y^3 per side over
//     1 - 4y; l1, A (128 h^2 + 40 pi h) y_b^3 over 1 - 4 y_b;
//   * float rounding: phase 2^-24 (3 + 8 (A + B) h) turns (th0 and the two phase sums;
//     a and b from rounded offsets, and h, each a few ulp of |a h| or |b h|); amplitude
//     1.2 2^-21 (c1, the two series amplitudes, the two phasor scalings, the products);
//   * the hardware sin / cos of the two factors, 2.5e-7.
__host__ __device__ inline double sep_pair_error(const DevPlane& P, double s, double cphi,
                                                 double sphi, int bx, int by) {
    const double kt = P.kturn;
    const double ex = 0.5 * (bx - 1) * fabs(P.sx) * kt;
    const double ey = 0.5 * (by - 1) * fabs(P.sy) * kt;
    const double ez = 0.5 * (by - 1) * fabs(P.sz) * kt;
    const double z = kt * plane_min_abs_z(P.z0, P.sz, P.ny);
    if (!(z > 0.0) || !(s >= 0.0 && s < 1.0)) return INFINITY;
    const double cz = sqrt(1.0 - s * s), R0 = z / cz, h = 0.5 / R0, rho = R0 * s;
    const double A = 2.0 * ex * rho * cphi + ex * ex;
    const double B = 2.0 * ey * rho * sphi + ey * ey + 2.0 * R0 * cz * ez + ez * ez;
    const double h2 = h * h, h3 = h2 * h, ya = A * h2, yb = B * h2;
    if (!(ya < 0.02 && yb < 0.02)) return INFINITY;
    const double pi = 0.5 * kTwoPiD;
    const double l1 = yb * (8.0 * h2 + 4.0 * pi * h) / (1.0 - 4.0 * yb);
    const double l2 = yb * (32.0 * h2 * h2 + 12.0 * pi * h3) / (1.0 - 6.0 * yb);
    const double l3 = yb * (128.0 * h3 * h3 + 40.0 * pi * h3 * h2) / (1.0 - 8.0 * yb);
    const double t1 = A * l1, t2 = A * A * l2, t3 = A * A * A * l3 * (1.0 + 16.0 * ya);
    const double ell = t1 + t2 + t3;
    const double rank = ell * ell * exp(ell) / 2.0 + t2 + t3;
    const double ya3 = ya * ya * ya, yb3 = yb * yb * yb;
    const double dl1 = A * (128.0 * h2 + 40.0 * pi * h) * yb3 / (1.0 - 4.0 * yb);
    const double side_phase = 5.0 * (A * h * ya3 / (1.0 - 4.0 * ya) + B * h * yb3 / (1.0 - 4.0 * yb));
    const double side_amp = 20.0 * (ya3 / (1.0 - 4.0 * ya) + yb3 / (1.0 - 4.0 * yb));
    const double round_phase = 5.9604644775390625e-8 * (3.0 + 8.0 * (A + B) * h);
    return kTwoPiD * (side_phase + round_phase) + side_amp + 1.1 * (rank + dl1) +
           1.2 * 4.76837158203125e-7 + 2.5e-7;
}

// Largest eps of bx x by separable blocks of P over pillars at every obliquity and azimuth
// (a grid of 256 obliquities x 33 azimuths in [0, pi/2]; eps varies slowly with both).
inline double sep_plane_error(const DevPlane& P, int bx, int by) {
    double worst = 0.0;
    for (int j = 0; j <= 32; ++j) {
        const double phi = 0.25 * kTwoPiD * j / 32.0, cphi = cos(phi), sphi = sin(phi);
        for (int i = 0; i < 256; ++i) {
            const double e = sep_pair_error(P, 0.999 * i / 255.0, cphi, sphi, bx, by);
            if (!(e <= worst)) worst = e;
        }
    }
    return worst;
}

// Whether bx x by blocks of P take the separable form (the exp(ikR)/R kernel only).
inline bool sep_plane_admitted(const DevPlane& P, int bx, int by) {
    return P.kernel == 0 && sep_plane_error(P, bx, by) <= kPairBudget;
}

}  // namespace crs
