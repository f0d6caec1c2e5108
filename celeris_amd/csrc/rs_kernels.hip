// rs_kernels.hip — gfx950 kernels for the Rayleigh–Sommerfeld far-field sum
//
//     E_p = sum_q t_q exp(i k R_pq) / R_pq,   R_pq = |(fx_p - x_q, fy_p - y_q, z_p)|
//
// (reference: src/celeris/cuda/propagate.cu:36-84 psf_kernel, :86-130 zx_kernel, and
// the FP64 CPU loop src/celeris/analysis/focal.cpp:229-233) and its adjoint
// g_q = sum_p conj(G_pq) u_p (no reference counterpart).
//
// Three arithmetic modes share one work decomposition: a work-group of 256 threads
// (4 wave64s) owns a set of pixels (forward) or pillars (adjoint) and streams the
// other side through LDS in tiles that every lane reads by broadcast.  Sums are split
// over pillar chunks (forward) or pixel quads (adjoint) so that small focal grids
// still launch thousands of work-groups across the 256 CUs; each split writes a
// partial (float64, or for FP32C a float pair rounded once from its FP64 combine) and a
// second kernel adds the partials in FP64 in a fixed order, so results are bitwise
// repeatable run to run.
//
// FP32C ("FP32-compensated").  Lengths are measured in turns (x' = k x / 2 pi), so a
// distance is directly a phase.  Pixels are taken in 4 x 4 blocks.  For each pillar q
// and block b the distance R0 from q to the block centre and the phase
// frac(R0 - z'_ref + arg t_q / 2 pi) are formed in FP64 (one v_rsq_f32 plus one FP64
// Newton step).  Every pixel p of the block sits at a small, exactly known offset e
// from the centre, so
//     R^2 = R0^2 + D,   D = e_x (2 dx0 + e_x) + e_y (2 dy0 + e_y) + (z_p'^2 - z_c'^2)
// where D is tiny next to R0^2 and is formed in float without cancellation (each
// term is one fma against per-block constants).  Then, with h = 1/(2 R0),
// x1 = D h = a[c] + b[r] (separable per column / row) and y = x1 h = D / (4 R0^2):
//     R - R0 = x1 (1 - y + 2 y^2 - ...)          (first order keeps x1 w, w = 1 - y)
//     R0 / R = 1 - 2 y + 6 y^2 - ...              (first order keeps 2 w - 1)
// and the pair is
//     E += |t| (k/2 pi) / R0 * (R0/R) * (cos 2 pi th, sin 2 pi th),  th = th0 + (R - R0)
// with v_cos_f32 / v_sin_f32 (which take turns): 6 VALU ops and 2 transcendentals per
// pair.  No float ever holds kR (~10^3 rad) or R^2 (~10^4 turns^2): the phase is carried
// to ~1e-7 turns.  Each work-group checks, from its plane and the pillars' extent,
// that the first-order terms meet the error budget and otherwise takes the second
// order (fp32c_first_order); the host flags each plane where even that would not, and
// such a plane is taken in 1 x 1 blocks (every pair at its own FP64 R0).  The factor
// exp(2 pi i z'_ref) is restored in FP64 by the reduction; split partials are added in
// FP64.  Storing the FP32C partials as floats halves the reductions' HBM reads (they
// are bandwidth-bound) for one more float rounding of each chunk's sum, ~6e-8 of a
// partial.
#include "rs_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>

namespace crs {

namespace {

constexpr double kInv2Pi = 0.15915494309189533577;  // 1 / (2 pi)
constexpr double kTwoPi = 6.28318530717958647693;

// Distance from a pillar to a block centre, R0 = sqrt(R0sq), to FP64 accuracy: the
// float rsq seeds one FP64 Newton step.  Returns R0; rf = 1/R0 as a float.
__device__ inline double block_distance(double R0sq, float& rf) {
    const float s = static_cast<float>(R0sq);
    rf = __builtin_amdgcn_rsqf(s);
    const double R0f = static_cast<double>(s * rf);
    const double res = fma(-R0f, R0f, R0sq);
    return fma(res, static_cast<double>(0.5f * rf), R0f);
}

// LDS record of one pillar for the FP32C forward: FP64 turn-scaled position, the
// phase offset z'_ref - arg(t)/2pi, and |t| k/2pi.
struct alignas(16) PillarRec {
    double x, y, zeta;
    float amp, pad;
};

}  // namespace

// ---------------------------------------------------------------------------------
// Staging: float64 pillar lists -> the layout each mode's pair loop reads.

// FP32C: positions stay FP64 (pos[npil], shared by all sets; each plane scales them
// by its own k/(2 pi) when it loads a tile) and {|t|, arg(t)/(2 pi)} per pillar and set.
// t == NULL stages positions only.
__global__ __launch_bounds__(kThreads) void stage_pillars_fp32c(
    const double* __restrict__ px, const double* __restrict__ py, const double2* __restrict__ t,
    int npil, int64_t total, double2* __restrict__ pos, float2* __restrict__ ap) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    if (idx < npil) pos[idx] = make_double2(px[idx], py[idx]);
    if (t) {
        const double2 v = t[idx];
        ap[idx] = make_float2(static_cast<float>(hypot(v.x, v.y)),
                              static_cast<float>(atan2(v.y, v.x) * kInv2Pi));
    }
}

// FP32: {x, y, Re t, Im t} in float (the reference's host staging, propagate.cu:145-152).
// t == NULL stages positions with zero transmission (the adjoint reads x, y only).
__global__ void stage_pillars_fp32(const double* __restrict__ px, const double* __restrict__ py,
                                   const double2* __restrict__ t, int npil, int64_t total,
                                   float4* __restrict__ out) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int q = static_cast<int>(idx % npil);
    const double2 v = t ? t[idx] : make_double2(0.0, 0.0);
    out[idx] = make_float4(static_cast<float>(px[q]), static_cast<float>(py[q]),
                           static_cast<float>(v.x), static_cast<float>(v.y));
}

// FP64: {x, y, Re t, Im t} in double.
__global__ void stage_pillars_fp64(const double* __restrict__ px, const double* __restrict__ py,
                                   const double2* __restrict__ t, int npil, int64_t total,
                                   double4* __restrict__ out) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int q = static_cast<int>(idx % npil);
    const double2 v = t ? t[idx] : make_double2(0.0, 0.0);
    out[idx] = make_double4(px[q], py[q], v.x, v.y);
}

// ---------------------------------------------------------------------------------
// Forward, FP32C.  A work-group owns kFwdTiles thread tiles (one BX x BY pixel block per
// lane) of one plane; its 4 waves split the pillar chunk: pillars stream through LDS in
// tiles of 256, scaled to turns by the plane's k/(2 pi) as they are loaded, and wave w
// takes entries [64w, 64w + 64) of each.  The waves' sums are added in FP64 at the end.
// BX = BY = 1 evaluates every pair at its own FP64 reference distance (no expansion).

// Per-lane constants of a pixel block: its centre (turns, FP64) and the offsets of its
// columns / rows, te = 2 e, e2 = e^2 (+ z_row'^2 - z_centre'^2 for the rows).
template <int BX, int BY>
struct BlockGeom {
    double FXc, FYc, ZC2;
    float tex[BX], ex2[BX], tey[BY], eyz[BY];
};

// One pillar against one block: th0 = frac(R0 - zeta), h = 1/(2 R0), c1 = |t| (k/2pi)
// / R0', and the separable parts of D (first order) or of x1 = D h (second order):
// a[c] = 2 e_c dx0 + e_c^2, b[r] = 2 e_r dy0 + e_r^2 + dz2_r, times h for GEN = 1.
//   GEN = 0:  p = h^3, q = -2 c1 h^2, c = c1   (block_pairs folds h into them)
//   GEN = 1:  p = 2 c1, q = -c1
//   1 x 1 blocks: c = c1 only (R = R0).
template <int BX, int BY>
struct PairState {
    float th0, h, p, q, c;
    float a[BX], b[BY];
};

template <int GEN, int BX, int BY>
__device__ __forceinline__ void pair_state(const PillarRec& T, const BlockGeom<BX, BY>& G,
                                           PairState<BX, BY>& S) {
    const double dX0 = G.FXc - T.x, dY0 = G.FYc - T.y;
    float rf;
    const double R0 = block_distance(fma(dX0, dX0, fma(dY0, dY0, G.ZC2)), rf);
    S.th0 = static_cast<float>(__builtin_amdgcn_fract(R0 - T.zeta));
    S.h = 0.5f * rf;
    const float c1 = T.amp * rf;
    if constexpr (BX * BY == 1) {
        S.c = c1;
    } else if constexpr (GEN) {
        S.p = c1 + c1;
        S.q = -c1;
    } else {
        const float h2 = S.h * S.h;
        S.p = h2 * S.h;
        S.q = -2.0f * c1 * h2;
        S.c = c1;
    }
    if constexpr (BX * BY > 1) {
        const float dx0 = static_cast<float>(dX0), dy0 = static_cast<float>(dY0);
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            S.a[c] = fmaf(G.tex[c], dx0, G.ex2[c]);
            if constexpr (GEN) S.a[c] *= S.h;
        }
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            S.b[r] = fmaf(G.tey[r], dy0, G.eyz[r]);
            if constexpr (GEN) S.b[r] *= S.h;
        }
    }
}

// The BX x BY pairs of one pillar.  With x1 = D h, y = x1 h = D / (4 R0^2):
//   GEN = 0:  R - R0 = x1 w,  w = 1 - y;             R0/R = 2 w - 1 = 1 - 2 y
//   GEN = 1:  R - R0 = x1 w,  w = 1 - y (1 - 2 y);   R0/R = 1 + y (6 y - 2)
// and E += c1 (R0/R) (cos 2 pi th, sin 2 pi th), th = th0 + (R - R0), in turns.
// GEN = 0 works on D itself: th = th0 + D (h - D h^3), A = c1 + (-2 c1 h^2) D (h folded
// into per-pillar constants); GEN = 1 on x1 in Horner form.  6 and 9 FP32 ops per pair
// besides the two transcendentals.
template <int GEN, int BX, int BY>
__device__ __forceinline__ void block_pairs(const PairState<BX, BY>& S, float (&ar)[BX * BY],
                                            float (&ai)[BX * BY]) {
    if constexpr (BX * BY == 1) {
        const float A = S.c;
        ar[0] = fmaf(A, __builtin_amdgcn_cosf(S.th0), ar[0]);
        ai[0] = fmaf(A, __builtin_amdgcn_sinf(S.th0), ai[0]);
    } else {
#pragma unroll
        for (int r = 0; r < BY; ++r) {
#pragma unroll
            for (int c = 0; c < BX; ++c) {
                const float d = S.a[c] + S.b[r];  // D (GEN = 0) or x1 (GEN = 1)
                float th, A;
                if constexpr (GEN) {
                    const float y = d * S.h;
                    const float w = fmaf(-y, fmaf(-2.0f, y, 1.0f), 1.0f);
                    A = fmaf(fmaf(3.0f * S.p, y, -S.p), y, -S.q);
                    th = fmaf(d, w, S.th0);
                } else {
                    th = fmaf(d, fmaf(-d, S.p, S.h), S.th0);
                    A = fmaf(S.q, d, S.c);
                }
                const int e = r * BX + c;
                ar[e] = fmaf(A, __builtin_amdgcn_cosf(th), ar[e]);
                ai[e] = fmaf(A, __builtin_amdgcn_sinf(th), ai[e]);
            }
        }
    }
}

// Pillars tile[k0 .. k1) against one block, one pillar state at a time.  (Issuing the
// next pillar's FP64 distance work alongside the current pillar's pairs, with two
// states alternating, took 143 VGPRs and 3 work-groups per CU and ran 3% slower.)
template <int GEN, int BX, int BY>
__device__ __forceinline__ void pillar_run(const PillarRec* tile, int k0, int k1,
                                                 const BlockGeom<BX, BY>& G,
                                                 float (&ar)[BX * BY], float (&ai)[BX * BY]) {
    for (int k = k0; k < k1; ++k) {
        PairState<BX, BY> st;
        pair_state<GEN>(tile[k], G, st);
        block_pairs<GEN>(st, ar, ai);
    }
}

// Whether the first-order pair terms meet the error budget for every 4 x 4 block of
// plane P against pillars within |x| <= ext.x, |y| <= ext.y (um).  With e' the block's
// half diagonal, e_z' its half depth, z' the nearest row and s = rho / R0 a pair's
// obliquity (at most sn, the steepest pillar-to-pixel direction): |D| / R0 <= B(s) =
// a s + c, a = 2 e', c = 2 e_z' + (e'^2 + e_z'^2) / z', and the dropped terms are
// 2 x1 y^2 <= B^3 / (16 R0^2) turns and 6 y^2 <= 3 B^2 / (8 R0^2).  An oblique pair is
// also a far one, R0^2 >= z'^2 / (1 - s^2), so the bounds are (a s + c)^k (1 - s^2) /
// z'^2 (k = 3, 2), which rise up to the root s_k of their derivative,
// s_3 = (sqrt(c^2 + 15 a^2) - c) / (5 a) and s_2 = (sqrt(c^2 + 8 a^2) - c) / (4 a), and
// fall after it: each is checked at min(s_k, sn).
__device__ inline bool fp32c_first_order(const DevPlane& P, double2 ext, int bx, int by) {
    const double xl = P.x0 + (P.nx - 1) * P.sx, yl = P.y0 + (P.ny - 1) * P.sy;
    const double rx = fmax(fabs(P.x0), fabs(xl)) + ext.x;
    const double ry = fmax(fabs(P.y0), fabs(yl)) + ext.y;
    const double zmin = plane_min_abs_z(P.z0, P.sz, P.ny);
    const double rho2 = rx * rx + ry * ry;
    const double sn = sqrt(rho2 / (rho2 + zmin * zmin));
    const double hx = 0.5 * (bx - 1) * fabs(P.sx) * P.kturn;
    const double hy = 0.5 * (by - 1) * fabs(P.sy) * P.kturn;
    const double hz = 0.5 * (by - 1) * fabs(P.sz) * P.kturn;
    const double e = sqrt(hx * hx + hy * hy);
    const double z = zmin * P.kturn;
    const double a = 2.0 * e, c = 2.0 * hz + (e * e + hz * hz) / z;
    const double s3 = a > 0.0 ? fmin(sn, (sqrt(c * c + 15.0 * a * a) - c) / (5.0 * a)) : 0.0;
    const double s2 = a > 0.0 ? fmin(sn, (sqrt(c * c + 8.0 * a * a) - c) / (4.0 * a)) : 0.0;
    const double B3 = a * s3 + c, B2 = a * s2 + c;
    const double z2 = z * z;
    return B3 * B3 * B3 * (1.0 - s3 * s3) / (16.0 * z2) <= 1.6e-7 &&
           3.0 * B2 * B2 * (1.0 - s2 * s2) / (8.0 * z2) <= 4e-6;
}

// Pixel origin of thread tile t of a plane's tile class (shape): 0 = the 4 x 4 tiles
// of the largest multiple-of-4 corner of the plane (tx x ty of them), 1 = 1 x 4 tiles
// of the rx = nx - 4 tx leftover columns, 2 = 4 x 1 tiles of the ry = ny - 4 ty
// leftover rows, 3 = 1 x 1 tiles of the leftover corner.  With tx = ty = 0 (no
// blocks) class 3 covers the whole plane.  No tile reaches outside the plane.
__device__ inline int tile_count(const DevPlane& P, int shape) {
    const int rx = P.nx - 4 * P.tx, ry = P.ny - 4 * P.ty;
    switch (shape) {
        case 0: return P.tx * P.ty;
        case 1: return rx * P.ty;
        case 2: return P.tx * ry;
        default: return rx * ry;
    }
}

__device__ inline void tile_origin(const DevPlane& P, int shape, int t, int& ox, int& oy) {
    const int rx = P.nx - 4 * P.tx;
    switch (shape) {
        case 0: ox = 4 * (t % P.tx); oy = 4 * (t / P.tx); break;
        case 1: ox = 4 * P.tx + t % rx; oy = 4 * (t / rx); break;
        case 2: ox = 4 * (t % P.tx); oy = 4 * P.ty + t / P.tx; break;
        default: ox = 4 * P.tx + t % rx; oy = 4 * P.ty + t / rx; break;
    }
}

// One work-group's pixel tiles of one class, all BX x BY, against its pillar chunk.
template <int BX, int BY>
__device__ __forceinline__ void fwd_tiles(
    const double2* __restrict__ pos, const float2* __restrict__ ap, int npil,
    const DevPlane& P, const FwdBlock& B, PillarRec* tile, float2* run,
    float2* __restrict__ partial) {
    constexpr int N = BX * BY;
    const int nx = P.nx;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int my_tile = B.tile0 + lane;
    const bool active = my_tile < tile_count(P, B.shape);
    int ox, oy;
    tile_origin(P, B.shape, active ? my_tile : B.tile0, ox, oy);
    const double kt = P.kturn;
    const double zref = P.zref;

    // Series order from the work-group's own pillar chunk: its extent max |x|, max |y|
    // (um), reduced through the running-sum buffer before that is cleared.  Uniform
    // across the work-group; readfirstlane makes the branch on it scalar.
    bool gen = false;
    if constexpr (N > 1) {
        double mx = 0.0, my = 0.0;
        for (int q = B.q0 + tid; q < B.q1; q += kThreads) {
            const double2 xy = pos[q];
            mx = fmax(mx, fabs(xy.x));
            my = fmax(my, fabs(xy.y));
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mx = fmax(mx, __shfl_xor(mx, o));
            my = fmax(my, __shfl_xor(my, o));
        }
        double* red = reinterpret_cast<double*>(run);
        if (lane == 0) {
            red[2 * wave] = mx;
            red[2 * wave + 1] = my;
        }
        __syncthreads();
        double2 E = make_double2(red[0], red[1]);
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
            E.x = fmax(E.x, red[2 * w]);
            E.y = fmax(E.y, red[2 * w + 1]);
        }
        gen = __builtin_amdgcn_readfirstlane(!fp32c_first_order(P, E, BX, BY));
        __syncthreads();  // every wave has read the maxima before run is cleared
    }

    BlockGeom<BX, BY> G;
    {
        const double ic = ox + (BX - 1) * 0.5, jc = oy + (BY - 1) * 0.5;
        G.FXc = kt * (P.x0 + ic * P.sx);
        G.FYc = kt * (P.y0 + jc * P.sy);
        const double zc = kt * (P.z0 + jc * P.sz);
        G.ZC2 = zc * zc;
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            const float e = static_cast<float>(kt * P.sx * (c - (BX - 1) * 0.5));
            G.tex[c] = 2.0f * e;
            G.ex2[c] = e * e;
        }
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            const float e = static_cast<float>(kt * P.sy * (r - (BY - 1) * 0.5));
            const double zr = kt * (P.z0 + (oy + r) * P.sz);
            G.tey[r] = 2.0f * e;
            G.eyz[r] = e * e + static_cast<float>(zr * zr - G.ZC2);
        }
    }
    const float amp_scale = static_cast<float>(kt);  // 1/R = (k/2pi) / R'
    const float2* __restrict__ aps = ap + static_cast<int64_t>(P.tset) * npil;

    // Each lane's running sums live in LDS (float, one slot per wave, pixel and lane),
    // fed by register sums over at most 64 pillars per LDS tile.
#pragma unroll
    for (int e = 0; e < N; ++e) run[(wave * N + e) * 64 + lane] = make_float2(0.0f, 0.0f);

    for (int base = B.q0; base < B.q1; base += kThreads) {
        const int n = min(kThreads, B.q1 - base);
        __syncthreads();
        if (tid < n) {
            const double2 xy = pos[base + tid];
            const float2 a = aps[base + tid];
            PillarRec rec;
            rec.x = kt * xy.x;
            rec.y = kt * xy.y;
            rec.zeta = zref - static_cast<double>(a.y);
            rec.amp = a.x * amp_scale;
            rec.pad = 0.0f;
            tile[tid] = rec;
        }
        __syncthreads();
        const int k0 = (wave * n) / kWaves, k1 = ((wave + 1) * n) / kWaves;
        if (k0 < k1) {
            float ar[N], ai[N];
#pragma unroll
            for (int e = 0; e < N; ++e) ar[e] = ai[e] = 0.0f;
            if (gen)
                pillar_run<1>(tile, k0, k1, G, ar, ai);
            else
                pillar_run<0>(tile, k0, k1, G, ar, ai);
#pragma unroll
            for (int e = 0; e < N; ++e) {
                float2& v = run[(wave * N + e) * 64 + lane];
                v.x += ar[e];
                v.y += ai[e];
            }
        }
    }

    // Wave 0 adds the 4 waves' sums in FP64, in wave order, and stores the float pair.
    __syncthreads();
    if (wave != 0 || !active) return;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            const int e = r * BX + c;
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const float2 o = run[(w * N + e) * 64 + lane];
                vr += o.x;
                vi += o.y;
            }
            partial[B.part + static_cast<int64_t>(oy + r) * nx + (ox + c)] =
                make_float2(static_cast<float>(vr), static_cast<float>(vi));
        }
    }
}

// BLK = 4: 4 x 4 tiles plus the edge classes (1 x 4, 4 x 1, 1 x 1), chosen per
// work-group from its descriptor (a plane flagged exact has 1 x 1 tiles only); BLK = 1,
// for calls whose planes are all flagged: 1 x 1 tiles only (every pair at its own
// FP64 reference distance).  123 VGPRs and 40 KB of LDS: 4 work-groups per CU.
template <int BLK>
__global__ __launch_bounds__(kThreads, 4) void fwd_fp32c(
    const double2* __restrict__ pos, const float2* __restrict__ ap, int npil,
    const DevPlane* __restrict__ planes, const FwdBlock* __restrict__ blocks,
    float2* __restrict__ partial) {
    constexpr int NMAX = BLK * BLK;
    __shared__ PillarRec tile[kThreads];
    __shared__ float2 run[kWaves * NMAX * 64];
    const FwdBlock B = blocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    if constexpr (BLK == kBlk) {
        switch (B.shape) {
            case 0: fwd_tiles<kBlk, kBlk>(pos, ap, npil, P, B, tile, run, partial); break;
            case 1: fwd_tiles<1, kBlk>(pos, ap, npil, P, B, tile, run, partial); break;
            case 2: fwd_tiles<kBlk, 1>(pos, ap, npil, P, B, tile, run, partial); break;
            default: fwd_tiles<1, 1>(pos, ap, npil, P, B, tile, run, partial); break;
        }
    } else {
        fwd_tiles<1, 1>(pos, ap, npil, P, B, tile, run, partial);
    }
}

// ---------------------------------------------------------------------------------
// Forward, FP32 (the reference kernel's arithmetic, propagate.cu:68-78).  One pixel
// per thread.

__global__ __launch_bounds__(kThreads) void fwd_fp32(
    const float4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const FwdBlock* __restrict__ blocks, double2* __restrict__ partial) {
    __shared__ float4 tile[kThreads];
    const FwdBlock B = blocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int nx = P.nx, ny = P.ny;
    const int tid = threadIdx.x;
    const int pix = B.tile0 + tid;
    const bool active = pix < nx * ny;
    const int pc = active ? pix : 0;
    const int i = pc % nx, j = pc / nx;
    const float fx = static_cast<float>(P.x0) + static_cast<float>(i) * static_cast<float>(P.sx);
    const float fy = static_cast<float>(P.y0) + static_cast<float>(j) * static_cast<float>(P.sy);
    const float z = static_cast<float>(P.z0) + static_cast<float>(j) * static_cast<float>(P.sz);
    const float k = static_cast<float>(P.k);
    const float4* __restrict__ src = pil + static_cast<int64_t>(P.tset) * npil;
    float er = 0.0f, ei = 0.0f;
    for (int base = B.q0; base < B.q1; base += kThreads) {
        const int n = min(kThreads, B.q1 - base);
        __syncthreads();
        if (tid < n) tile[tid] = src[base + tid];
        __syncthreads();
        for (int q = 0; q < n; ++q) {
            const float4 Q = tile[q];
            const float dx = fx - Q.x, dy = fy - Q.y;
            const float R = sqrtf(dx * dx + dy * dy + z * z);
            float s, c;
            sincosf(k * R, &s, &c);
            const float inv = 1.0f / R;
            er += (Q.z * c - Q.w * s) * inv;
            ei += (Q.z * s + Q.w * c) * inv;
        }
    }
    if (active) partial[B.part + pix] = make_double2(er, ei);
}

// ---------------------------------------------------------------------------------
// Forward, FP64: the reference algorithm in double (focal.cpp:229-233).  One pixel
// per thread.

__global__ __launch_bounds__(kThreads) void fwd_fp64(
    const double4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const FwdBlock* __restrict__ blocks, double2* __restrict__ partial) {
    __shared__ double4 tile[kThreads];
    const FwdBlock B = blocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int nx = P.nx, ny = P.ny;
    const int tid = threadIdx.x;
    const int pix = B.tile0 + tid;
    const bool active = pix < nx * ny;
    const int pc = active ? pix : 0;
    const int i = pc % nx, j = pc / nx;
    const double fx = P.x0 + i * P.sx;
    const double fy = P.y0 + j * P.sy;
    const double z = P.z0 + j * P.sz;
    const double k = P.k;
    const double4* __restrict__ src = pil + static_cast<int64_t>(P.tset) * npil;
    double er = 0.0, ei = 0.0;
    for (int base = B.q0; base < B.q1; base += kThreads) {
        const int n = min(kThreads, B.q1 - base);
        __syncthreads();
        if (tid < n) tile[tid] = src[base + tid];
        __syncthreads();
        for (int q = 0; q < n; ++q) {
            const double4 Q = tile[q];
            const double dx = fx - Q.x, dy = fy - Q.y;
            const double R = sqrt(dx * dx + dy * dy + z * z);
            double s, c;
            sincos(k * R, &s, &c);
            const double rho = 1.0 / R;
            const double pr = rho * c, pim = rho * s;
            er += Q.z * pr - Q.w * pim;
            ei += Q.z * pim + Q.w * pr;
        }
    }
    if (active) partial[B.part + pix] = make_double2(er, ei);
}

// ---------------------------------------------------------------------------------
// Forward reduction: add the plane's pillar-chunk partials, restore the factor
// exp(2 pi i z'_ref) that FP32C leaves out, write field and/or |E|^2.  A work-group
// takes kReduceCols pixels; wave w adds the w-th quarter of the slots, and wave 0 adds
// the four quarter sums in order (a fixed order: results are bitwise repeatable).

// T: float2 (FP32C partials, each one chunk's sum rounded once) or double2.
template <typename T>
__global__ __launch_bounds__(kThreads) void fwd_reduce(
    const DevPlane* __restrict__ planes, const PixBlock* __restrict__ pblocks,
    const T* __restrict__ partial, int restore_phase, double2* __restrict__ field,
    double* __restrict__ intensity) {
    __shared__ double2 red[kWaves][kReduceCols];
    const PixBlock B = pblocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int npix = P.nx * P.ny;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = B.pix0 + lane;
    const bool valid = p < npix;
    const int s0 = (P.nslots * wave) / kWaves, s1 = (P.nslots * (wave + 1)) / kWaves;
    double er = 0.0, ei = 0.0;
    if (valid) {
        const T* __restrict__ src = partial + P.part + p;
#pragma unroll 4
        for (int s = s0; s < s1; ++s) {
            const T v = src[static_cast<int64_t>(s) * npix];
            er += v.x;
            ei += v.y;
        }
    }
    red[wave][lane] = make_double2(er, ei);
    __syncthreads();
    if (wave != 0 || !valid) return;
    er = 0.0;
    ei = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        er += red[w][lane].x;
        ei += red[w][lane].y;
    }
    if (restore_phase) {
        double sn, cs;
        sincos(kTwoPi * (P.zref - floor(P.zref)), &sn, &cs);
        const double r = er * cs - ei * sn;
        ei = er * sn + ei * cs;
        er = r;
    }
    const int64_t o = P.offset + p;
    if (field) field[o] = make_double2(er, ei);
    if (intensity) intensity[o] = er * er + ei * ei;
}

// ---------------------------------------------------------------------------------
// Adjoint staging (FP32C).  Pixel records are laid out quad by quad (4 rows), column
// by column, 4 records per column: record (quad u, column i, row r) sits at
// sbase + (u*nxp + i)*4 + r, so an LDS tile of 64 columns x 4 rows is one contiguous
// run.  Record = {2U, -U, arg(u)/2pi - z'_ref mod 1, 0} with U = |u| k/2pi; padding
// (rows >= ny, columns >= nx) has U = 0 and contributes nothing.

__global__ __launch_bounds__(kThreads) void stage_pixels_fp32c(
    const DevPlane* __restrict__ planes, const PixBlock* __restrict__ pblocks,
    const double2* __restrict__ u, float4* __restrict__ pix) {
    const PixBlock B = pblocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int nrec = P.nquads * P.nxp * kQuad;
    const int rix = B.pix0 + static_cast<int>(threadIdx.x);
    if (rix >= nrec) return;
    const int quad = rix / (P.nxp * kQuad);
    const int rem = rix - quad * P.nxp * kQuad;
    const int i = rem / kQuad, j = quad * kQuad + rem % kQuad;
    float4 rec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < P.nx && j < P.ny) {
        const double2 v = u[P.offset + static_cast<int64_t>(j) * P.nx + i];
        double chi = atan2(v.y, v.x) * kInv2Pi - P.zref;
        chi -= rint(chi);
        const float U = static_cast<float>(hypot(v.x, v.y) * P.kturn);
        rec = make_float4(2.0f * U, -U, static_cast<float>(chi), 0.0f);
    }
    pix[P.sbase + rix] = rec;
}

// ---------------------------------------------------------------------------------
// Adjoint, FP32C.  Thread owns kPQ = 1 pillar (FP64 position in registers); a quad's pixel
// records stream through LDS 64 columns at a time and are consumed in BX x BY blocks
// (4 x 4; 1 x 1 for planes flagged exact, where every pair needs its own reference
// distance, which the host sends to the <1, 1> instance).  Per pair
//     conj(G) u = U / R0' (R0/R) exp(2 pi i (chi - th0 - (R - R0))),   U = |u| k/2pi,
// with the same expansion of R - R0 and R0/R as the forward; the block's sum is scaled
// by 1/R0' once.

// One pillar against one block: th0 = frac(R0 - z'_ref), h = 1/(2 R0), rf = 1/R0 and
// the separable parts a[c], b[r] of x1 = D h.
template <int BX, int BY>
struct AdjState {
    float th0, h, rf;
    float a[BX], b[BY];
};

template <int BX, int BY>
__device__ __forceinline__ void adj_state(double dX0, double dY0, double zref,
                                          const BlockGeom<BX, BY>& G, AdjState<BX, BY>& S) {
    float rf;
    const double R0 = block_distance(fma(dX0, dX0, fma(dY0, dY0, G.ZC2)), rf);
    S.th0 = static_cast<float>(__builtin_amdgcn_fract(R0 - zref));
    S.rf = rf;
    S.h = 0.5f * rf;
    if constexpr (BX * BY > 1) {
        const float dx0 = static_cast<float>(dX0), dy0 = static_cast<float>(dY0);
#pragma unroll
        for (int c = 0; c < BX; ++c) S.a[c] = fmaf(G.tex[c], dx0, G.ex2[c]) * S.h;
#pragma unroll
        for (int r = 0; r < BY; ++r) S.b[r] = fmaf(G.tey[r], dy0, G.eyz[r]) * S.h;
    }
}

// The pairs of one pillar with one block's records W (column-major: W[c*BY + r] =
// {2U, -U, chi}); adds the block's sum to (gr, gi).  GEN as in block_pairs.  For 4 x 4
// blocks the pair phases are chi - (R - R0) and the block's sum is turned by
// exp(-2 pi i th0) and scaled by 1/R0' once; 1 x 1 blocks use chi - th0 directly.
template <int GEN, int BX, int BY>
__device__ __forceinline__ void adj_pairs(const AdjState<BX, BY>& S, const float3 (&W)[BX * BY],
                                          float& gr, float& gi) {
    if constexpr (BX * BY == 1) {
        const float phi = W[0].z - S.th0;
        const float A = S.rf * (W[0].x + W[0].y);
        gr = fmaf(A, __builtin_amdgcn_cosf(phi), gr);
        gi = fmaf(A, __builtin_amdgcn_sinf(phi), gi);
    } else {
        float br = 0.0f, bi = 0.0f;
#pragma unroll
        for (int c = 0; c < BX; ++c) {
#pragma unroll
            for (int r = 0; r < BY; ++r) {
                const float3 w3 = W[c * BY + r];
                const float x1 = S.a[c] + S.b[r];
                float w, A;
                if constexpr (GEN) {  // U = -w3.y, 2U = w3.x: A = U (1 + y (6 y - 2))
                    const float y = x1 * S.h;
                    w = fmaf(-y, fmaf(-2.0f, y, 1.0f), 1.0f);
                    A = fmaf(fmaf(3.0f * w3.x, y, -w3.x), y, -w3.y);
                } else {
                    w = fmaf(-x1, S.h, 1.0f);
                    A = fmaf(w3.x, w, w3.y);
                }
                const float phi = fmaf(-x1, w, w3.z);
                br = fmaf(A, __builtin_amdgcn_cosf(phi), br);
                bi = fmaf(A, __builtin_amdgcn_sinf(phi), bi);
            }
        }
        const float cr = S.rf * __builtin_amdgcn_cosf(S.th0);
        const float sn = S.rf * __builtin_amdgcn_sinf(S.th0);
        gr = fmaf(br, cr, fmaf(bi, sn, gr));
        gi = fmaf(bi, cr, fmaf(-br, sn, gi));
    }
}

// One 64-column LDS tile of a quad against the thread's pillars.
template <int GEN, int BX, int BY>
__device__ __forceinline__ void adj_tile(const float4* tile, int ncol, const DevPlane& P, int j0,
                                         int cb, const double (&qx)[kPQ],
                                         const double (&qy)[kPQ], float (&ar)[kPQ],
                                         float (&ai)[kPQ]) {
    const double kt = P.kturn;
    BlockGeom<BX, BY> G;
#pragma unroll
    for (int c = 0; c < BX; ++c) {
        const float e = static_cast<float>(kt * P.sx * (c - (BX - 1) * 0.5));
        G.tex[c] = 2.0f * e;
        G.ex2[c] = e * e;
    }
#pragma unroll
    for (int sr0 = 0; sr0 < kQuad; sr0 += BY) {
        const double jc = j0 + sr0 + (BY - 1) * 0.5;
        G.FYc = kt * (P.y0 + jc * P.sy);
        const double zc = kt * (P.z0 + jc * P.sz);
        G.ZC2 = zc * zc;
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            const float e = static_cast<float>(kt * P.sy * (r - (BY - 1) * 0.5));
            const double zr = kt * (P.z0 + (j0 + sr0 + r) * P.sz);
            G.tey[r] = 2.0f * e;
            G.eyz[r] = e * e + static_cast<float>(zr * zr - G.ZC2);
        }
        double dY0[kPQ];
#pragma unroll
        for (int k = 0; k < kPQ; ++k) dY0[k] = fma(-kt, qy[k], G.FYc);
        for (int g = 0; g < ncol; g += BX) {
            G.FXc = kt * (P.x0 + (cb + g + (BX - 1) * 0.5) * P.sx);
            float3 W[BX * BY];
#pragma unroll
            for (int c = 0; c < BX; ++c)
#pragma unroll
                for (int r = 0; r < BY; ++r) {
                    const float4 v = tile[(g + c) * kQuad + sr0 + r];
                    W[c * BY + r] = make_float3(v.x, v.y, v.z);
                }
#pragma unroll
            for (int k = 0; k < kPQ; ++k) {
                AdjState<BX, BY> S;
                adj_state(fma(-kt, qx[k], G.FXc), dY0[k], P.zref, G, S);
                adj_pairs<GEN>(S, W, ar[k], ai[k]);
            }
        }
    }
}

template <int BX, int BY>
__global__ __launch_bounds__(kThreads, 3) void adj_fp32c(
    const double* __restrict__ px, const double* __restrict__ py, int npil,
    const DevPlane* __restrict__ planes, const AdjQuad* __restrict__ quads,
    const AdjBlock* __restrict__ blocks, const float4* __restrict__ pix,
    float2* __restrict__ partial) {
    __shared__ float4 tile[kThreads];
    __shared__ double2 red[kWaves];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    double qx[kPQ], qy[kPQ];
    double mx = 0.0, my = 0.0;
#pragma unroll
    for (int k = 0; k < kPQ; ++k) {
        const int q = min(B.q0 + tid + k * kThreads, npil - 1);
        qx[k] = px[q];
        qy[k] = py[q];
        mx = fmax(mx, fabs(qx[k]));
        my = fmax(my, fabs(qy[k]));
    }
    // The work-group's own pillar extent decides the series order (no global pass).
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        my = fmax(my, __shfl_xor(my, o));
    }
    if ((tid & 63) == 0) red[tid >> 6] = make_double2(mx, my);
    __syncthreads();
    double2 E = red[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
        E.x = fmax(E.x, red[w].x);
        E.y = fmax(E.y, red[w].y);
    }
    // Kahan-compensated running sums, fed by one float sum per LDS tile.
    float sr[kPQ], si[kPQ], cr[kPQ], ci[kPQ];
#pragma unroll
    for (int k = 0; k < kPQ; ++k) sr[k] = si[k] = cr[k] = ci[k] = 0.0f;

    for (int u = B.u0; u < B.u1; ++u) {
        const AdjQuad Q = quads[u];
        const DevPlane& P = planes[Q.plane];
        const bool gen =
            __builtin_amdgcn_readfirstlane(BX * BY > 1 && !fp32c_first_order(P, E, BX, kQuad));
        const float4* __restrict__ prow =
            pix + P.sbase + static_cast<int64_t>(Q.j0 / kQuad) * P.nxp * kQuad;
        for (int cb = 0; cb < P.nxp; cb += kAdjCols) {
            const int ncol = min(kAdjCols, P.nxp - cb);
            __syncthreads();
            if (tid < ncol * kQuad) tile[tid] = prow[cb * kQuad + tid];
            __syncthreads();
            float ar[kPQ], ai[kPQ];
#pragma unroll
            for (int k = 0; k < kPQ; ++k) ar[k] = ai[k] = 0.0f;
            if (gen)
                adj_tile<1, BX, BY>(tile, ncol, P, Q.j0, cb, qx, qy, ar, ai);
            else
                adj_tile<0, BX, BY>(tile, ncol, P, Q.j0, cb, qx, qy, ar, ai);
#pragma unroll
            for (int k = 0; k < kPQ; ++k) {
                const float yr = ar[k] - cr[k];
                const float tr = sr[k] + yr;
                cr[k] = (tr - sr[k]) - yr;
                sr[k] = tr;
                const float yi = ai[k] - ci[k];
                const float tq = si[k] + yi;
                ci[k] = (tq - si[k]) - yi;
                si[k] = tq;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kPQ; ++k) {
        const int q = B.q0 + tid + k * kThreads;
        if (q < npil)
            partial[B.part + q] =
                make_float2(static_cast<float>(static_cast<double>(sr[k]) - static_cast<double>(cr[k])),
                            static_cast<float>(static_cast<double>(si[k]) - static_cast<double>(ci[k])));
    }
}

// ---------------------------------------------------------------------------------
// Adjoint, FP32 and FP64 (reference pair arithmetic, transposed).  One pillar per
// thread; each row's u streams through LDS.

__global__ __launch_bounds__(kThreads) void adj_fp32(
    const float4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const AdjQuad* __restrict__ quads, const AdjBlock* __restrict__ blocks,
    const double2* __restrict__ u, double2* __restrict__ partial) {
    __shared__ float2 tile[kThreads];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    const int q = B.q0 + tid;
    const float4 v = pil[min(q, npil - 1)];
    float gr = 0.0f, gi = 0.0f;
    for (int uq = B.u0; uq < B.u1; ++uq) {
        const AdjQuad Q = quads[uq];
        const DevPlane& P = planes[Q.plane];
        const int nx = P.nx;
        const float k = static_cast<float>(P.k);
        for (int j = Q.j0; j < min(Q.j0 + kQuad, P.ny); ++j) {
            const float fy = static_cast<float>(P.y0) + static_cast<float>(j) * static_cast<float>(P.sy);
            const float z = static_cast<float>(P.z0) + static_cast<float>(j) * static_cast<float>(P.sz);
            const float dy = fy - v.y;
            const double2* __restrict__ urow = u + P.offset + static_cast<int64_t>(j) * nx;
            for (int base = 0; base < nx; base += kThreads) {
                const int n = min(kThreads, nx - base);
                __syncthreads();
                if (tid < n) {
                    const double2 w = urow[base + tid];
                    tile[tid] = make_float2(static_cast<float>(w.x), static_cast<float>(w.y));
                }
                __syncthreads();
                for (int p = 0; p < n; ++p) {
                    const float2 U = tile[p];
                    const float fx = static_cast<float>(P.x0) +
                                     static_cast<float>(base + p) * static_cast<float>(P.sx);
                    const float dx = fx - v.x;
                    const float R = sqrtf(dx * dx + dy * dy + z * z);
                    float s, c;
                    sincosf(k * R, &s, &c);
                    const float inv = 1.0f / R;
                    gr += (c * U.x + s * U.y) * inv;
                    gi += (c * U.y - s * U.x) * inv;
                }
            }
        }
    }
    if (q < npil) partial[B.part + q] = make_double2(gr, gi);
}

__global__ __launch_bounds__(kThreads) void adj_fp64(
    const double4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const AdjQuad* __restrict__ quads, const AdjBlock* __restrict__ blocks,
    const double2* __restrict__ u, double2* __restrict__ partial) {
    __shared__ double2 tile[kThreads];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    const int q = B.q0 + tid;
    const double4 v = pil[min(q, npil - 1)];
    double gr = 0.0, gi = 0.0;
    for (int uq = B.u0; uq < B.u1; ++uq) {
        const AdjQuad Q = quads[uq];
        const DevPlane& P = planes[Q.plane];
        const int nx = P.nx;
        for (int j = Q.j0; j < min(Q.j0 + kQuad, P.ny); ++j) {
            const double fy = P.y0 + j * P.sy;
            const double z = P.z0 + j * P.sz;
            const double dy = fy - v.y;
            const double2* __restrict__ urow = u + P.offset + static_cast<int64_t>(j) * nx;
            for (int base = 0; base < nx; base += kThreads) {
                const int n = min(kThreads, nx - base);
                __syncthreads();
                if (tid < n) tile[tid] = urow[base + tid];
                __syncthreads();
                for (int p = 0; p < n; ++p) {
                    const double2 U = tile[p];
                    const double fx = P.x0 + (base + p) * P.sx;
                    const double dx = fx - v.x;
                    const double R = sqrt(dx * dx + dy * dy + z * z);
                    double s, c;
                    sincos(P.k * R, &s, &c);
                    const double rho = 1.0 / R;
                    const double pr = rho * c, pim = rho * s;
                    gr += pr * U.x + pim * U.y;
                    gi += pr * U.y - pim * U.x;
                }
            }
        }
    }
    if (q < npil) partial[B.part + q] = make_double2(gr, gi);
}

// Adjoint reduction: g[idx] = sum over slots of partial[slot*n + idx]; kReduceCols
// outputs per work-group, slot quarters per wave, added in order as in fwd_reduce.
template <typename T>
__global__ __launch_bounds__(kThreads) void adj_reduce(const T* __restrict__ partial,
                                                       int nslots, int64_t n,
                                                       double2* __restrict__ g) {
    __shared__ double2 red[kWaves][kReduceCols];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * kReduceCols + lane;
    const bool valid = idx < n;
    const int s0 = (nslots * wave) / kWaves, s1 = (nslots * (wave + 1)) / kWaves;
    double gr = 0.0, gi = 0.0;
    if (valid) {
#pragma unroll 4
        for (int s = s0; s < s1; ++s) {
            const T v = partial[static_cast<int64_t>(s) * n + idx];
            gr += v.x;
            gi += v.y;
        }
    }
    red[wave][lane] = make_double2(gr, gi);
    __syncthreads();
    if (wave != 0 || !valid) return;
    gr = 0.0;
    gi = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        gr += red[w][lane].x;
        gi += red[w][lane].y;
    }
    g[idx] = make_double2(gr, gi);
}

// ---------------------------------------------------------------------------------
// MTF (analysis/mtf.cpp:10-63): autocorrelation of the exit pupil along x.  One
// work-group per shift s sums P[iy][ix] conj(P[iy][ix+s]) and P0[iy][ix] P0[iy][ix+s]
// over all rows and columns ix < n - s in FP64 (each thread a column stride, then a
// fixed-order tree), so the result is repeatable; mtf_finish normalises by shift 0 and
// casts to float as the reference's MtfCurve does.

__global__ __launch_bounds__(kThreads) void mtf_autocorr(const double2* __restrict__ P,
                                                         const double* __restrict__ P0, int n,
                                                         double2* __restrict__ acc,
                                                         double* __restrict__ acc0) {
    __shared__ double red[3][kThreads];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int m = n - s;
    double ar = 0.0, ai = 0.0, a0 = 0.0;
    for (int iy = 0; iy < n; ++iy) {
        const double2* __restrict__ row = P + static_cast<int64_t>(iy) * n;
        const double* __restrict__ row0 = P0 + static_cast<int64_t>(iy) * n;
        for (int ix = tid; ix < m; ix += kThreads) {
            const double2 a = row[ix], b = row[ix + s];
            ar += a.x * b.x + a.y * b.y;
            ai += a.y * b.x - a.x * b.y;
            a0 += row0[ix] * row0[ix + s];
        }
    }
    red[0][tid] = ar;
    red[1][tid] = ai;
    red[2][tid] = a0;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] += red[0][tid + w];
            red[1][tid] += red[1][tid + w];
            red[2][tid] += red[2][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        acc[s] = make_double2(red[0][0], red[1][0]);
        acc0[s] = red[2][0];
    }
}

__global__ __launch_bounds__(kThreads) void mtf_finish(const double2* __restrict__ acc,
                                                       const double* __restrict__ acc0, int ns,
                                                       double p, double f, double lam,
                                                       float* __restrict__ freq,
                                                       float* __restrict__ mtf,
                                                       float* __restrict__ ideal) {
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= ns) return;
    double nP = hypot(acc[0].x, acc[0].y);
    double n0 = fabs(acc0[0]);
    if (nP <= 0.0) nP = 1.0;
    if (n0 <= 0.0) n0 = 1.0;
    freq[s] = static_cast<float>((s * p) / (lam * f) * 1000.0);
    mtf[s] = static_cast<float>(hypot(acc[s].x, acc[s].y) / nP);
    ideal[s] = static_cast<float>(fabs(acc0[s]) / n0);
}

// ---------------------------------------------------------------------------------
// Launchers.

namespace {
inline unsigned grid_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }
}  // namespace

hipError_t launch_stage_pillars(int mode, const double* px, const double* py, const double* t,
                                int npil, int nsets, void* out, hipStream_t s) {
    const int64_t total = static_cast<int64_t>(npil) * nsets;
    const double2* tt = reinterpret_cast<const double2*>(t);
    if (mode == 0) {
        const StagedLayout L = fp32c_staged_layout(npil, nsets);
        unsigned char* base = static_cast<unsigned char*>(out);
        stage_pillars_fp32c<<<grid_for(total), kThreads, 0, s>>>(
            px, py, tt, npil, total, reinterpret_cast<double2*>(base + L.pos),
            reinterpret_cast<float2*>(base + L.ap));
    } else if (mode == 2) {
        stage_pillars_fp32<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<float4*>(out));
    } else {
        stage_pillars_fp64<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<double4*>(out));
    }
    return hipGetLastError();
}

hipError_t launch_forward(int mode, int order, const void* pil, int npil, int nsets,
                          const DevPlane* planes, const FwdBlock* blocks, int nblocks,
                          void* partial_v, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 0) {
        float2* partial = static_cast<float2*>(partial_v);
        const StagedLayout L = fp32c_staged_layout(npil, nsets);
        const unsigned char* base = static_cast<const unsigned char*>(pil);
        const double2* pos = reinterpret_cast<const double2*>(base + L.pos);
        const float2* ap = reinterpret_cast<const float2*>(base + L.ap);
        // order kExactOrder: 1 x 1 blocks (every pair at its own reference distance);
        // otherwise 4 x 4 blocks, series order chosen per pillar tile from its extent.
        if (order == kExactOrder)
            fwd_fp32c<1><<<nblocks, kThreads, 0, s>>>(pos, ap, npil, planes, blocks, partial);
        else
            fwd_fp32c<kBlk><<<nblocks, kThreads, 0, s>>>(pos, ap, npil, planes, blocks, partial);
    } else if (mode == 2) {
        fwd_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes,
                                              blocks, static_cast<double2*>(partial_v));
    } else {
        fwd_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes,
                                              blocks, static_cast<double2*>(partial_v));
    }
    return hipGetLastError();
}

hipError_t launch_fwd_reduce(int mode, const DevPlane* planes, const PixBlock* pblocks,
                             int nblocks, const void* partial, double2* field,
                             double* intensity, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 0)  // FP32C: float partials; the reduction restores exp(2 pi i z'_ref)
        fwd_reduce<float2><<<nblocks, kThreads, 0, s>>>(
            planes, pblocks, static_cast<const float2*>(partial), 1, field, intensity);
    else
        fwd_reduce<double2><<<nblocks, kThreads, 0, s>>>(
            planes, pblocks, static_cast<const double2*>(partial), 0, field, intensity);
    return hipGetLastError();
}

hipError_t launch_stage_pixels(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                               const double2* u, float4* pix, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    stage_pixels_fp32c<<<nblocks, kThreads, 0, s>>>(planes, pblocks, u, pix);
    return hipGetLastError();
}

hipError_t launch_adjoint(int mode, int order, const void* pil, int npil,
                          const DevPlane* planes, const AdjQuad* quads, const AdjBlock* blocks,
                          int nblocks, const void* pix_or_u, void* partial_v, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 0) {
        float2* partial = static_cast<float2*>(partial_v);
        // FP32C reads the caller's positions directly: pil = {px, py}.
        const double* const* xy = static_cast<const double* const*>(pil);
        const float4* pix = static_cast<const float4*>(pix_or_u);
        if (order == kExactOrder)
            adj_fp32c<1, 1><<<nblocks, kThreads, 0, s>>>(xy[0], xy[1], npil, planes, quads, blocks, pix, partial);
        else
            adj_fp32c<kBlk, kQuad><<<nblocks, kThreads, 0, s>>>(xy[0], xy[1], npil, planes, quads, blocks, pix, partial);
    } else if (mode == 2) {
        adj_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes, quads,
                                              blocks, static_cast<const double2*>(pix_or_u),
                                              static_cast<double2*>(partial_v));
    } else {
        adj_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes, quads,
                                              blocks, static_cast<const double2*>(pix_or_u),
                                              static_cast<double2*>(partial_v));
    }
    return hipGetLastError();
}

hipError_t launch_adj_reduce(int mode, const void* partial, int nslots, int64_t n, double2* g,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>((n + kReduceCols - 1) / kReduceCols);
    if (mode == 0)
        adj_reduce<float2><<<grid, kThreads, 0, s>>>(static_cast<const float2*>(partial), nslots,
                                                     n, g);
    else
        adj_reduce<double2><<<grid, kThreads, 0, s>>>(static_cast<const double2*>(partial),
                                                      nslots, n, g);
    return hipGetLastError();
}

hipError_t launch_mtf(const double2* P, const double* P0, int n, int ns, double p, double f,
                      double lam, double2* acc, double* acc0, float* freq, float* mtf,
                      float* ideal, hipStream_t s) {
    if (n <= 0 || ns <= 0) return hipSuccess;
    mtf_autocorr<<<ns, kThreads, 0, s>>>(P, P0, n, acc, acc0);
    mtf_finish<<<grid_for(ns), kThreads, 0, s>>>(acc, acc0, ns, p, f, lam, freq, mtf, ideal);
    return hipGetLastError();
}

}  // namespace crs
