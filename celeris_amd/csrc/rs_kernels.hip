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
// other side through LDS in 256-entry tiles that every lane reads by broadcast.
// Sums are split over pillar chunks (forward) or pixel rows (adjoint) so that small
// focal grids still launch thousands of work-groups across the 256 CUs; each split
// writes a float64 partial and a second kernel adds the partials in a fixed order,
// so results are bitwise repeatable run to run.
//
// FP32C ("FP32-compensated") pair arithmetic.  Every phase-carrying coordinate is
// formed in FP64 and scaled to turns, x' = k x / (2 pi) (likewise y, z, fx, fy), then
// rounded to float once; z'^2 is carried as a float hi+lo pair.  Per pillar-pixel pair:
//   R2  = dx'^2 + [(fy'-y')^2 + z'^2]       R' = sqrt(R2) is k R / (2 pi): the phase in turns
//   iR  = rsq(R2), R = R2*iR                 v_rsq_f32
//   th  = (R - (z'_hi - arg t/2pi)) - (R*R - R2)/(2R)
//                                            one Newton step for R (removes any bias of
//                                            the hardware root), with z'_hi subtracted
//                                            exactly (Sterbenz), so th = k(R - z)/2pi + arg t/2pi
//                                            carries float rounding of th itself, not of kR
//   th  = fract(th)                          v_fract_f32, exact
//   E  += |t| k/(2 pi) * iR * (cos 2pi th, sin 2pi th)   v_cos_f32 / v_sin_f32 take turns
// The dropped factor exp(2 pi i z'_hi) is applied in FP64 by the reduction kernel.  Each
// pillar tile's float partial is added into a Kahan-compensated running sum, and the
// partials of the pillar-chunk split are added in FP64.
#include "rs_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>

namespace crs {

namespace {

constexpr double kInv2Pi = 0.15915494309189533577;  // 1 / (2 pi)
constexpr double kTwoPi = 6.28318530717958647693;

// w^2 in FP64, split into the float nearest it and the float nearest the rest.
__device__ inline void split_sq(double w, float& hi, float& lo) {
    const double w2 = w * w;
    hi = static_cast<float>(w2);
    lo = static_cast<float>(w2 - static_cast<double>(hi));
}

// Phase (turns) of the pair: R - zeta with one Newton correction of R = sqrt(R2);
// also returns 1/R.  `sign` = +1 gives (R_true - zeta), -1 gives (zeta - R_true).
template <int SIGN>
__device__ inline float pair_turns(float R2, float zeta, float& iR) {
    iR = __builtin_amdgcn_rsqf(R2);
    const float R = R2 * iR;
    const float resh = 0.5f * fmaf(R, R, -R2);
    if constexpr (SIGN > 0)
        return fmaf(-resh, iR, R - zeta);
    else
        return fmaf(resh, iR, zeta - R);
}

}  // namespace

// ---------------------------------------------------------------------------------
// Staging: float64 pillar lists -> the layout each mode's pair loop reads.

// FP32C: positions stay FP64 (pos[npil], shared by all sets; each plane scales them
// by its own k/(2 pi) when it loads a tile) and {|t|, arg(t)/(2 pi)} per pillar and set.
__global__ void stage_pillars_fp32c(const double* __restrict__ px, const double* __restrict__ py,
                                    const double2* __restrict__ t, int npil, int64_t total,
                                    double2* __restrict__ pos, float2* __restrict__ ap) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int q = static_cast<int>(idx % npil);
    if (idx < npil) pos[q] = make_double2(px[q], py[q]);
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
// Forward, FP32C.  Thread owns a PX x PY pixel block; pillars stream through LDS,
// scaled to turns by the plane's k/(2 pi) as they are loaded.

template <int PX, int PY>
__global__ __launch_bounds__(kThreads, 2) void fwd_fp32c(
    const double2* __restrict__ pos, const float2* __restrict__ ap, int npil,
    const DevPlane* __restrict__ planes, const FwdBlock* __restrict__ blocks,
    double2* __restrict__ partial) {
    __shared__ float4 tile[kThreads];
    const FwdBlock B = blocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int nx = P.nx, ny = P.ny;
    const int tid = threadIdx.x;
    const int my_tile = B.tile0 + tid;
    const bool active = my_tile < P.tx * P.ty;
    const int tt = active ? my_tile : 0;
    const int ti = tt % P.tx, tj = tt / P.tx;
    const double kt = P.kturn;

    float fx[PX];
#pragma unroll
    for (int c = 0; c < PX; ++c) {
        const int i = min(ti * PX + c, nx - 1);
        fx[c] = static_cast<float>(kt * (P.x0 + static_cast<double>(i) * P.sx));
    }
    float fy[PY], zh[PY], z2h[PY], z2l[PY];
#pragma unroll
    for (int r = 0; r < PY; ++r) {
        const int j = min(tj * PY + r, ny - 1);
        fy[r] = static_cast<float>(kt * (P.y0 + static_cast<double>(j) * P.sy));
        const double zt = kt * (P.z0 + static_cast<double>(j) * P.sz);
        zh[r] = static_cast<float>(zt);
        split_sq(zt, z2h[r], z2l[r]);
    }
    const float amp_scale = static_cast<float>(kt);  // 1/R = (k/2pi) / R'
    const float2* __restrict__ aps = ap + static_cast<int64_t>(P.tset) * npil;

    float sr[PX * PY], si[PX * PY], cr[PX * PY], ci[PX * PY];
#pragma unroll
    for (int e = 0; e < PX * PY; ++e) sr[e] = si[e] = cr[e] = ci[e] = 0.0f;

    for (int base = B.q0; base < B.q1; base += kThreads) {
        const int n = min(kThreads, B.q1 - base);
        __syncthreads();
        if (tid < n) {
            const double2 xy = pos[base + tid];
            const float2 a = aps[base + tid];
            tile[tid] = make_float4(static_cast<float>(kt * xy.x), static_cast<float>(kt * xy.y),
                                    a.x * amp_scale, a.y);
        }
        __syncthreads();
        float ar[PX * PY], ai[PX * PY];
#pragma unroll
        for (int e = 0; e < PX * PY; ++e) ar[e] = ai[e] = 0.0f;
        for (int q = 0; q < n; ++q) {
            const float4 Q = tile[q];
            float dx[PX];
#pragma unroll
            for (int c = 0; c < PX; ++c) dx[c] = fx[c] - Q.x;
#pragma unroll
            for (int r = 0; r < PY; ++r) {
                const float dy = fy[r] - Q.y;
                const float dyz2 = fmaf(dy, dy, z2l[r]) + z2h[r];
                const float zeta = zh[r] - Q.w;
#pragma unroll
                for (int c = 0; c < PX; ++c) {
                    const float R2 = fmaf(dx[c], dx[c], dyz2);
                    float iR;
                    const float th = __builtin_amdgcn_fractf(pair_turns<1>(R2, zeta, iR));
                    const float A = Q.z * iR;
                    const int e = r * PX + c;
                    ar[e] = fmaf(A, __builtin_amdgcn_cosf(th), ar[e]);
                    ai[e] = fmaf(A, __builtin_amdgcn_sinf(th), ai[e]);
                }
            }
        }
        // Kahan: fold this pillar tile's partial into the running sum.
#pragma unroll
        for (int e = 0; e < PX * PY; ++e) {
            const float yr = ar[e] - cr[e];
            const float tr = sr[e] + yr;
            cr[e] = (tr - sr[e]) - yr;
            sr[e] = tr;
            const float yi = ai[e] - ci[e];
            const float tq = si[e] + yi;
            ci[e] = (tq - si[e]) - yi;
            si[e] = tq;
        }
    }

    if (!active) return;
#pragma unroll
    for (int r = 0; r < PY; ++r) {
        const int j = tj * PY + r;
        if (j >= ny) break;
#pragma unroll
        for (int c = 0; c < PX; ++c) {
            const int i = ti * PX + c;
            if (i >= nx) break;
            const int e = r * PX + c;
            partial[B.part + static_cast<int64_t>(j) * nx + i] =
                make_double2(static_cast<double>(sr[e]) - static_cast<double>(cr[e]),
                             static_cast<double>(si[e]) - static_cast<double>(ci[e]));
        }
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
// Forward reduction: add the plane's pillar-chunk partials in slot order, restore
// the factor exp(2 pi i z'_hi) that FP32C leaves out, write field and/or |E|^2.

__global__ __launch_bounds__(kThreads) void fwd_reduce(
    const DevPlane* __restrict__ planes, const PixBlock* __restrict__ pblocks,
    const double2* __restrict__ partial, int restore_phase, double2* __restrict__ field,
    double* __restrict__ intensity) {
    const PixBlock B = pblocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int npix = P.nx * P.ny;
    const int p = B.pix0 + static_cast<int>(threadIdx.x);
    if (p >= npix) return;
    double er = 0.0, ei = 0.0;
    for (int s = 0; s < P.nslots; ++s) {
        const double2 v = partial[P.part + static_cast<int64_t>(s) * npix + p];
        er += v.x;
        ei += v.y;
    }
    if (restore_phase) {
        // exp(2 pi i z'_hi), z'_hi = float(k z / 2pi) exactly as the pair kernel used it.
        const int j = p / P.nx;
        const double zh = static_cast<double>(static_cast<float>(P.kturn * (P.z0 + j * P.sz)));
        double sn, cs;
        sincos(kTwoPi * (zh - floor(zh)), &sn, &cs);
        const double r = er * cs - ei * sn;
        ei = er * sn + ei * cs;
        er = r;
    }
    const int64_t o = P.offset + p;
    if (field) field[o] = make_double2(er, ei);
    if (intensity) intensity[o] = er * er + ei * ei;
}

// ---------------------------------------------------------------------------------
// Adjoint staging (FP32C): per pixel {fx', |u| k/(2 pi), Z, 0} with fx' = k fx/(2 pi) and
// Z = z'_hi + (arg(u)/(2 pi) - z'_hi mod 1), so that Z - R' is the phase (turns) of
// conj(G) u up to whole turns: the pair loop then needs no per-row phase offset.

__global__ __launch_bounds__(kThreads) void stage_pixels_fp32c(
    const DevPlane* __restrict__ planes, const PixBlock* __restrict__ pblocks,
    const double2* __restrict__ u, float4* __restrict__ pix) {
    const PixBlock B = pblocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int npix = P.nx * P.ny;
    const int p = B.pix0 + static_cast<int>(threadIdx.x);
    if (p >= npix) return;
    const int i = p % P.nx, j = p / P.nx;
    const int64_t o = P.offset + p;
    const double kt = P.kturn;
    const double2 v = u[o];
    const double zh = static_cast<double>(static_cast<float>(kt * (P.z0 + j * P.sz)));
    double zeta = atan2(v.y, v.x) * kInv2Pi - zh;
    zeta -= rint(zeta);
    pix[o] = make_float4(static_cast<float>(kt * (P.x0 + static_cast<double>(i) * P.sx)),
                         static_cast<float>(hypot(v.x, v.y) * kt), static_cast<float>(zeta + zh),
                         0.0f);
}

// ---------------------------------------------------------------------------------
// Adjoint, FP32C.  Thread owns PQ pillars (FP64 positions in registers, scaled to
// turns by each row's plane); each row's pixels stream through LDS.

template <int PQ>
__global__ __launch_bounds__(kThreads, 2) void adj_fp32c(
    const double2* __restrict__ pos, int npil, const DevPlane* __restrict__ planes,
    const AdjRow* __restrict__ rows, const AdjBlock* __restrict__ blocks,
    const float4* __restrict__ pix, double2* __restrict__ partial) {
    __shared__ float4 tile[kThreads];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    double qx[PQ], qy[PQ];
#pragma unroll
    for (int k = 0; k < PQ; ++k) {
        const double2 v = pos[min(B.q0 + tid + k * kThreads, npil - 1)];
        qx[k] = v.x;
        qy[k] = v.y;
    }
    float sr[PQ], si[PQ], cr[PQ], ci[PQ];
#pragma unroll
    for (int k = 0; k < PQ; ++k) sr[k] = si[k] = cr[k] = ci[k] = 0.0f;

    for (int r = B.r0; r < B.r1; ++r) {
        const AdjRow rw = rows[r];
        const DevPlane& P = planes[rw.plane];
        const int nx = P.nx;
        const double kt = P.kturn;
        const double fy = P.y0 + static_cast<double>(rw.j) * P.sy;
        float z2h, z2l;
        split_sq(kt * (P.z0 + static_cast<double>(rw.j) * P.sz), z2h, z2l);
        float xq[PQ], dyz2[PQ];
#pragma unroll
        for (int k = 0; k < PQ; ++k) {
            xq[k] = static_cast<float>(kt * qx[k]);
            const float dy = static_cast<float>(kt * fy) - static_cast<float>(kt * qy[k]);
            dyz2[k] = fmaf(dy, dy, z2l) + z2h;
        }
        const float4* __restrict__ prow = pix + P.offset + static_cast<int64_t>(rw.j) * nx;
        for (int base = 0; base < nx; base += kThreads) {
            const int n = min(kThreads, nx - base);
            __syncthreads();
            if (tid < n) tile[tid] = prow[base + tid];
            __syncthreads();
            float ar[PQ], ai[PQ];
#pragma unroll
            for (int k = 0; k < PQ; ++k) ar[k] = ai[k] = 0.0f;
            for (int p = 0; p < n; ++p) {
                const float4 X = tile[p];
#pragma unroll
                for (int k = 0; k < PQ; ++k) {
                    const float dx = X.x - xq[k];
                    const float R2 = fmaf(dx, dx, dyz2[k]);
                    float iR;
                    const float th = __builtin_amdgcn_fractf(pair_turns<-1>(R2, X.z, iR));
                    const float A = X.y * iR;
                    ar[k] = fmaf(A, __builtin_amdgcn_cosf(th), ar[k]);
                    ai[k] = fmaf(A, __builtin_amdgcn_sinf(th), ai[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < PQ; ++k) {
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
    for (int k = 0; k < PQ; ++k) {
        const int q = B.q0 + tid + k * kThreads;
        if (q < npil)
            partial[B.part + q] =
                make_double2(static_cast<double>(sr[k]) - static_cast<double>(cr[k]),
                             static_cast<double>(si[k]) - static_cast<double>(ci[k]));
    }
}

// ---------------------------------------------------------------------------------
// Adjoint, FP32 and FP64 (reference pair arithmetic, transposed).  One pillar per
// thread; u rows stream through LDS.

__global__ __launch_bounds__(kThreads) void adj_fp32(
    const float4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const AdjRow* __restrict__ rows, const AdjBlock* __restrict__ blocks,
    const double2* __restrict__ u, double2* __restrict__ partial) {
    __shared__ float2 tile[kThreads];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    const int q = B.q0 + tid;
    const float4 v = pil[min(q, npil - 1)];
    float gr = 0.0f, gi = 0.0f;
    for (int r = B.r0; r < B.r1; ++r) {
        const AdjRow rw = rows[r];
        const DevPlane& P = planes[rw.plane];
        const int nx = P.nx;
        const float fy = static_cast<float>(P.y0) + static_cast<float>(rw.j) * static_cast<float>(P.sy);
        const float z = static_cast<float>(P.z0) + static_cast<float>(rw.j) * static_cast<float>(P.sz);
        const float k = static_cast<float>(P.k);
        const float dy = fy - v.y;
        const double2* __restrict__ urow = u + P.offset + static_cast<int64_t>(rw.j) * nx;
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
    if (q < npil) partial[B.part + q] = make_double2(gr, gi);
}

__global__ __launch_bounds__(kThreads) void adj_fp64(
    const double4* __restrict__ pil, int npil, const DevPlane* __restrict__ planes,
    const AdjRow* __restrict__ rows, const AdjBlock* __restrict__ blocks,
    const double2* __restrict__ u, double2* __restrict__ partial) {
    __shared__ double2 tile[kThreads];
    const AdjBlock B = blocks[blockIdx.x];
    const int tid = threadIdx.x;
    const int q = B.q0 + tid;
    const double4 v = pil[min(q, npil - 1)];
    double gr = 0.0, gi = 0.0;
    for (int r = B.r0; r < B.r1; ++r) {
        const AdjRow rw = rows[r];
        const DevPlane& P = planes[rw.plane];
        const int nx = P.nx;
        const double fy = P.y0 + rw.j * P.sy;
        const double z = P.z0 + rw.j * P.sz;
        const double dy = fy - v.y;
        const double2* __restrict__ urow = u + P.offset + static_cast<int64_t>(rw.j) * nx;
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
    if (q < npil) partial[B.part + q] = make_double2(gr, gi);
}

// Adjoint reduction: g[idx] = sum over slots of partial[slot*n + idx], slot order.
__global__ __launch_bounds__(kThreads) void adj_reduce(const double2* __restrict__ partial,
                                                       int nslots, int64_t n,
                                                       double2* __restrict__ g) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (idx >= n) return;
    double gr = 0.0, gi = 0.0;
    for (int s = 0; s < nslots; ++s) {
        const double2 v = partial[static_cast<int64_t>(s) * n + idx];
        gr += v.x;
        gi += v.y;
    }
    g[idx] = make_double2(gr, gi);
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
        // Layout: pos (double2 x npil) then ap (float2 x nsets*npil).
        double2* pos = static_cast<double2*>(out);
        float2* ap = reinterpret_cast<float2*>(pos + npil);
        stage_pillars_fp32c<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total, pos, ap);
    }
    else if (mode == 2)
        stage_pillars_fp32<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<float4*>(out));
    else
        stage_pillars_fp64<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<double4*>(out));
    return hipGetLastError();
}

hipError_t launch_forward(int mode, const void* pil, int npil, const DevPlane* planes,
                          const FwdBlock* blocks, int nblocks, double2* partial, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 0)
        fwd_fp32c<kPX, kPY><<<nblocks, kThreads, 0, s>>>(
            static_cast<const double2*>(pil),
            reinterpret_cast<const float2*>(static_cast<const double2*>(pil) + npil), npil, planes,
            blocks, partial);
    else if (mode == 2)
        fwd_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes,
                                              blocks, partial);
    else
        fwd_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes,
                                              blocks, partial);
    return hipGetLastError();
}

hipError_t launch_fwd_reduce(int restore_phase, const DevPlane* planes, const PixBlock* pblocks,
                             int nblocks, const double2* partial, double2* field,
                             double* intensity, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    fwd_reduce<<<nblocks, kThreads, 0, s>>>(planes, pblocks, partial, restore_phase, field,
                                            intensity);
    return hipGetLastError();
}

hipError_t launch_stage_pixels(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                               const double2* u, float4* pix, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    stage_pixels_fp32c<<<nblocks, kThreads, 0, s>>>(planes, pblocks, u, pix);
    return hipGetLastError();
}

hipError_t launch_adjoint(int mode, const void* pil, int npil, const DevPlane* planes,
                          const AdjRow* rows, const AdjBlock* blocks, int nblocks,
                          const void* pix_or_u, double2* partial, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 0)
        adj_fp32c<kPQ><<<nblocks, kThreads, 0, s>>>(static_cast<const double2*>(pil), npil, planes,
                                                    rows, blocks,
                                                    static_cast<const float4*>(pix_or_u), partial);
    else if (mode == 2)
        adj_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes, rows,
                                              blocks, static_cast<const double2*>(pix_or_u),
                                              partial);
    else
        adj_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes, rows,
                                              blocks, static_cast<const double2*>(pix_or_u),
                                              partial);
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const double2* partial, int nslots, int64_t n, double2* g,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    adj_reduce<<<grid_for(n), kThreads, 0, s>>>(partial, nslots, n, g);
    return hipGetLastError();
}

}  // namespace crs
