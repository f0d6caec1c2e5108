// rs_kernels.hip — gfx950 kernels for the Rayleigh–Sommerfeld far-field sum
//
//     E_p = sum_q t_q exp(i k R_pq) / R_pq,   R_pq = |(fx_p - x_q, fy_p - y_q, z_p)|
//
// (reference: src/celeris/cuda/propagate.cu:36-84 psf_kernel, :86-130 zx_kernel, and
// the FP64 CPU loop src/celeris/analysis/focal.cpp:229-233) and its adjoint
// g_q = sum_p conj(G_pq) u_p (no reference counterpart).
//
// Three arithmetic modes.  A work-group of 256 threads (4 wave64s) owns a set of
// pixels (forward) or pillars (adjoint) and streams the other side through LDS in tiles
// that every lane reads by broadcast.  Partial sums land in per-work-item slots and a
// second kernel adds them in FP64 in a fixed order, so results are bitwise repeatable.
//
// FP32C ("FP32-compensated").  Lengths are measured in turns (x' = k x / 2 pi), so a
// distance is directly a phase.  Pixels are taken in BX x BY blocks (4 or 8 each way).
// For each pillar q and block b the distance R0 from q to the block centre and the
// phase frac(R0 - z'_ref + arg t_q / 2 pi) are formed in FP64 (one v_rsq_f32 plus one
// FP64 Newton step).  Every pixel p of the block sits at a small, exactly known offset e
// from the centre, so
//     R^2 = R0^2 + D,   D = e_x (2 dx0 + e_x) + e_y (2 dy0 + e_y) + (z_p'^2 - z_c'^2)
// where D is tiny next to R0^2 and is formed in float without cancellation (each term
// is one fma against per-block constants, separable into a column part a[c] and a row
// part b[r]).  With h = 1/(2 R0), x1 = D h and y = x1 h = D / (4 R0^2):
//     R - R0 = x1 (1 - y + 2 y^2 - ...),      R0 / R = 1 - 2 y + 6 y^2 - ...
// and the pair is
//     E += |t| (k/2 pi) / R0 * (R0/R) * (cos 2 pi th, sin 2 pi th),  th = th0 + (R - R0)
// with v_cos_f32 / v_sin_f32 (which take turns): 6 FP32 ops and 2 transcendentals per
// pair at the lowest series order.  No float ever holds kR (~10^3 rad) or R^2 (~10^4
// turns^2).  Each work item chooses, from its plane and the pillars' extent, the
// cheapest series order whose worst-case pair error stays within kPairBudget of
// |t|/R (rs_internal.h: truncation, float rounding and the hardware sin/cos error); the
// host sends a plane to 1 x 1 blocks (every pair at its own FP64 R0) where even the
// highest order would not.  The factor exp(2 pi i z'_ref) is restored in FP64 by the
// reduction.
//
// FP32C work is split stream-K style: the (pixel tile, pillar) work of a forward call
// and the (pillar group, pixel strip column) work of an adjoint call are laid on one
// line and cut into pieces, one per work-group, of costs that fall along the line (the
// host's piece_bounds), so the last work-groups the hardware dispatches are short and
// every CU stays busy nearly to the end; each piece writes one partial slot.
#include "rs_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

// Scratch of the allocation-free forward entry point (crs_forward_noalloc).  Each
// device holds its own copy for the life of the process; the host's name for it is a
// stand-in, so host code uses only the device address noalloc_scratch returns.  Calls
// that overlap on one device share it and race.
static __device__ __align__(256) unsigned char scratch[crs::kNoAllocScratchBytes];

namespace crs {

// Device address of the current device's copy of `scratch`, looked up once per device
// with that device current (hipGetSymbolAddress), then cached.
hipError_t noalloc_scratch(void** ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

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

// Wave index as a scalar (every lane of a wave has the same one).
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

}  // namespace

// ---------------------------------------------------------------------------------
// Staging: float64 pillar lists -> the layout each mode's pair loop reads.

// FP32C: positions stay FP64 (pos[npil], shared by all sets; each plane scales them
// by its own k/(2 pi) when it loads a tile), {|t|, arg(t)/(2 pi)} per pillar and set,
// and per kExtTile pillars their extent {max |x|, max |y|} (one work-group's max).
// t == NULL stages positions (and extents) only.
__global__ __launch_bounds__(kThreads) void stage_pillars_fp32c(
    const double* __restrict__ px, const double* __restrict__ py, const double2* __restrict__ t,
    int npil, int64_t total, double2* __restrict__ pos, float2* __restrict__ ap,
    double2* __restrict__ ext) {
    __shared__ double2 red[kWaves];
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double mx = 0.0, my = 0.0;
    if (idx < npil) {
        const double x = px[idx], y = py[idx];
        pos[idx] = make_double2(x, y);
        mx = fabs(x);
        my = fabs(y);
    }
    if (t && idx < total) {
        const double2 v = t[idx];
        ap[idx] = make_float2(static_cast<float>(hypot(v.x, v.y)),
                              static_cast<float>(atan2(v.y, v.x) * kInv2Pi));
    }
    // Work-groups of the first npil / kThreads ... blocks cover the positions; each
    // writes its kExtTile pillars' extent.
    if (static_cast<int64_t>(blockIdx.x) * blockDim.x >= npil) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        my = fmax(my, __shfl_xor(my, o));
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double2(mx, my);
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 E = red[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
            E.x = fmax(E.x, red[w].x);
            E.y = fmax(E.y, red[w].y);
        }
        ext[blockIdx.x] = E;
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
// Forward, FP32C.  A work-group runs its segments (a tile of 64 thread tiles against a
// pillar range) one after the other.  Within a segment its 4 waves split each LDS tile
// of 256 pillars (wave w takes entries [n w / 4, n (w+1) / 4)), each lane summing its
// BX x BY pixels in float registers; every kWaveFold pillars per wave (and at the end)
// the four waves' sums are added, wave by wave in order, into FP64 running sums in LDS,
// which the segment finally stores as float pairs in its own partial slot.

// Constants of a pixel block.  Its centre (turns, FP64) is per lane; the offsets of its
// columns and rows enter D = a[c] + b[r] as
//     a[c] = m_c u + Ex(|m_c|),  u = sxt dx0,              m_c = 2c - (BX - 1)
//     b[r] = m_r v + Ey(|m_r|),  v = syt dy0 + szt zc',    m_r = 2r - (BY - 1)
// with Ex(m) = (m sxt / 2)^2 and Ey(m) = (m / 2)^2 (syt^2 + szt^2) (sxt = k sx / 2 pi etc.;
// the rows' depth change z_r'^2 - zc'^2 = m_r szt zc' + (m_r szt / 2)^2 folds into v and
// Ey).  sxt, syt and the Ex / Ey tables are the same for every lane (scalar registers),
// so a block costs one multiply per distinct |m| and an add per column or row, and no
// vector registers for its offsets.
template <int BX, int BY>
struct BlockGeom {
    double FXc, FYc, ZC2;
    float wz;                    // szt zc' (per lane in the forward)
    float zk;                    // (k / 2 pi) zc', the RS-I factor's depth
    float mz;                    // szt / (2 zc'): RS-I rows' depth ratio z'_r / zc' = 1 + m_r mz
    float sxt, syt;              // uniform
    float ex[(BX + 1) / 2];      // uniform: Ex(|m|) for |m| = 2k + (BX + 1) % 2
    float ey[(BY + 1) / 2];      // uniform: Ey(|m|)
};

__device__ __forceinline__ float uniform_f(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// The plane-wide offset constants of BX x BY blocks (sxt, syt, Ex, Ey), and wz for a
// block centred at depth zc' (turns).
template <int BX, int BY>
__device__ __forceinline__ void block_offsets_setup(const DevPlane& P, double zc,
                                                    BlockGeom<BX, BY>& G) {
    const double kt = P.kturn;
    const double sxt = kt * P.sx, syt = kt * P.sy, szt = kt * P.sz;
    G.sxt = uniform_f(static_cast<float>(sxt));
    G.syt = uniform_f(static_cast<float>(syt));
    G.wz = static_cast<float>(szt * zc);
    G.zk = static_cast<float>(kt * zc);
    G.mz = static_cast<float>(0.5 * szt / zc);
#pragma unroll
    for (int k = 0; k < (BX + 1) / 2; ++k) {
        const double m = 2 * k + (BX + 1) % 2;
        G.ex[k] = uniform_f(static_cast<float>(0.25 * m * m * sxt * sxt));
    }
#pragma unroll
    for (int k = 0; k < (BY + 1) / 2; ++k) {
        const double m = 2 * k + (BY + 1) % 2;
        G.ey[k] = uniform_f(static_cast<float>(0.25 * m * m * (syt * syt + szt * szt)));
    }
}

// a[c], b[r] of a block for a pillar at (dx0, dy0) turns from its centre.
template <int BX, int BY>
__device__ __forceinline__ void block_offsets(const BlockGeom<BX, BY>& G, float dx0, float dy0,
                                              float (&a)[BX], float (&b)[BY]) {
    const float u = G.sxt * dx0;
    const float v = fmaf(G.syt, dy0, G.wz);
#pragma unroll
    for (int c = 0; c < BX; ++c) {
        const int m = 2 * c - (BX - 1), am = m < 0 ? -m : m;
        const float mu = am == 1 ? u : static_cast<float>(am) * u;
        a[c] = m == 0 ? G.ex[0] : (m < 0 ? G.ex[am / 2] - mu : G.ex[am / 2] + mu);
    }
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int m = 2 * r - (BY - 1), am = m < 0 ? -m : m;
        const float mv = am == 1 ? v : static_cast<float>(am) * v;
        b[r] = m == 0 ? G.ey[0] : (m < 0 ? G.ey[am / 2] - mv : G.ey[am / 2] + mv);
    }
}

// One pillar against one block: th0 = frac(R0 - zeta), h = 1/(2 R0), c1 = |t| (k/2pi)
// / R0', and the separable parts of D (orders kP1A1 / kP1A2) or of x1 = D h (kP2A2):
// a[c] = 2 e_c dx0 + e_c^2, b[r] = 2 e_r dy0 + e_r^2 + dz2_r (times h for kP2A2).
//   kP1A1:  p = h^3, q = -2 c1 h^2, c = c1
//   kP1A2:  as kP1A1, and r = 6 c1 h^4
//   kP2A2:  p = 2 c1, q = -c1
//   1 x 1 blocks: c = c1 only (R = R0).
//   RSI (RS-I planes, the expansion in rs_internal.h): th0 - 1/4 + h / 2pi^2, c1 times
//   2 pi zk rf (1 + e^2/2), the phase's D-linear coefficient h (1 - e^2) (hp; w0 = 1 - e^2
//   for kP2A2), e^2 = h^2 / pi^2, the amplitude series of (1 + 4y)^-1 (p = 4 c1 for kP2A2,
//   q = -4 c1 h^2 and r = 16 c1 h^4 for the first-order phase) and each row's depth ratio mu.
template <int BX, int BY>
struct PairState {
    float th0, h, p, q, c, r;
    float hp, w0;  // RS-I only
    float a[BX], b[BY];
    float mu[BY];  // RS-I only
};

template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void pair_state(const PillarRec& T, const BlockGeom<BX, BY>& G,
                                           PairState<BX, BY>& S) {
    const double dX0 = G.FXc - T.x, dY0 = G.FYc - T.y;
    float rf;
    const double R0 = block_distance(fma(dX0, dX0, fma(dY0, dY0, G.ZC2)), rf);
    S.th0 = static_cast<float>(__builtin_amdgcn_fract(R0 - T.zeta));
    S.h = 0.5f * rf;
    float c1 = T.amp * rf;
    if constexpr (RSI && BX * BY > 1) {
        const float e2 = S.h * S.h * 0.10132118364233778f;  // h^2 / pi^2
        S.th0 += fmaf(S.h, 0.050660591821168886f, -0.25f);  // + h / (2 pi^2) - 1/4
        c1 *= 6.28318530717958648f * G.zk * rf * fmaf(0.5f, e2, 1.0f);
        S.hp = S.h * (1.0f - e2);
        S.w0 = 1.0f - e2;
#pragma unroll
        for (int r = 0; r < BY; ++r) S.mu[r] = fmaf(static_cast<float>(2 * r - (BY - 1)), G.mz, 1.0f);
    }
    if constexpr (BX * BY == 1) {
        S.c = c1;
    } else if constexpr (ORD == kP2A2) {
        S.p = RSI ? 4.0f * c1 : c1 + c1;
        S.q = -c1;
    } else {
        const float h2 = S.h * S.h;
        S.p = h2 * S.h;
        S.q = (RSI ? -4.0f : -2.0f) * c1 * h2;
        S.c = c1;
        if constexpr (ORD == kP1A2) S.r = (RSI ? 16.0f : 6.0f) * c1 * (h2 * h2);
    }
    if constexpr (BX * BY > 1) {
        block_offsets(G, static_cast<float>(dX0), static_cast<float>(dY0), S.a, S.b);
        if constexpr (ORD == kP2A2) {
#pragma unroll
            for (int c = 0; c < BX; ++c) S.a[c] *= S.h;
#pragma unroll
            for (int r = 0; r < BY; ++r) S.b[r] *= S.h;
        }
    }
}

// The BX x BY pairs of one pillar (element e = r BX + c).
//   kP1A1 (on D):  th = th0 + D (h - D h^3),  A = c1 + (-2 c1 h^2) D
//   kP1A2 (on D):  th as kP1A1,                A = c1 + D (-2 c1 h^2 + 6 c1 h^4 D)
//   kP2A2 (on x1): th = th0 + x1 w, w = 1 - y (1 - 2 y),  A = c1 (1 + y (6 y - 2))
// and E += A (cos 2 pi th, sin 2 pi th): 6, 7 and 9 FP32 ops per pair besides the two
// transcendentals.
// Phase and amplitude of pair e of a block (ORD as above).
template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void pair_phase(const PairState<BX, BY>& S, int e, float& th, float& A) {
    const int r = e / BX, c = e % BX;
    const float d = S.a[c] + S.b[r];
    if constexpr (ORD == kP2A2) {
        const float y = d * S.h;
        if constexpr (RSI) {
            const float w = fmaf(-y, fmaf(-2.0f, y, 1.0f), S.w0);
            A = fmaf(fmaf(4.0f * S.p, y, -S.p), y, -S.q);
            th = fmaf(d, w, S.th0);
        } else {
            const float w = fmaf(-y, fmaf(-2.0f, y, 1.0f), 1.0f);
            A = fmaf(fmaf(3.0f * S.p, y, -S.p), y, -S.q);
            th = fmaf(d, w, S.th0);
        }
    } else {
        th = fmaf(d, fmaf(-d, S.p, RSI ? S.hp : S.h), S.th0);
        if constexpr (ORD == kP1A2)
            A = fmaf(fmaf(S.r, d, S.q), d, S.c);
        else
            A = fmaf(S.q, d, S.c);
    }
    if constexpr (RSI) A *= S.mu[r];
}

template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void block_pairs(const PairState<BX, BY>& S, float (&ar)[BX * BY],
                                            float (&ai)[BX * BY]) {
    constexpr int N = BX * BY;
    if constexpr (N == 1) {
        const float A = S.c;
        ar[0] = fmaf(A, __builtin_amdgcn_cosf(S.th0), ar[0]);
        ai[0] = fmaf(A, __builtin_amdgcn_sinf(S.th0), ai[0]);
    } else {
        // Groups of G pairs: the group's phases, then its transcendentals, then its sums,
        // with scheduling barriers so that no sum waits on a transcendental issued just
        // before it (left to itself the compiler chains each pair's sin and cos straight
        // into its sums, measured 12% slower for 4 x 4 blocks).
        // (4 x 8 blocks: groups of 4, so the kernel fits 168 VGPRs, 3 waves per SIMD.)
        constexpr int G = N < 8 ? N : (BX == 4 && BY == 8 ? 4 : 8);
#pragma unroll
        for (int g0 = 0; g0 < N; g0 += G) {
            float th[G], A[G], cs[G], sn[G];
#pragma unroll
            for (int k = 0; k < G; ++k) pair_phase<ORD, BX, BY, RSI>(S, g0 + k, th[k], A[k]);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                cs[k] = __builtin_amdgcn_cosf(th[k]);
                sn[k] = __builtin_amdgcn_sinf(th[k]);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < G; ++k) {
                ar[g0 + k] = fmaf(A[k], cs[k], ar[g0 + k]);
                ai[g0 + k] = fmaf(A[k], sn[k], ai[g0 + k]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Pillars tile[k0 .. k1) against one block, one pillar state at a time.
template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void pillar_run(const PillarRec* tile, int k0, int k1,
                                           const BlockGeom<BX, BY>& G, float (&ar)[BX * BY],
                                           float (&ai)[BX * BY]) {
    for (int k = k0; k < k1; ++k) {
        PairState<BX, BY> st;
        pair_state<ORD, BX, BY, RSI>(tile[k], G, st);
        block_pairs<ORD, BX, BY, RSI>(st, ar, ai);
    }
}

// Pillars tile[k0 .. k1) against one pixel with the RS-I kernel (exact planes only):
// t exp(ikR)/R (z/R)(1/R - ik) = [A (cos, sin) th0] (zk/R')(1/R' - 2 pi i), with
// A = |t| k/2pi / R', zk = (k/2pi) z', R' = R k/2pi.
__device__ __forceinline__ void pillar_run_rsi(const PillarRec* tile, int k0, int k1,
                                               const BlockGeom<1, 1>& G, float (&ar)[1],
                                               float (&ai)[1]) {
    for (int k = k0; k < k1; ++k) {
        PairState<1, 1> st;
        pair_state<kP1A1>(tile[k], G, st);
        const float rf = st.h + st.h;
        const float f = G.zk * rf;
        const float fr = f * rf, fi = f * -6.28318530717958648f;
        const float cr = st.c * __builtin_amdgcn_cosf(st.th0);
        const float ci = st.c * __builtin_amdgcn_sinf(st.th0);
        ar[0] = fmaf(cr, fr, fmaf(-ci, fi, ar[0]));
        ai[0] = fmaf(ci, fr, fmaf(cr, fi, ai[0]));
    }
}

// One segment: tile T (class of block shape BX x BY within a plane blocked KBX x KBY)
// against pillars [q0, q1) at series order `order`; partial sums to `slot`
// (element e * 64 + lane).
template <int BX, int BY, int KBX, int KBY, bool RSI = false>
__device__ __forceinline__ void fwd_segment(
    const double2* __restrict__ pos, const float2* __restrict__ aps, const DevPlane& P,
    const FwdTile& T, int q0, int q1, int order, PillarRec* tile, double2* run,
    float2* __restrict__ slot) {
    constexpr int N = BX * BY;
    const int tid = threadIdx.x, wave = wave_id(), lane = tid & 63;
    const int my_tile = T.tile0 + lane;
    const bool active = my_tile < tile_count(P, T.shape, KBX, KBY);
    int ox, oy;
    tile_origin(P, T.shape, active ? my_tile : T.tile0, KBX, KBY, ox, oy);
    const double kt = P.kturn;
    const double zref = P.zref;

    BlockGeom<BX, BY> G;
    {
        const double ic = ox + (BX - 1) * 0.5, jc = oy + (BY - 1) * 0.5;
        G.FXc = kt * (P.x0 + ic * P.sx);
        G.FYc = kt * (P.y0 + jc * P.sy);
        const double zc = kt * (P.z0 + jc * P.sz);
        G.ZC2 = zc * zc;
        block_offsets_setup(P, zc, G);
    }
    const float amp_scale = static_cast<float>(kt);  // 1/R = (k/2pi) / R'

    float ar[N], ai[N];
#pragma unroll
    for (int e = 0; e < N; ++e) ar[e] = ai[e] = 0.0f;
    bool first = true;
    int since_fold = 0;
    for (int base = q0; base < q1; base += kFwdThreads) {
        const int n = min(kFwdThreads, q1 - base);
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
        const int k0 = (wave * n) / kFwdWaves, k1 = ((wave + 1) * n) / kFwdWaves;
        if constexpr (N == 1) {
            // RS-I planes only ever run in the exact-plane launch (KBX = KBY = 1); keeping
            // the branch out of the block kernels leaves their code as it was (a change
            // there moved the 4 x 8 kernel's register assignment and cost it 6% at 321^2).
            if constexpr (KBX * KBY == 1 || RSI) {
                if (P.kernel == kKernelRsi)
                    pillar_run_rsi(tile, k0, k1, G, ar, ai);
                else
                    pillar_run<kP1A1>(tile, k0, k1, G, ar, ai);
            } else {
                pillar_run<kP1A1>(tile, k0, k1, G, ar, ai);
            }
        } else {
            if (order == kP1A1)
                pillar_run<kP1A1, BX, BY, RSI>(tile, k0, k1, G, ar, ai);
            else if (order == kP1A2)
                pillar_run<kP1A2, BX, BY, RSI>(tile, k0, k1, G, ar, ai);
            else
                pillar_run<kP2A2, BX, BY, RSI>(tile, k0, k1, G, ar, ai);
        }
        since_fold += kFwdThreads / kFwdWaves;
        if (since_fold >= kWaveFold || base + kFwdThreads >= q1) {
            // Fold the four waves' float sums into the FP64 running sums, in wave order.
#pragma unroll
            for (int w = 0; w < kFwdWaves; ++w) {
                __syncthreads();
                if (wave == w) {
#pragma unroll
                    for (int e = 0; e < N; ++e) {
                        double2 v = make_double2(0.0, 0.0);
                        if (!first || w != 0) v = run[e * 64 + lane];
                        v.x += static_cast<double>(ar[e]);
                        v.y += static_cast<double>(ai[e]);
                        run[e * 64 + lane] = v;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < N; ++e) ar[e] = ai[e] = 0.0f;
            first = false;
            since_fold = 0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (N * 64 + kFwdThreads - 1) / kFwdThreads; ++k) {
        const int i = tid + k * kFwdThreads;
        if (N * 64 % kFwdThreads == 0 || i < N * 64) {
            const double2 v = run[i];
            slot[i] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
        }
    }
}

// Waves per SIMD each forward instance is compiled for (a register budget of 512 /
// waves VGPRs; the accumulators alone take 2 BX BY registers).  4 x 4: 4 waves (122
// VGPRs).  4 x 8: 3 waves (168 VGPRs, groups of 4 pairs; its first-order loop keeps
// every value in registers, the rarer higher orders reload three FP64 values per pillar
// from scratch): 2.59 against 2.65 ms for the best 2-wave shape (8 x 4) at 91,809 pillars
// x 321^2.  8 x 4: 2 waves (194 VGPRs); at 3 waves its loops spill and it ran 4% slower.
constexpr int fwd_waves_per_simd(int bx, int by) { return bx * by <= 16 ? 4 : (bx == 4 ? 3 : 2); }

// KBX x KBY: the launch's block shape; its work-groups take the four tile classes of
// the planes blocked that way (and the 1 x 1 tiles of exact planes).  <1, 1>: calls
// whose planes are all exact.
template <int KBX, int KBY, bool RSI>
__device__ __forceinline__ void fwd_fp32c_body(
    const double2* __restrict__ pos, const float2* __restrict__ ap,
    const double2* __restrict__ ext, int npil, const DevPlane* __restrict__ planes,
    const FwdTile* __restrict__ tiles, const FwdSeg* __restrict__ segs,
    const int32_t* __restrict__ wg_seg, float2* __restrict__ partial, int64_t slot_elems) {
    __shared__ PillarRec tile[kFwdThreads];
    __shared__ double2 run[KBX * KBY * 64];
    const int s0 = wg_seg[blockIdx.x], s1 = wg_seg[blockIdx.x + 1];
    for (int s = s0; s < s1; ++s) {
        const FwdSeg S = segs[s];
        const FwdTile T = tiles[S.tile];
        const DevPlane& P = planes[T.plane];
        const float2* __restrict__ aps = ap + static_cast<int64_t>(P.tset) * npil;
        float2* __restrict__ slot = partial + static_cast<int64_t>(s) * slot_elems;
        if constexpr (KBX * KBY == 1) {
            fwd_segment<1, 1, 1, 1>(pos, aps, P, T, S.q0, S.q1, kP1A1, tile, run, slot);
        } else {
            // Series order from the segment's pillar extent (the staged extents of the
            // kExtTile runs it covers) and the plane's obliquity limits for the full
            // KBX x KBY block; the 1 x KBY and KBX x 1 edge blocks share every offset
            // bound of the full block and take the same order.  The extent loop runs on
            // uniform addresses (scalar loads, no lane shuffles: shuffle indices held
            // across the segment loop used to spill to scratch).
            int order = kP2A2;
            if (T.shape != 3) {
                double mx = 0.0, my = 0.0;
                const int e0 = S.q0 / kExtTile, e1 = (S.q1 - 1) / kExtTile;
                for (int i = e0; i <= e1; ++i) {
                    const double2 v = ext[i];
                    mx = fmax(mx, v.x);
                    my = fmax(my, v.y);
                }
                order = __builtin_amdgcn_readfirstlane(fp32c_series_order(P, mx, my));
            }
            switch (T.shape) {
                case 0:
                    fwd_segment<KBX, KBY, KBX, KBY, RSI>(pos, aps, P, T, S.q0, S.q1, order, tile, run, slot);
                    break;
                case 1:
                    fwd_segment<1, KBY, KBX, KBY, RSI>(pos, aps, P, T, S.q0, S.q1, order, tile, run, slot);
                    break;
                case 2:
                    fwd_segment<KBX, 1, KBX, KBY, RSI>(pos, aps, P, T, S.q0, S.q1, order, tile, run, slot);
                    break;
                default:
                    fwd_segment<1, 1, KBX, KBY, RSI>(pos, aps, P, T, S.q0, S.q1, kP1A1, tile, run, slot);
                    break;
            }
        }
    }
}

template <int KBX, int KBY>
__global__ __launch_bounds__(kFwdThreads, fwd_waves_per_simd(KBX, KBY)) void fwd_fp32c(
    const double2* __restrict__ pos, const float2* __restrict__ ap,
    const double2* __restrict__ ext, int npil, const DevPlane* __restrict__ planes,
    const FwdTile* __restrict__ tiles, const FwdSeg* __restrict__ segs,
    const int32_t* __restrict__ wg_seg, float2* __restrict__ partial, int64_t slot_elems) {
    fwd_fp32c_body<KBX, KBY, false>(pos, ap, ext, npil, planes, tiles, segs, wg_seg, partial,
                                    slot_elems);
}

// RS-I planes in KBX x KBY blocks (a launch of their own, so the reference kernel's
// instances keep their code).
template <int KBX, int KBY>
__global__ __launch_bounds__(kFwdThreads, fwd_waves_per_simd(KBX, KBY)) void fwd_fp32c_rsi(
    const double2* __restrict__ pos, const float2* __restrict__ ap,
    const double2* __restrict__ ext, int npil, const DevPlane* __restrict__ planes,
    const FwdTile* __restrict__ tiles, const FwdSeg* __restrict__ segs,
    const int32_t* __restrict__ wg_seg, float2* __restrict__ partial, int64_t slot_elems) {
    fwd_fp32c_body<KBX, KBY, true>(pos, ap, ext, npil, planes, tiles, segs, wg_seg, partial,
                                   slot_elems);
}

// ---------------------------------------------------------------------------------
// Pair kernels G(R) of the FP32 / FP64 paths.  RSI = false: exp(ikR)/R, the reference's
// (focal.cpp:232, propagate.cu:71-77).  RSI = true: Rayleigh-Sommerfeld I,
// exp(ikR)/R * (z/R) * (1/R - ik) = exp(ikR) * (z/R^2) * (1/R - ik), opt-in per plane.

template <bool RSI>
__device__ __forceinline__ void green64(double dx, double dy, double z, double k, double& gr,
                                        double& gi) {
    const double R = sqrt(dx * dx + dy * dy + z * z);
    double s, c;
    sincos(k * R, &s, &c);
    const double rho = 1.0 / R;
    if constexpr (RSI) {
        const double a = z * rho * rho;           // z / R^2
        const double fr = a * rho, fi = -a * k;   // (z / R^2) (1/R - ik)
        gr = c * fr - s * fi;
        gi = s * fr + c * fi;
    } else {
        gr = rho * c;
        gi = rho * s;
    }
}

template <bool RSI>
__device__ __forceinline__ void green32(float dx, float dy, float z, float k, float& gr,
                                        float& gi) {
    const float R = sqrtf(dx * dx + dy * dy + z * z);
    float s, c;
    sincosf(k * R, &s, &c);
    const float inv = 1.0f / R;
    if constexpr (RSI) {
        const float a = z * inv * inv;
        const float fr = a * inv, fi = -a * k;
        gr = c * fr - s * fi;
        gi = s * fr + c * fi;
    } else {
        gr = c * inv;
        gi = s * inv;
    }
}

// One LDS tile of pillars against one pixel (forward): E += t G; the reference
// kernel in the reference's own operation order, (t_r c - t_i s) / R (propagate.cu:76-77).
template <bool RSI>
__device__ __forceinline__ void fwd_tile32(const float4* tile, int n, float fx, float fy, float z,
                                           float k, float& er, float& ei) {
    for (int q = 0; q < n; ++q) {
        const float4 Q = tile[q];
        if constexpr (RSI) {
            float gr, gi;
            green32<true>(fx - Q.x, fy - Q.y, z, k, gr, gi);
            er += Q.z * gr - Q.w * gi;
            ei += Q.z * gi + Q.w * gr;
        } else {
            const float dx = fx - Q.x, dy = fy - Q.y;
            const float R = sqrtf(dx * dx + dy * dy + z * z);
            float s, c;
            sincosf(k * R, &s, &c);
            const float inv = 1.0f / R;
            er += (Q.z * c - Q.w * s) * inv;
            ei += (Q.z * s + Q.w * c) * inv;
        }
    }
}

template <bool RSI>
__device__ __forceinline__ void fwd_tile64(const double4* tile, int n, double fx, double fy,
                                           double z, double k, double& er, double& ei) {
    for (int q = 0; q < n; ++q) {
        const double4 Q = tile[q];
        double gr, gi;
        green64<RSI>(fx - Q.x, fy - Q.y, z, k, gr, gi);
        er += Q.z * gr - Q.w * gi;
        ei += Q.z * gi + Q.w * gr;
    }
}

// One LDS tile of pixel cotangents against one pillar (adjoint): g += conj(G) u.
template <bool RSI>
__device__ __forceinline__ void adj_tile32(const float2* tile, int n, float x0, float sx, int base,
                                           float qx, float dy, float z, float k, float& gr,
                                           float& gi) {
    for (int p = 0; p < n; ++p) {
        const float2 U = tile[p];
        const float fx = x0 + static_cast<float>(base + p) * sx;
        if constexpr (RSI) {
            float cr, ci;
            green32<true>(fx - qx, dy, z, k, cr, ci);
            gr += cr * U.x + ci * U.y;
            gi += cr * U.y - ci * U.x;
        } else {
            const float dx = fx - qx;
            const float R = sqrtf(dx * dx + dy * dy + z * z);
            float s, c;
            sincosf(k * R, &s, &c);
            const float inv = 1.0f / R;
            gr += (c * U.x + s * U.y) * inv;
            gi += (c * U.y - s * U.x) * inv;
        }
    }
}

template <bool RSI>
__device__ __forceinline__ void adj_tile64(const double2* tile, int n, double x0, double sx,
                                           int base, double qx, double dy, double z, double k,
                                           double& gr, double& gi) {
    for (int p = 0; p < n; ++p) {
        const double2 U = tile[p];
        const double fx = x0 + (base + p) * sx;
        double cr, ci;
        green64<RSI>(fx - qx, dy, z, k, cr, ci);
        gr += cr * U.x + ci * U.y;
        gi += cr * U.y - ci * U.x;
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
        if (P.kernel == kKernelRsi)
            fwd_tile32<true>(tile, n, fx, fy, z, k, er, ei);
        else
            fwd_tile32<false>(tile, n, fx, fy, z, k, er, ei);
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
        if (P.kernel == kKernelRsi)
            fwd_tile64<true>(tile, n, fx, fy, z, k, er, ei);
        else
            fwd_tile64<false>(tile, n, fx, fy, z, k, er, ei);
    }
    if (active) partial[B.part + pix] = make_double2(er, ei);
}

// ---------------------------------------------------------------------------------
// Forward reduction (FP32 / FP64): add the plane's pillar-chunk partials, write field
// and/or |E|^2 (restore_phase: turn by exp(2 pi i z'_ref) first).  A work-group
// takes kReduceCols pixels; wave w adds the w-th quarter of the slots, and wave 0 adds
// the four quarter sums in order (a fixed order: results are bitwise repeatable).

// T: the partial type (double2).
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
// Forward reduction, FP32C: for every element (e * 64 + lane) of a tile, the sum over
// the tile's segments' partial slots (FP64), turned by exp(2 pi i z'_ref) and written at
// the element's pixel.  A work-group takes 64 elements of one tile; wave w sums the w-th
// quarter of the tile's slots (eight loads in flight), and wave 0 adds the four quarter
// sums in wave order, so the result does not depend on timing.

namespace {
__device__ __forceinline__ void sum_slots(const float2* __restrict__ src, int64_t stride, int k0,
                                          int k1, double& er, double& ei) {
    constexpr int U = 8;
    int k = k0;
    for (; k + U <= k1; k += U) {
        float2 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) v[j] = src[static_cast<int64_t>(k + j) * stride];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            er += v[j].x;
            ei += v[j].y;
        }
    }
    for (; k < k1; ++k) {
        const float2 v = src[static_cast<int64_t>(k) * stride];
        er += v.x;
        ei += v.y;
    }
}
}  // namespace

__global__ __launch_bounds__(kThreads) void fwd_reduce_fp32c(
    const DevPlane* __restrict__ planes, const FwdTile* __restrict__ tiles,
    const FwdRed* __restrict__ reds, const float2* __restrict__ partial, int64_t slot_elems,
    double2* __restrict__ field, double* __restrict__ intensity) {
    __shared__ double2 red[kWaves][64];
    const FwdRed R = reds[blockIdx.x];
    const FwdTile T = tiles[R.tile];
    const DevPlane& P = planes[T.plane];
    const int wave = wave_id(), lane = threadIdx.x & 63;
    const int i = R.el0 + lane;  // element of the tile (el0 is a multiple of 64)
    double er = 0.0, ei = 0.0;
    if (i < T.nel * 64) {
        const int k0 = (wave * T.nseg) / kWaves, k1 = ((wave + 1) * T.nseg) / kWaves;
        sum_slots(partial + static_cast<int64_t>(T.seg0) * slot_elems + i, slot_elems, k0, k1,
                  er, ei);
    }
    red[wave][lane] = make_double2(er, ei);
    __syncthreads();
    if (wave != 0 || i >= T.nel * 64) return;
    er = red[0][lane].x;
    ei = red[0][lane].y;
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
        er += red[w][lane].x;
        ei += red[w][lane].y;
    }
    const int e = i >> 6;  // element e of thread tile T.tile0 + lane
    const int t = T.tile0 + lane;
    if (t >= tile_count(P, T.shape, P.bx, P.by)) return;
    int ox, oy;
    tile_origin(P, T.shape, t, P.bx, P.by, ox, oy);
    const int bw = (T.shape == 0 || T.shape == 2) ? P.bx : 1;
    const int c = e % bw, r = e / bw;
    double sn, cs;
    sincos(kTwoPi * (P.zref - floor(P.zref)), &sn, &cs);
    const double xr = er * cs - ei * sn;
    const double xi = er * sn + ei * cs;
    const int64_t o = P.offset + static_cast<int64_t>(oy + r) * P.nx + (ox + c);
    if (field) field[o] = make_double2(xr, xi);
    if (intensity) intensity[o] = xr * xr + xi * xi;
}

// ---------------------------------------------------------------------------------
// Adjoint staging (FP32C).  Pixel records are laid out strip by strip (kStrip rows),
// column by column, kStrip records per column: record (strip u, column i, row r) sits at
// sbase + (u*nxp + i)*kStrip + r, so an LDS tile of 64 columns x kStrip rows is one
// contiguous run.  Record = {2U, -U, arg(u)/2pi - z'_ref mod 1, 0} with U = |u| k/2pi, or
// on planes whose blocks take the separable form {Re w, Im w, -Re w, 0} with
// w = u (k/2pi) exp(-2 pi i z'_ref); padding (rows >= ny, columns >= nx) is all zero and
// contributes nothing.

__global__ __launch_bounds__(kThreads) void stage_pixels_fp32c(
    const DevPlane* __restrict__ planes, const PixBlock* __restrict__ pblocks,
    const double2* __restrict__ u, float4* __restrict__ pix) {
    const PixBlock B = pblocks[blockIdx.x];
    const DevPlane& P = planes[B.plane];
    const int nrec = P.nquads * P.nxp * kStrip;
    const int rix = B.pix0 + static_cast<int>(threadIdx.x);
    if (rix >= nrec) return;
    const int strip = rix / (P.nxp * kStrip);
    const int rem = rix - strip * P.nxp * kStrip;
    const int i = rem / kStrip, j = strip * kStrip + rem % kStrip;
    float4 rec = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < P.nx && j < P.ny) {
        const double2 v = u[P.offset + static_cast<int64_t>(j) * P.nx + i];
        if (P.sep) {
            // Separable planes: u (k/2pi) exp(-2 pi i z'_ref) as {Re, Im, -Re}.
            double sn, cs;
            sincos(kTwoPi * (P.zref - floor(P.zref)), &sn, &cs);
            const double wr = P.kturn * (v.x * cs + v.y * sn);
            const double wi = P.kturn * (v.y * cs - v.x * sn);
            rec = make_float4(static_cast<float>(wr), static_cast<float>(wi),
                              static_cast<float>(-wr), 0.0f);
        } else {
            double chi = atan2(v.y, v.x) * kInv2Pi - P.zref;
            chi -= rint(chi);
            const float U = static_cast<float>(hypot(v.x, v.y) * P.kturn);
            rec = make_float4(2.0f * U, -U, static_cast<float>(chi), 0.0f);
        }
    }
    pix[P.sbase + rix] = rec;
}

// ---------------------------------------------------------------------------------
// Adjoint, FP32C.  A thread owns one pillar (FP64 position in registers) of a group of
// 256; a work-group runs its segments (a group against a run of strip columns) one after
// the other.  A strip's records stream through LDS 64 columns at a time (as three float
// arrays, so a column of records is a few vector reads) and are consumed in BX x BY
// blocks (BY = 4 takes a strip as two halves; 1 x 1 for exact planes, in a launch of
// their own).  Per pair
//     conj(G) u = U / R0' (R0/R) exp(2 pi i (chi - th0 - (R - R0))),   U = |u| k/2pi,
// with the forward's expansion of R - R0 and R0/R; each block's sum (kept per column,
// so the accumulation chains stay short) is turned by exp(-2 pi i th0) and scaled by
// 1/R0' once.  Sums over LDS tiles are Kahan-compensated.

template <int BX, int BY>
struct AdjState {
    float th0, h, rf;
    float zk;  // RS-I planes: (k/2pi) z' of the pixel row (1 x 1 blocks)
    bool rsi;
    float w0;      // RS-I blocks: 1 - h^2/pi^2
    float mu[BY];  // RS-I blocks: the rows' depth ratios z'_r / zc'
    float a[BX], b[BY];
};

// RSI: RS-I planes in blocks (the expansion in rs_internal.h): th0 - 1/4 + h / 2pi^2, the
// block's scale rf times 2 pi zk rf (1 + e^2/2), e^2 = h^2 / pi^2.
template <int BX, int BY, bool RSI = false>
__device__ __forceinline__ void adj_state(double dX0, double dY0, double zref,
                                          const BlockGeom<BX, BY>& G, AdjState<BX, BY>& S) {
    float rf;
    const double R0 = block_distance(fma(dX0, dX0, fma(dY0, dY0, G.ZC2)), rf);
    S.th0 = static_cast<float>(__builtin_amdgcn_fract(R0 - zref));
    S.rf = rf;
    S.h = 0.5f * rf;
    S.zk = 0.0f;
    S.rsi = false;
    if constexpr (RSI && BX * BY > 1) {
        const float e2 = S.h * S.h * 0.10132118364233778f;  // h^2 / pi^2
        S.th0 += fmaf(S.h, 0.050660591821168886f, -0.25f);  // + h / (2 pi^2) - 1/4
        S.rf = rf * (6.28318530717958648f * G.zk * rf * fmaf(0.5f, e2, 1.0f));
        S.w0 = 1.0f - e2;
#pragma unroll
        for (int r = 0; r < BY; ++r) S.mu[r] = fmaf(static_cast<float>(2 * r - (BY - 1)), G.mz, 1.0f);
    }
    if constexpr (BX * BY > 1) {
        block_offsets(G, static_cast<float>(dX0), static_cast<float>(dY0), S.a, S.b);
#pragma unroll
        for (int c = 0; c < BX; ++c) S.a[c] *= S.h;
#pragma unroll
        for (int r = 0; r < BY; ++r) S.b[r] *= S.h;
    }
}

// The pairs of one pillar with one block whose first record (column 0, row 0 of the
// block) is LDS record k0 (record of column c, row r: k0 + c kStrip + r); adds the
// block's turned sum to (gr, gi).
//   kP1A1:  w = 1 - x1 h,             A = 2U w - U,                phase chi - x1 w
//   kP2A2:  w = 1 - y (1 - 2 y),      A = U (1 + y (6 y - 2)),     phase chi - x1 w
// (y = x1 h); 1 x 1 blocks use chi - th0 and A = U directly.  RSI: w starts from w0, and
// A = U (1 - 4y) (kP1A1) or U (1 + y (16 y - 4)) (kP2A2), times the row's mu.
template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void adj_pairs(const AdjState<BX, BY>& S, const float* sU2,
                                          const float* sUm, const float* sChi, int k0,
                                          float& gr, float& gi) {
    if constexpr (BX * BY == 1) {
        const float phi = sChi[k0] - S.th0;
        const float A = S.rf * (sU2[k0] + sUm[k0]);
        if (S.rsi) {
            // RS-I: conj(G) u carries conj((zk/R')(1/R' - 2 pi i)) = f (rf, +2 pi).
            const float f = S.zk * S.rf;
            const float fr = f * S.rf, fi = f * 6.28318530717958648f;
            const float cr = A * __builtin_amdgcn_cosf(phi), ci = A * __builtin_amdgcn_sinf(phi);
            gr = fmaf(cr, fr, fmaf(-ci, fi, gr));
            gi = fmaf(ci, fr, fmaf(cr, fi, gi));
        } else {
            gr = fmaf(A, __builtin_amdgcn_cosf(phi), gr);
            gi = fmaf(A, __builtin_amdgcn_sinf(phi), gi);
        }
    } else {
        float br[BX], bi[BX];
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            br[c] = 0.0f;
            bi[c] = 0.0f;
#pragma unroll
            for (int r = 0; r < BY; ++r) {
                const int k = k0 + c * kStrip + r;
                const float U2 = sU2[k], Um = sUm[k], chi = sChi[k];
                const float x1 = S.a[c] + S.b[r];
                float w, A;
                if constexpr (RSI) {
                    if constexpr (ORD == kP2A2) {
                        const float y = x1 * S.h;
                        w = fmaf(-y, fmaf(-2.0f, y, 1.0f), S.w0);
                        A = fmaf(fmaf(8.0f * U2, y, -2.0f * U2), y, -Um);
                    } else {
                        w = fmaf(-x1, S.h, S.w0);
                        A = fmaf(U2, fmaf(-x1, S.h + S.h, 1.0f), Um);
                    }
                    A *= S.mu[r];
                } else if constexpr (ORD == kP2A2) {
                    const float y = x1 * S.h;
                    w = fmaf(-y, fmaf(-2.0f, y, 1.0f), 1.0f);
                    A = fmaf(fmaf(3.0f * U2, y, -U2), y, -Um);
                } else {
                    w = fmaf(-x1, S.h, 1.0f);
                    A = fmaf(U2, w, Um);
                }
                const float phi = fmaf(-x1, w, chi);
                br[c] = fmaf(A, __builtin_amdgcn_cosf(phi), br[c]);
                bi[c] = fmaf(A, __builtin_amdgcn_sinf(phi), bi[c]);
            }
        }
        float sr = br[0], si = bi[0];
#pragma unroll
        for (int c = 1; c < BX; ++c) {
            sr += br[c];
            si += bi[c];
        }
        const float cr = S.rf * __builtin_amdgcn_cosf(S.th0);
        const float sn = S.rf * __builtin_amdgcn_sinf(S.th0);
        gr = fmaf(sr, cr, fmaf(si, sn, gr));
        gi = fmaf(si, cr, fmaf(-sr, sn, gi));
    }
}

// Columns [0, ncol) of one LDS tile (strip starting at row j0, tile starting at plane
// column cb) against the thread's pillar (qx, qy in um).
template <int ORD, int BX, int BY, bool RSI = false>
__device__ __forceinline__ void adj_tile(const float* sU2, const float* sUm, const float* sChi,
                                         int ncol, const DevPlane& P, int j0, int cb, double qx,
                                         double qy, float& ar, float& ai) {
    const double kt = P.kturn;
    BlockGeom<BX, BY> G;
    for (int sr0 = 0; sr0 < kStrip; sr0 += BY) {
        if (BY < kStrip && j0 + sr0 >= P.ny) break;  // all-padding rows (uniform)
        const double jc = j0 + sr0 + (BY - 1) * 0.5;
        G.FYc = kt * (P.y0 + jc * P.sy);
        const double zc = kt * (P.z0 + jc * P.sz);
        G.ZC2 = zc * zc;
        block_offsets_setup(P, zc, G);
        const double dY0 = fma(-kt, qy, G.FYc);
        for (int g = 0; g < ncol; g += BX) {
            G.FXc = kt * (P.x0 + (cb + g + (BX - 1) * 0.5) * P.sx);
            AdjState<BX, BY> S;
            adj_state<BX, BY, RSI>(fma(-kt, qx, G.FXc), dY0, P.zref, G, S);
            if constexpr (BX * BY == 1) {
                S.rsi = P.kernel == kKernelRsi;
                S.zk = G.zk;
            }
            adj_pairs<ORD, BX, BY, RSI>(S, sU2, sUm, sChi, g * kStrip + sr0, ar, ai);
        }
    }
}

// Separable form (planes the host marks `sep`; the expansion and its bound: rs_internal.h).
// For one pillar and one block whose first record is k0,
//     sum_(c,r) conj(G) u = sum_c Ubar_c S_c,   S_c = sum_r conj(V_r + a_c W_r) w_cr,
//     Ubar_c = rf (1 - 2y + 6y^2) exp(-2 pi i (th0 + a_c h (1 - y + 2y^2))),  y = a_c h^2,
// V_r = (1 - 2y + 6y^2) exp(2 pi i b_r h (1 - y + 2y^2)) (y = b_r h^2), W_r = V_r l1(b_r) and
// w the staged records {Re, Im, -Re} of u (k/2pi) exp(-2 pi i z'_ref): per pair two FMAs
// for V_r + a_c W_r and four for S_c; per pillar and block one v_cos_f32 / v_sin_f32 pair
// per row and per column.  RAG: a block cut by the plane's last column or row, whose
// columns c >= nc and rows r >= nr (all-zero padding records) are skipped, at the cost of
// uniform branches the full blocks do not take.
//
// The series factors are written with inline constants (+-1, +-2, +-4 and literals in
// two-operand forms): a constant the three-operand FMA encoding cannot hold goes to a
// scalar register, and an FMA with a scalar operand issues at ~4.5 cycles per wave64
// instead of ~3.1.

// (1 - 2y + 6y^2, 1 - y + 2y^2): amplitude and phase factors of F at y = D h^2.
__device__ __forceinline__ void sep_factors(float y, float& amp, float& ph) {
#pragma clang fp contract(off)  // keep y * 6 a multiply by a literal, not an FMA taking 6 from an SGPR
    const float y6 = y * 6.0f;
    amp = fmaf(y, y6 - 2.0f, 1.0f);
    ph = fmaf(y, fmaf(2.0f, y, -1.0f), 1.0f);
}

template <int BX, int BY, bool RAG>
__device__ __forceinline__ void adj_pairs_sep(double dX0, double dY0, double zref,
                                              const BlockGeom<BX, BY>& G, const float* sRe,
                                              const float* sIm, const float* sNRe, int k0,
                                              int nc, int nr, float& gr, float& gi) {
    float rf;
    const double R0 = block_distance(fma(dX0, dX0, fma(dY0, dY0, G.ZC2)), rf);
    const float th0 = static_cast<float>(__builtin_amdgcn_fract(R0 - zref));
    const float h = 0.5f * rf, h2 = h * h;
    float a[BX], b[BY];
    block_offsets(G, static_cast<float>(dX0), static_cast<float>(dY0), a, b);
    const float p1r = 8.0f * h2 * h2, p1i = -12.5663706143591729f * h2 * h;
    float vr[BY], vi[BY], wr[BY], wi[BY];
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        if (RAG && r >= nr) continue;
        const float y = b[r] * h2;
        float g, pf;
        sep_factors(y, g, pf);
        const float ph = b[r] * h * pf;
        const float cs = g * __builtin_amdgcn_cosf(ph), sn = g * __builtin_amdgcn_sinf(ph);
        const float l1r = b[r] * p1r * fmaf(-4.0f, y, 1.0f);
        const float l1i = b[r] * p1i * fmaf(-2.0f, y, 1.0f - y);  // (1 - 3y)
        vr[r] = cs;
        vi[r] = sn;
        wr[r] = fmaf(cs, l1r, -sn * l1i);
        wi[r] = fmaf(cs, l1i, sn * l1r);
    }
#pragma unroll
    for (int c = 0; c < BX; ++c) {
        if (RAG && c >= nc) continue;
        float sr = 0.0f, si = 0.0f;
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            if (RAG && r >= nr) continue;
            const int k = k0 + c * kStrip + r;
            const float ur = sRe[k], ui = sIm[k], nur = sNRe[k];
            const float yr = fmaf(a[c], wr[r], vr[r]);
            const float yi = fmaf(a[c], wi[r], vi[r]);
            sr = fmaf(yr, ur, sr);
            sr = fmaf(yi, ui, sr);
            si = fmaf(yr, ui, si);
            si = fmaf(yi, nur, si);
        }
        const float y = a[c] * h2;
        float g, pf;
        sep_factors(y, g, pf);
        const float ph = fmaf(a[c] * h, pf, th0);
        g *= rf;
        const float cu = g * __builtin_amdgcn_cosf(ph), su = g * __builtin_amdgcn_sinf(ph);
        gr = fmaf(cu, sr, fmaf(su, si, gr));
        gi = fmaf(cu, si, fmaf(-su, sr, gi));
    }
}

// Columns [0, ncol) of one LDS tile in the separable form (BY = kStrip: one block row per
// strip).
template <int BX, int BY>
__device__ __forceinline__ void adj_tile_sep(const float* sRe, const float* sIm, const float* sNRe,
                                             int ncol, const DevPlane& P, int j0, int cb,
                                             double qx, double qy, float& ar, float& ai) {
    static_assert(BY == kStrip, "separable adjoint blocks span a whole strip");
    const double kt = P.kturn;
    BlockGeom<BX, BY> G;
    const double jc = j0 + (BY - 1) * 0.5;
    G.FYc = kt * (P.y0 + jc * P.sy);
    const double zc = kt * (P.z0 + jc * P.sz);
    G.ZC2 = zc * zc;
    block_offsets_setup(P, zc, G);
    const double dY0 = fma(-kt, qy, G.FYc);
    const int nr = min(BY, P.ny - j0);  // real rows of the strip (uniform)
    for (int g = 0; g < ncol; g += BX) {
        G.FXc = kt * (P.x0 + (cb + g + (BX - 1) * 0.5) * P.sx);
        const double dX0 = fma(-kt, qx, G.FXc);
        const int nc = min(BX, P.nx - (cb + g));  // real columns of the block (uniform)
        if (nc == BX && nr == BY)
            adj_pairs_sep<BX, BY, false>(dX0, dY0, P.zref, G, sRe, sIm, sNRe, g * kStrip, BX, BY,
                                         ar, ai);
        else
            adj_pairs_sep<BX, BY, true>(dX0, dY0, P.zref, G, sRe, sIm, sNRe, g * kStrip, nc, nr,
                                        ar, ai);
    }
}

// The adjoint kernels' body (RSI: RS-I planes in blocks).  Every adjoint kernel below is
// held to 128 VGPRs, 4 waves per SIMD: VALU issue slows with fewer waves (FP32 FMAs
// measured 3.1 cycles per wave64 instruction per SIMD at 4 waves, 3.5 at 2 and 3, 5.3 at 1).
template <int BX, int BY, bool RSI>
__device__ __forceinline__ void adj_fp32c_body(
    const double* __restrict__ px, const double* __restrict__ py, int npil,
    const DevPlane* __restrict__ planes, const AdjStrip* __restrict__ strips,
    const AdjSeg* __restrict__ segs, const int32_t* __restrict__ wg_seg,
    const float4* __restrict__ pix, float2* __restrict__ partial) {
    __shared__ float sU2[kAdjCols * kStrip], sUm[kAdjCols * kStrip], sChi[kAdjCols * kStrip];
    __shared__ double2 red[kWaves];
    const int tid = threadIdx.x;
    const int s0 = wg_seg[blockIdx.x], s1 = wg_seg[blockIdx.x + 1];
    for (int s = s0; s < s1; ++s) {
        const AdjSeg S = segs[s];
        const int q = min(S.q0 + tid, npil - 1);
        const double qx = px[q], qy = py[q];
        // The group's own pillar extent decides the series order of each plane.
        double2 E = make_double2(0.0, 0.0);
        if constexpr (BX * BY > 1) {
            double mx = fabs(qx), my = fabs(qy);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                mx = fmax(mx, __shfl_xor(mx, o));
                my = fmax(my, __shfl_xor(my, o));
            }
            __syncthreads();  // the previous segment has read red
            if ((tid & 63) == 0) red[tid >> 6] = make_double2(mx, my);
            __syncthreads();
            E = red[0];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) {
                E.x = fmax(E.x, red[w].x);
                E.y = fmax(E.y, red[w].y);
            }
        }
        float sr = 0.0f, si = 0.0f, cr = 0.0f, ci = 0.0f;
        int prev_plane = -1, order = kP1A1;
        for (int u = S.u0; u <= S.u1; ++u) {
            const AdjStrip st = strips[u];
            const DevPlane& P = planes[st.plane];
            const int ncols = (P.nx + BX - 1) / BX * BX;
            const int cs = u == S.u0 ? S.c0 : 0;
            const int ce = u == S.u1 ? S.c1 : ncols;
            if (cs >= ce) continue;
            if constexpr (BX * BY > 1) {
                if (st.plane != prev_plane) {
                    // The adjoint plan's limits count rows up to the strips' padded end
                    // (obl2 = obl1: no kP1A2 in the adjoint).
                    DevPlane Q = P;
                    Q.ny = P.nquads * kStrip;
                    order = __builtin_amdgcn_readfirstlane(fp32c_series_order(Q, E.x, E.y));
                    if (BX * BY == 64 && P.sep) order = kSep;
                    prev_plane = st.plane;
                }
            }
            const float4* __restrict__ prow =
                pix + P.sbase + static_cast<int64_t>(st.j0 / kStrip) * P.nxp * kStrip;
            for (int cb = cs; cb < ce; cb += kAdjCols) {
                const int ncol = min(kAdjCols, ce - cb);
                __syncthreads();
                for (int k = tid; k < ncol * kStrip; k += kThreads) {
                    const float4 v = prow[cb * kStrip + k];
                    sU2[k] = v.x;
                    sUm[k] = v.y;
                    sChi[k] = v.z;
                }
                __syncthreads();
                float ar = 0.0f, ai = 0.0f;
                int ord = order;
                if constexpr (BX * BY == 64 && !RSI) {
                    if (order == kSep) {
                        adj_tile_sep<BX, BY>(sU2, sUm, sChi, ncol, P, st.j0, cb, qx, qy, ar, ai);
                        ord = -1;
                    }
                }
                if (ord == kP1A1)
                    adj_tile<kP1A1, BX, BY, RSI>(sU2, sUm, sChi, ncol, P, st.j0, cb, qx, qy, ar, ai);
                else if (ord >= 0)
                    adj_tile<kP2A2, BX, BY, RSI>(sU2, sUm, sChi, ncol, P, st.j0, cb, qx, qy, ar, ai);
                const float yr = ar - cr;
                const float tr = sr + yr;
                cr = (tr - sr) - yr;
                sr = tr;
                const float yi = ai - ci;
                const float tq = si + yi;
                ci = (tq - si) - yi;
                si = tq;
            }
        }
        partial[static_cast<int64_t>(S.slot) * kThreads + tid] =
            make_float2(static_cast<float>(static_cast<double>(sr) - static_cast<double>(cr)),
                        static_cast<float>(static_cast<double>(si) - static_cast<double>(ci)));
    }
}

template <int BX, int BY>
__global__ __launch_bounds__(kThreads, 4) void adj_fp32c(
    const double* __restrict__ px, const double* __restrict__ py, int npil,
    const DevPlane* __restrict__ planes, const AdjStrip* __restrict__ strips,
    const AdjSeg* __restrict__ segs, const int32_t* __restrict__ wg_seg,
    const float4* __restrict__ pix, float2* __restrict__ partial) {
    adj_fp32c_body<BX, BY, false>(px, py, npil, planes, strips, segs, wg_seg, pix, partial);
}

// RS-I planes in BX x BY blocks (a launch of their own).
template <int BX, int BY>
__global__ __launch_bounds__(kThreads, 4) void adj_fp32c_rsi(
    const double* __restrict__ px, const double* __restrict__ py, int npil,
    const DevPlane* __restrict__ planes, const AdjStrip* __restrict__ strips,
    const AdjSeg* __restrict__ segs, const int32_t* __restrict__ wg_seg,
    const float4* __restrict__ pix, float2* __restrict__ partial) {
    adj_fp32c_body<BX, BY, true>(px, py, npil, planes, strips, segs, wg_seg, pix, partial);
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
                const float x0 = static_cast<float>(P.x0), sx = static_cast<float>(P.sx);
                if (P.kernel == kKernelRsi)
                    adj_tile32<true>(tile, n, x0, sx, base, v.x, dy, z, k, gr, gi);
                else
                    adj_tile32<false>(tile, n, x0, sx, base, v.x, dy, z, k, gr, gi);
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
                if (P.kernel == kKernelRsi)
                    adj_tile64<true>(tile, n, P.x0, P.sx, base, v.x, dy, z, P.k, gr, gi);
                else
                    adj_tile64<false>(tile, n, P.x0, P.sx, base, v.x, dy, z, P.k, gr, gi);
            }
        }
    }
    if (q < npil) partial[B.part + q] = make_double2(gr, gi);
}

// Adjoint reduction (FP32 / FP64): g[idx] = sum over slots of partial[slot*n + idx]; kReduceCols
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

// Adjoint reduction, FP32C: g[s npil + q] = sum over the slots listed for q's group
// (slot_off[s ngroups + grp] .. [+1]) of partial[slot * kThreads + q - grp kThreads]
// (FP64).  A work-group takes 64 pillars (a quarter of a group); wave w sums the w-th
// quarter of the group's slot list (eight loads in flight) and wave 0 adds the four
// quarter sums in wave order, so the result does not depend on timing.
__global__ __launch_bounds__(kThreads) void adj_reduce_fp32c(
    const float2* __restrict__ partial, const int32_t* __restrict__ slot_off,
    const int32_t* __restrict__ slots, int npil, int ngroups, double2* __restrict__ g) {
    __shared__ double2 red[kWaves][64];
    const int quarter = blockIdx.x & 3;
    const int sg = blockIdx.x >> 2;  // s * ngroups + grp
    const int s = sg / ngroups, grp = sg - s * ngroups;
    const int wave = wave_id(), lane = threadIdx.x & 63;
    const int lq = quarter * 64 + lane;  // pillar within the group
    const int k0 = slot_off[sg], n = slot_off[sg + 1] - k0;
    const int a0 = k0 + (wave * n) / kWaves, a1 = k0 + ((wave + 1) * n) / kWaves;
    double gr = 0.0, gi = 0.0;
    constexpr int U = 8;
    int k = a0;
    for (; k + U <= a1; k += U) {
        int sl[U];
        float2 v[U];
#pragma unroll
        for (int j = 0; j < U; ++j) sl[j] = slots[k + j];
#pragma unroll
        for (int j = 0; j < U; ++j) v[j] = partial[static_cast<int64_t>(sl[j]) * kThreads + lq];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            gr += v[j].x;
            gi += v[j].y;
        }
    }
    for (; k < a1; ++k) {
        const float2 v = partial[static_cast<int64_t>(slots[k]) * kThreads + lq];
        gr += v.x;
        gi += v.y;
    }
    red[wave][lane] = make_double2(gr, gi);
    __syncthreads();
    const int q = grp * kThreads + lq;
    if (wave != 0 || q >= npil) return;
    gr = red[0][lane].x;
    gi = red[0][lane].y;
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
        gr += red[w][lane].x;
        gi += red[w][lane].y;
    }
    g[static_cast<int64_t>(s) * npil + q] = make_double2(gr, gi);
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

template <int KBX, int KBY, bool RSI = false>
void launch_fwd_shape(const Fp32cFwdArgs& a, hipStream_t s) {
    if constexpr (RSI)
        fwd_fp32c_rsi<KBX, KBY><<<a.nwg, kFwdThreads, 0, s>>>(a.pos, a.ap, a.ext, a.npil, a.planes,
                                                           a.tiles, a.segs, a.wg_seg, a.partial,
                                                           a.slot_elems);
    else
        fwd_fp32c<KBX, KBY><<<a.nwg, kFwdThreads, 0, s>>>(a.pos, a.ap, a.ext, a.npil, a.planes,
                                                       a.tiles, a.segs, a.wg_seg, a.partial,
                                                       a.slot_elems);
}

template <int BX, int BY, bool RSI = false>
void launch_adj_shape(const Fp32cAdjArgs& a, hipStream_t s) {
    if constexpr (RSI)
        adj_fp32c_rsi<BX, BY><<<a.nwg, kThreads, 0, s>>>(a.px, a.py, a.npil, a.planes, a.strips,
                                                      a.segs, a.wg_seg, a.pix, a.partial);
    else
        adj_fp32c<BX, BY><<<a.nwg, kThreads, 0, s>>>(a.px, a.py, a.npil, a.planes, a.strips,
                                                  a.segs, a.wg_seg, a.pix, a.partial);
}
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
            reinterpret_cast<float2*>(base + L.ap), reinterpret_cast<double2*>(base + L.ext));
    } else if (mode == 2) {
        stage_pillars_fp32<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<float4*>(out));
    } else {
        stage_pillars_fp64<<<grid_for(total), kThreads, 0, s>>>(px, py, tt, npil, total,
                                                                 static_cast<double4*>(out));
    }
    return hipGetLastError();
}

// FP32C forward pair kernel of block shape bx x by: 4 x 4, 4 x 8, 8 x 4, or 1 x 1
// (exact planes only); rsi: RS-I planes in 4 x 4 blocks.
hipError_t launch_forward_fp32c(int bx, int by, bool rsi, const Fp32cFwdArgs& a, hipStream_t s) {
    if (a.nwg <= 0) return hipSuccess;
    if (rsi) {
        if (bx != 4 || by != 4) return hipErrorInvalidValue;
        launch_fwd_shape<4, 4, true>(a, s);
    } else if (bx == 4 && by == 4) launch_fwd_shape<4, 4>(a, s);
    else if (bx == 4 && by == 8) launch_fwd_shape<4, 8>(a, s);
    else if (bx == 8 && by == 4) launch_fwd_shape<8, 4>(a, s);
    else if (bx == 1 && by == 1) launch_fwd_shape<1, 1>(a, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fwd_reduce_fp32c(const DevPlane* planes, const FwdTile* tiles,
                                   const FwdRed* reds, int nred, const float2* partial,
                                   int64_t slot_elems, double2* field, double* intensity,
                                   hipStream_t s) {
    if (nred <= 0) return hipSuccess;
    fwd_reduce_fp32c<<<nred, kThreads, 0, s>>>(planes, tiles, reds, partial, slot_elems, field,
                                               intensity);
    return hipGetLastError();
}

hipError_t launch_adjoint_fp32c(int bx, int by, bool rsi, const Fp32cAdjArgs& a, hipStream_t s) {
    if (a.nwg <= 0) return hipSuccess;
    if (rsi) {
        if (bx != 4 || by != 4) return hipErrorInvalidValue;
        launch_adj_shape<4, 4, true>(a, s);
    } else if (bx == 4 && by == 4) launch_adj_shape<4, 4>(a, s);
    else if (bx == 4 && by == 8) launch_adj_shape<4, 8>(a, s);
    else if (bx == 8 && by == 4) launch_adj_shape<8, 4>(a, s);
    else if (bx == 8 && by == 8) launch_adj_shape<8, 8>(a, s);
    else if (bx == 1 && by == 1) launch_adj_shape<1, 1>(a, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_adj_reduce_fp32c(const float2* partial, const int32_t* slot_off,
                                   const int32_t* slots, int npil, int ngroups, int64_t n,
                                   double2* g, hipStream_t s) {
    if (n <= 0 || npil <= 0) return hipSuccess;
    const int64_t nsets = n / npil;
    adj_reduce_fp32c<<<static_cast<unsigned>(nsets * ngroups * 4), kThreads, 0, s>>>(
        partial, slot_off, slots, npil, ngroups, g);
    return hipGetLastError();
}

// Work-groups of the FP32C pair kernel of this shape that one CU holds at once (0 if
// unknown): the stream-K planner sizes its launches from it.
int fp32c_resident_per_cu(bool adjoint, int bx, int by, bool rsi) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    auto occ = [&](const void* fn, int threads) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, 0);
    };
    if (rsi) {
        if (adjoint) occ(reinterpret_cast<const void*>(adj_fp32c_rsi<4, 4>), kThreads);
        else occ(reinterpret_cast<const void*>(fwd_fp32c_rsi<4, 4>), kFwdThreads);
    } else if (!adjoint) {
        if (bx == 4 && by == 4) occ(reinterpret_cast<const void*>(fwd_fp32c<4, 4>), kFwdThreads);
        else if (bx == 4 && by == 8) occ(reinterpret_cast<const void*>(fwd_fp32c<4, 8>), kFwdThreads);
        else if (bx == 8 && by == 4) occ(reinterpret_cast<const void*>(fwd_fp32c<8, 4>), kFwdThreads);
        else occ(reinterpret_cast<const void*>(fwd_fp32c<1, 1>), kFwdThreads);
    } else {
        if (bx == 4 && by == 4) occ(reinterpret_cast<const void*>(adj_fp32c<4, 4>), kThreads);
        else if (bx == 4 && by == 8) occ(reinterpret_cast<const void*>(adj_fp32c<4, 8>), kThreads);
        else if (bx == 8 && by == 4) occ(reinterpret_cast<const void*>(adj_fp32c<8, 4>), kThreads);
        else if (bx == 8 && by == 8) occ(reinterpret_cast<const void*>(adj_fp32c<8, 8>), kThreads);
        else occ(reinterpret_cast<const void*>(adj_fp32c<1, 1>), kThreads);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// Profiler names of the FP32C pair kernels.
const char* fp32c_kernel_name(bool adjoint, int bx, int by, bool rsi) {
    if (rsi) return adjoint ? "crs::adj_fp32c_rsi<4, 4>" : "crs::fwd_fp32c_rsi<4, 4>";
    static const char* fwd[] = {"crs::fwd_fp32c<4, 4>", "crs::fwd_fp32c<4, 8>",
                                "crs::fwd_fp32c<8, 4>", "crs::fwd_fp32c<8, 8>",
                                "crs::fwd_fp32c<1, 1>"};  // <8, 8>: adjoint only
    static const char* adj[] = {"crs::adj_fp32c<4, 4>", "crs::adj_fp32c<4, 8>",
                                "crs::adj_fp32c<8, 4>", "crs::adj_fp32c<8, 8>",
                                "crs::adj_fp32c<1, 1>"};
    int i = 4;
    if (bx == 4 && by == 4) i = 0;
    else if (bx == 4 && by == 8) i = 1;
    else if (bx == 8 && by == 4) i = 2;
    else if (bx == 8 && by == 8) i = 3;
    return adjoint ? adj[i] : fwd[i];
}

hipError_t launch_forward(int mode, const void* pil, int npil, const DevPlane* planes,
                          const FwdBlock* blocks, int nblocks, void* partial_v, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 2) {
        fwd_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes,
                                              blocks, static_cast<double2*>(partial_v));
    } else {
        fwd_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes,
                                              blocks, static_cast<double2*>(partial_v));
    }
    return hipGetLastError();
}

hipError_t launch_fwd_reduce(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                             const void* partial, double2* field, double* intensity,
                             hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
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

hipError_t launch_adjoint(int mode, const void* pil, int npil, const DevPlane* planes,
                          const AdjQuad* quads, const AdjBlock* blocks, int nblocks,
                          const void* u, void* partial_v, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == 2) {
        adj_fp32<<<nblocks, kThreads, 0, s>>>(static_cast<const float4*>(pil), npil, planes, quads,
                                              blocks, static_cast<const double2*>(u),
                                              static_cast<double2*>(partial_v));
    } else {
        adj_fp64<<<nblocks, kThreads, 0, s>>>(static_cast<const double4*>(pil), npil, planes, quads,
                                              blocks, static_cast<const double2*>(u),
                                              static_cast<double2*>(partial_v));
    }
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const void* partial, int nslots, int64_t n, double2* g,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = static_cast<unsigned>((n + kReduceCols - 1) / kReduceCols);
    adj_reduce<double2><<<grid, kThreads, 0, s>>>(static_cast<const double2*>(partial), nslots, n,
                                                  g);
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
