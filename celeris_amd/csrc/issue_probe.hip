// VALU issue costs on this GPU, per SIMD: shader-clock cycles per wave64 instruction of
// the kinds the FP32C pair loops are made of (FP32 FMA, v_sin_f32, FP64 FMA), and of the
// loops' own mix (16 transcendentals + 48 FP32 FMAs, grouped as the kernels group them).
//
// A one-round launch of W waves per SIMD on average; each wave reads the shader clock
// (s_memtime) around its loop and its hardware id (XCC, SE, CU, SIMD).  Per SIMD, the
// cycles per instruction are (last end - first start) / (its waves x iterations x
// instructions per iteration); the figures reported are the medians over the SIMDs that
// held the most waves (the dispatcher does not spread waves perfectly evenly).
//
// Build: hipcc --offload-arch=gfx950 -O3 (packed FP32 off, as for the library).
// Usage: issue_probe [iterations]   -> one JSON line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

namespace {

constexpr int kW = 8;          // waves per SIMD
constexpr int kThreadsPerWg = 256;

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "issue_probe: %s at line %d\n", hipGetErrorString(e_), \
                         __LINE__);                                                 \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

// MODE 0: 48 FP32 FMAs per iteration (16 chains of 3), each taking the kernel argument s
// (a scalar register) as an operand; 1: 16 v_sin_f32 (16 chains); 2: 16 v_sin_f32 then 48
// FP32 FMAs (grouped, as the pair loops group them); 3: 48 FP64 FMAs (8 chains of 6);
// 4: 48 FP32 adds; 5: 48 FP32 multiplies; 6: 48 FP32 FMAs on vector operands only (as in
// the pair loops: acc += x y with x, y per lane).  At most 64 VGPRs, so that kW waves fit
// on a SIMD.
template <int MODE>
__global__ __launch_bounds__(kThreadsPerWg, kW) void probe(long long* cycles, float* sink,
                                                           int iters, float s) {
    unsigned hw_id, xcc_id;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw_id));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc_id));
    float x[16], acc[16];
    double xd[8], ad[8];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        x[i] = threadIdx.x * 1e-3f + i * 0.01f;
        acc[i] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        xd[i] = x[i];
        ad[i] = 0.0;
    }
    const double sd = s;
    const long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
        if constexpr (MODE == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[i] = fmaf(x[i], s, acc[i]);
                x[i] = fmaf(x[i], s, 0.001f);
                acc[i] = fmaf(x[i], acc[i], s);
            }
        } else if constexpr (MODE == 1) {
#pragma unroll
            for (int i = 0; i < 16; ++i) x[i] = __builtin_amdgcn_sinf(x[i]);
        } else if constexpr (MODE == 2) {
            float t[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) t[i] = __builtin_amdgcn_sinf(x[i]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[i & 7] = fmaf(t[i], s, acc[i & 7]);
                x[i] = fmaf(x[i], s, 0.001f);
                acc[i & 7] = fmaf(x[i], acc[i & 7], s);
            }
        } else if constexpr (MODE == 4) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[i] = acc[i] + x[i];
                x[i] = x[i] + s;
                acc[i] = acc[i] + x[i];
            }
        } else if constexpr (MODE == 5) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[i] = acc[i] * x[i];
                x[i] = x[i] * s;
                acc[i] = acc[i] * x[i];
            }
        } else if constexpr (MODE == 6) {
            // Every operand a vector register; x changes each iteration, so no product
            // is loop-invariant.
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[i] = fmaf(x[i], x[(i + 1) & 15], acc[i]);
                acc[i] = fmaf(x[(i + 2) & 15], x[(i + 3) & 15], acc[i]);
                x[i] = fmaf(x[i], x[(i + 5) & 15], x[(i + 7) & 15]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    ad[i] = fma(xd[i], sd, ad[i]);
                    xd[i] = fma(xd[i], sd, 0.001);
                    ad[i] = fma(xd[i], ad[i], sd);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    const long long t1 = clock64();
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) r += acc[i] + x[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) r += static_cast<float>(ad[i] + xd[i]);
    const int gid = blockIdx.x * kThreadsPerWg + threadIdx.x;
    sink[gid] = r;
    if ((threadIdx.x & 63) == 0) {
        // SIMD key: XCC id, then HW_ID bits 4..15 (SIMD, pipe, CU, SH, SE)
        const long long key = (static_cast<long long>(xcc_id & 0xf) << 12) | ((hw_id >> 4) & 0xfff);
        cycles[3 * (gid >> 6) + 0] = t0;
        cycles[3 * (gid >> 6) + 1] = t1;
        cycles[3 * (gid >> 6) + 2] = key;
    }
}

// Operand placement, with fixed registers (inline asm, VALU only): 48 instructions per
// iteration on 16 accumulators v40..v55 (each updated every 16th instruction); the first
// multiplicands are four registers of one bank (register number mod 4) in v8..v23, the
// second four of one bank in v24..v39.  BANKED 0: v_fmac_f32 with the multiplicands in
// different banks; 1: in the same bank; 2: the first multiplicand an SGPR; 3: v_mul_f32
// with both operands in one bank; 4: v_mov_b32_dpp quad_perm [1, 0, 3, 2]; 5: one
// dependent chain (every FMA updates v40).  The registers stay below v64, so 8 waves fit
// on a SIMD.  One round: 16 instructions, accumulator i (v40 + i) taking v(a + 4j) x
// v(b + 4k).
#define CRS_FMAC_DIFFER \
    "v_fmac_f32 v40, v9, v26\n" \
    "v_fmac_f32 v41, v13, v30\n" \
    "v_fmac_f32 v42, v17, v34\n" \
    "v_fmac_f32 v43, v21, v38\n" \
    "v_fmac_f32 v44, v9, v30\n" \
    "v_fmac_f32 v45, v13, v34\n" \
    "v_fmac_f32 v46, v17, v38\n" \
    "v_fmac_f32 v47, v21, v26\n" \
    "v_fmac_f32 v48, v9, v34\n" \
    "v_fmac_f32 v49, v13, v38\n" \
    "v_fmac_f32 v50, v17, v26\n" \
    "v_fmac_f32 v51, v21, v30\n" \
    "v_fmac_f32 v52, v9, v38\n" \
    "v_fmac_f32 v53, v13, v26\n" \
    "v_fmac_f32 v54, v17, v30\n" \
    "v_fmac_f32 v55, v21, v34\n"
#define CRS_FMAC_SAME \
    "v_fmac_f32 v40, v10, v26\n" \
    "v_fmac_f32 v41, v14, v30\n" \
    "v_fmac_f32 v42, v18, v34\n" \
    "v_fmac_f32 v43, v22, v38\n" \
    "v_fmac_f32 v44, v10, v30\n" \
    "v_fmac_f32 v45, v14, v34\n" \
    "v_fmac_f32 v46, v18, v38\n" \
    "v_fmac_f32 v47, v22, v26\n" \
    "v_fmac_f32 v48, v10, v34\n" \
    "v_fmac_f32 v49, v14, v38\n" \
    "v_fmac_f32 v50, v18, v26\n" \
    "v_fmac_f32 v51, v22, v30\n" \
    "v_fmac_f32 v52, v10, v38\n" \
    "v_fmac_f32 v53, v14, v26\n" \
    "v_fmac_f32 v54, v18, v30\n" \
    "v_fmac_f32 v55, v22, v34\n"
#define CRS_FMAC_SGPR \
    "v_fmac_f32 v40, s8, v26\n" \
    "v_fmac_f32 v41, s8, v30\n" \
    "v_fmac_f32 v42, s8, v34\n" \
    "v_fmac_f32 v43, s8, v38\n" \
    "v_fmac_f32 v44, s8, v30\n" \
    "v_fmac_f32 v45, s8, v34\n" \
    "v_fmac_f32 v46, s8, v38\n" \
    "v_fmac_f32 v47, s8, v26\n" \
    "v_fmac_f32 v48, s8, v34\n" \
    "v_fmac_f32 v49, s8, v38\n" \
    "v_fmac_f32 v50, s8, v26\n" \
    "v_fmac_f32 v51, s8, v30\n" \
    "v_fmac_f32 v52, s8, v38\n" \
    "v_fmac_f32 v53, s8, v26\n" \
    "v_fmac_f32 v54, s8, v30\n" \
    "v_fmac_f32 v55, s8, v34\n"
#define CRS_MUL_SAME \
    "v_mul_f32 v40, v10, v26\n" \
    "v_mul_f32 v41, v14, v30\n" \
    "v_mul_f32 v42, v18, v34\n" \
    "v_mul_f32 v43, v22, v38\n" \
    "v_mul_f32 v44, v10, v30\n" \
    "v_mul_f32 v45, v14, v34\n" \
    "v_mul_f32 v46, v18, v38\n" \
    "v_mul_f32 v47, v22, v26\n" \
    "v_mul_f32 v48, v10, v34\n" \
    "v_mul_f32 v49, v14, v38\n" \
    "v_mul_f32 v50, v18, v26\n" \
    "v_mul_f32 v51, v22, v30\n" \
    "v_mul_f32 v52, v10, v38\n" \
    "v_mul_f32 v53, v14, v26\n" \
    "v_mul_f32 v54, v18, v30\n" \
    "v_mul_f32 v55, v22, v34\n"
#define CRS_DPP \
    "v_mov_b32_dpp v40, v9 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v41, v13 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v42, v17 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v43, v21 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v44, v9 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v45, v13 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v46, v17 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v47, v21 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v48, v9 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v49, v13 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v50, v17 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v51, v21 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v52, v9 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v53, v13 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v54, v17 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_mov_b32_dpp v55, v21 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define CRS_CHAIN \
    "v_fmac_f32 v40, v9, v26\n" \
    "v_fmac_f32 v40, v13, v30\n" \
    "v_fmac_f32 v40, v17, v34\n" \
    "v_fmac_f32 v40, v21, v38\n" \
    "v_fmac_f32 v40, v9, v30\n" \
    "v_fmac_f32 v40, v13, v34\n" \
    "v_fmac_f32 v40, v17, v38\n" \
    "v_fmac_f32 v40, v21, v26\n" \
    "v_fmac_f32 v40, v9, v34\n" \
    "v_fmac_f32 v40, v13, v38\n" \
    "v_fmac_f32 v40, v17, v26\n" \
    "v_fmac_f32 v40, v21, v30\n" \
    "v_fmac_f32 v40, v9, v38\n" \
    "v_fmac_f32 v40, v13, v26\n" \
    "v_fmac_f32 v40, v17, v30\n" \
    "v_fmac_f32 v40, v21, v34\n"
#define CRS_CLOBBER                                                                         \
    "v8", "v9", "v10", "v11", "v12", "v13", "v14", "v15", "v16", "v17", "v18", "v19", "v20", \
        "v21", "v22", "v23", "v24", "v25", "v26", "v27", "v28", "v29", "v30", "v31", "v32",  \
        "v33", "v34", "v35", "v36", "v37", "v38", "v39", "v40", "v41", "v42", "v43", "v44",  \
        "v45", "v46", "v47", "v48", "v49", "v50", "v51", "v52", "v53", "v54", "v55", "s8"

template <int BANKED>
__global__ __launch_bounds__(kThreadsPerWg, kW) void probe_banked(long long* cycles, float* sink,
                                                                  int iters, float s) {
    unsigned hw_id, xcc_id;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw_id));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc_id));
    const float seed = threadIdx.x * 1e-3f + s;
    // Every register the blocks read gets a value first.
    asm volatile(
        "v_mov_b32 v8, %0\n v_mov_b32 v9, %0\n v_mov_b32 v10, %0\n v_mov_b32 v11, %0\n"
        "v_mov_b32 v12, %0\n v_mov_b32 v13, %0\n v_mov_b32 v14, %0\n v_mov_b32 v15, %0\n"
        "v_mov_b32 v16, %0\n v_mov_b32 v17, %0\n v_mov_b32 v18, %0\n v_mov_b32 v19, %0\n"
        "v_mov_b32 v20, %0\n v_mov_b32 v21, %0\n v_mov_b32 v22, %0\n v_mov_b32 v23, %0\n"
        "v_mov_b32 v24, %0\n v_mov_b32 v25, %0\n v_mov_b32 v26, %0\n v_mov_b32 v27, %0\n"
        "v_mov_b32 v28, %0\n v_mov_b32 v29, %0\n v_mov_b32 v30, %0\n v_mov_b32 v31, %0\n"
        "v_mov_b32 v32, %0\n v_mov_b32 v33, %0\n v_mov_b32 v34, %0\n v_mov_b32 v35, %0\n"
        "v_mov_b32 v36, %0\n v_mov_b32 v37, %0\n v_mov_b32 v38, %0\n v_mov_b32 v39, %0\n"
        "v_mov_b32 v40, 0\n v_mov_b32 v41, 0\n v_mov_b32 v42, 0\n v_mov_b32 v43, 0\n"
        "v_mov_b32 v44, 0\n v_mov_b32 v45, 0\n v_mov_b32 v46, 0\n v_mov_b32 v47, 0\n"
        "v_mov_b32 v48, 0\n v_mov_b32 v49, 0\n v_mov_b32 v50, 0\n v_mov_b32 v51, 0\n"
        "v_mov_b32 v52, 0\n v_mov_b32 v53, 0\n v_mov_b32 v54, 0\n v_mov_b32 v55, 0\n"
        "s_mov_b32 s8, 0x3f7fbe77\n" ::"v"(seed)
        : CRS_CLOBBER);
    const long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
        if constexpr (BANKED == 0)  // banks: multiplicands 1 and 2
            asm volatile(CRS_FMAC_DIFFER CRS_FMAC_DIFFER
                             CRS_FMAC_DIFFER ::: CRS_CLOBBER);
        else if constexpr (BANKED == 1)  // both multiplicands in bank 2
            asm volatile(CRS_FMAC_SAME CRS_FMAC_SAME
                             CRS_FMAC_SAME ::: CRS_CLOBBER);
        else if constexpr (BANKED == 2)
            asm volatile(CRS_FMAC_SGPR CRS_FMAC_SGPR
                             CRS_FMAC_SGPR ::: CRS_CLOBBER);
        else if constexpr (BANKED == 3)
            asm volatile(CRS_MUL_SAME CRS_MUL_SAME
                             CRS_MUL_SAME ::: CRS_CLOBBER);
        else if constexpr (BANKED == 4)
            asm volatile(CRS_DPP CRS_DPP
                             CRS_DPP ::: CRS_CLOBBER);
        else
            asm volatile(CRS_CHAIN CRS_CHAIN
                             CRS_CHAIN ::: CRS_CLOBBER);
    }
    const long long t1 = clock64();
    float r;
    asm volatile("v_add_f32 %0, v40, v41" : "=v"(r) : : CRS_CLOBBER);
    const int gid = blockIdx.x * kThreadsPerWg + threadIdx.x;
    sink[gid] = r;
    if ((threadIdx.x & 63) == 0) {
        const long long key = (static_cast<long long>(xcc_id & 0xf) << 12) | ((hw_id >> 4) & 0xfff);
        cycles[3 * (gid >> 6) + 0] = t0;
        cycles[3 * (gid >> 6) + 1] = t1;
        cycles[3 * (gid >> 6) + 2] = key;
    }
}

// Matrix cores beside the VALU.  MX 0: v_mfma_f32_32x32x2_f32, MX 1:
// v_mfma_f32_16x16x1_4b_f32 (accumulators in AGPRs, one dependent chain), NM of them per
// iteration, then NV x 16 vector-operand FP32 FMAs (16 chains); with both, each wave
// alternates the two, so the SIMD overlaps them only if one wave's MFMAs can run beside
// another's VALU work.
typedef float crs_f16v __attribute__((ext_vector_type(16)));
template <int MX, int NM, int NV>
__global__ __launch_bounds__(kThreadsPerWg, 4) void probe_mfma(long long* cycles, float* sink,
                                                              int iters, float s) {
    unsigned hw_id, xcc_id;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw_id));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc_id));
    const int l = threadIdx.x;
    crs_f16v acc = {};
    const float a = l * 1e-3f + s, b = 0.5f + l * 1e-4f;
    float x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = l * 1e-3f + i;
    const long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            if constexpr (MX == 0)
                asm volatile("v_mfma_f32_32x32x2_f32 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
            else
                asm volatile("v_mfma_f32_16x16x1_4b_f32 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int i = 0; i < 16; ++i) x[i] = fmaf(x[i], x[(i + 5) & 15], x[(i + 2) & 15]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    const long long t1 = clock64();
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) r += acc[i] + x[i];
    const int gid = blockIdx.x * kThreadsPerWg + threadIdx.x;
    sink[gid] = r;
    if ((threadIdx.x & 63) == 0) {
        const long long key = (static_cast<long long>(xcc_id & 0xf) << 12) | ((hw_id >> 4) & 0xfff);
        cycles[3 * (gid >> 6) + 0] = t0;
        cycles[3 * (gid >> 6) + 1] = t1;
        cycles[3 * (gid >> 6) + 2] = key;
    }
}

struct Result {
    double med, lo, hi;  // over the fullest SIMDs
    int waves, simds;    // waves on those SIMDs, and how many SIMDs held that many
};

template <int MODE>
Result run(int nwg, int iters, int ninstr, long long* d_rec, float* d_sink) {
    const int nwaves = nwg * kThreadsPerWg / 64;
    // MODE >= 200: probe_mfma<MX, NM, NV> with MODE = 200 + 100 MX + 10 (NM / 16) + NV / 8
    // (NM 0 or 16, NV 0 or 8); MODE >= 100: probe_banked<MODE - 100>.
    if constexpr (MODE >= 200) {
        constexpr int MX = (MODE - 200) / 100, NM = ((MODE / 10) % 10) * 16, NV = (MODE % 10) * 8;
        probe_mfma<MX, NM, NV><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, 8, 0.5f);  // warm-up
        CHECK(hipGetLastError());
        probe_mfma<MX, NM, NV><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, iters, 0.5f);
    } else if constexpr (MODE >= 100) {
        probe_banked<MODE - 100><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, 8, 0.5f);  // warm-up
        CHECK(hipGetLastError());
        probe_banked<MODE - 100><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, iters, 0.5f);
    } else {
        probe<MODE><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, 8, 0.999f);  // warm-up
        CHECK(hipGetLastError());
        probe<MODE><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, iters, 0.999f);
    }
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    std::vector<long long> r(3 * static_cast<size_t>(nwaves));
    CHECK(hipMemcpy(r.data(), d_rec, r.size() * sizeof(long long), hipMemcpyDeviceToHost));
    struct Simd {
        long long t0 = -1, t1 = 0;
        int n = 0;
    };
    std::map<long long, Simd> simd;
    for (int w = 0; w < nwaves; ++w) {
        Simd& S = simd[r[3 * w + 2]];
        S.t0 = S.n == 0 ? r[3 * w] : std::min(S.t0, r[3 * w]);
        S.t1 = std::max(S.t1, r[3 * w + 1]);
        ++S.n;
    }
    int nmax = 0;
    for (const auto& kv : simd) nmax = std::max(nmax, kv.second.n);
    std::vector<double> c;
    for (const auto& kv : simd)
        if (kv.second.n == nmax)
            c.push_back(static_cast<double>(kv.second.t1 - kv.second.t0) /
                        (static_cast<double>(nmax) * iters * ninstr));
    std::sort(c.begin(), c.end());
    return {c[c.size() / 2], c.front(), c.back(), nmax, static_cast<int>(c.size())};
}

}  // namespace

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 4000;
    if (iters <= 0) {
        std::fprintf(stderr, "issue_probe: iterations must be positive\n");
        return 2;
    }
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int simds = prop.multiProcessorCount * 4;
    const int nwg = simds * kW * 64 / kThreadsPerWg;  // one round: kW waves on every SIMD
    long long* d_cyc = nullptr;
    float* d_sink = nullptr;
    CHECK(hipMalloc(&d_cyc, static_cast<size_t>(nwg) * (kThreadsPerWg / 64) * 3 * sizeof(long long)));
    CHECK(hipMalloc(&d_sink, static_cast<size_t>(nwg) * kThreadsPerWg * sizeof(float)));
    const Result fma32 = run<0>(nwg, iters, 48, d_cyc, d_sink);
    const Result sin32 = run<1>(nwg, iters, 16, d_cyc, d_sink);
    const Result mix = run<2>(nwg, iters, 64, d_cyc, d_sink);
    const Result fma64 = run<3>(nwg, iters, 48, d_cyc, d_sink);
    const Result add32 = run<4>(nwg, iters, 48, d_cyc, d_sink);
    const Result mul32 = run<5>(nwg, iters, 48, d_cyc, d_sink);
    // Vector-operand FMAs with 1, 2, 3, 4 and kW waves on every SIMD (one round each).
    const int occ[5] = {1, 2, 3, 4, kW};
    Result fmav[5];
    for (int i = 0; i < 5; ++i)
        fmav[i] = run<6>(simds * occ[i] * 64 / kThreadsPerWg, iters, 48, d_cyc, d_sink);
    // Operand placement (probe_banked), at 4 waves per SIMD (the pair kernels' occupancy)
    // and kW.
    const char* bnames[6] = {"fmac_banks_differ", "fmac_same_bank", "fmac_sgpr",
                             "mul_same_bank", "dpp_mov", "fmac_one_chain"};
    Result banked[6][2];
    for (int o = 0; o < 2; ++o) {
        const int w = o == 0 ? 4 : kW;
        const int g = simds * w * 64 / kThreadsPerWg;
        banked[0][o] = run<100>(g, iters, 48, d_cyc, d_sink);
        banked[1][o] = run<101>(g, iters, 48, d_cyc, d_sink);
        banked[2][o] = run<102>(g, iters, 48, d_cyc, d_sink);
        banked[3][o] = run<103>(g, iters, 48, d_cyc, d_sink);
        banked[4][o] = run<104>(g, iters, 48, d_cyc, d_sink);
        banked[5][o] = run<105>(g, iters, 48, d_cyc, d_sink);
    }
    // Matrix cores beside the VALU, at 4 waves per SIMD: cycles per wave-iteration per SIMD
    // (16 MFMAs; 128 FMAs; both).
    const int g4 = simds * 4 * 64 / kThreadsPerWg, mi = iters / 20 > 0 ? iters / 20 : 1;
    const Result mx32 = run<210>(g4, mi, 1, d_cyc, d_sink), vx32 = run<201>(g4, mi, 1, d_cyc, d_sink),
                 bx32 = run<211>(g4, mi, 1, d_cyc, d_sink);
    const Result mx16 = run<310>(g4, mi, 1, d_cyc, d_sink), vx16 = run<301>(g4, mi, 1, d_cyc, d_sink),
                 bx16 = run<311>(g4, mi, 1, d_cyc, d_sink);
    CHECK(hipFree(d_cyc));
    CHECK(hipFree(d_sink));
    // mix: cycles per 64-instruction iteration against the sum of its parts
    const double mix_iter = mix.med * 64, parts = 16 * sin32.med + 48 * fma32.med;
    auto item = [](const char* name, const Result& R) {
        std::printf("\"%s\": {\"cycles\": %.3f, \"range\": [%.3f, %.3f], \"waves\": %d, \"simds\": %d}",
                    name, R.med, R.lo, R.hi, R.waves, R.simds);
    };
    std::printf("{\"arch\": \"%s\", \"cus\": %d, \"launch_waves_per_simd\": %d, \"iterations\": %d, ",
                prop.gcnArchName, prop.multiProcessorCount, kW, iters);
    item("v_fma_f32_sgpr_operand", fma32);
    std::printf(", ");
    for (int i = 0; i < 5; ++i) {
        char name[48];
        std::snprintf(name, sizeof(name), "v_fma_f32_vgpr_%d_waves", occ[i]);
        item(name, fmav[i]);
        std::printf(", ");
    }
    item("v_sin_f32", sin32);
    std::printf(", ");
    item("v_fma_f64", fma64);
    std::printf(", ");
    item("v_add_f32", add32);
    std::printf(", ");
    item("v_mul_f32", mul32);
    std::printf(", ");
    item("mix_16sin_48fma", mix);
    std::printf(", \"mix_iteration_cycles\": %.1f, \"sum_of_parts_cycles\": %.1f", mix_iter, parts);
    for (int b = 0; b < 6; ++b)
        for (int o = 0; o < 2; ++o) {
            char name[64];
            std::snprintf(name, sizeof(name), "%s_%d_waves", bnames[b], banked[b][o].waves);
            std::printf(", ");
            item(name, banked[b][o]);
        }
    std::printf(", \"mfma_32x32x2_f32\": {\"cycles_each\": %.1f, \"with_128_fma\": %.1f, \"sum_of_parts\": %.1f}",
                mx32.med / 16, bx32.med, mx32.med + vx32.med);
    std::printf(", \"mfma_16x16x1_4b_f32\": {\"cycles_each\": %.1f, \"with_128_fma\": %.1f, \"sum_of_parts\": %.1f}",
                mx16.med / 16, bx16.med, mx16.med + vx16.med);
    std::printf("}\n");
    return 0;
}
