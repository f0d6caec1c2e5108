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

struct Result {
    double med, lo, hi;  // over the fullest SIMDs
    int waves, simds;    // waves on those SIMDs, and how many SIMDs held that many
};

template <int MODE>
Result run(int nwg, int iters, int ninstr, long long* d_rec, float* d_sink) {
    const int nwaves = nwg * kThreadsPerWg / 64;
    probe<MODE><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, 8, 0.999f);  // warm-up
    CHECK(hipGetLastError());
    probe<MODE><<<nwg, kThreadsPerWg>>>(d_rec, d_sink, iters, 0.999f);
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
    std::printf(", \"mix_iteration_cycles\": %.1f, \"sum_of_parts_cycles\": %.1f}\n", mix_iter, parts);
    return 0;
}
