// Work layout of the plain FP32 pair sum, measured.  The reference's psf_kernel
// (propagate.cu:36-84) stages pillars 256 at a time in shared memory and gives each work
// group 256 pixels and the whole pillar list: ceil(n^2 / 256) work-groups, each reading
// every pillar from global memory once.  This library's plain kernels keep that tile and
// also split the pillar list over work-groups (partial sums, reduced afterwards), so a
// launch has enough work-groups for 256 CUs.  The probe runs one LDS-tiled kernel (the
// same float arithmetic as the reference: sqrtf, sincosf, 1/R) in both layouts:
//
//   whole: grid.y = 1, every work-group sums all pillars (the reference's layout);
//   split: grid.y = ceil(npil / chunk) pillar chunks, partial sums per chunk.
//
// It times each with HIP events (median of the timed launches), checks that the split
// partials add up to the whole sums, and prints one JSON line; run it under rocprofv3
// (--kernel-trace, --pmc FETCH_SIZE, ...) for the counters of each layout.
//
// Build: hipcc --offload-arch=gfx950 -O3 (packed FP32 off, as for the library).
// Usage: reread_probe [pillars_per_side pixels_per_side chunk reps]
//        (default 303 161 2048 10: the headline lens and grid; 64 81 512 gives config 1's
//        sizes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int kThreads = 256;

#define CHECK(x)                                                                         \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "reread_probe: %s at line %d\n", hipGetErrorString(e_), \
                         __LINE__);                                                      \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

struct Grid {
    float x0, y0, step, z, k;
    int n;  // n x n pixels
};

// blockIdx.x: 256 pixels; blockIdx.y: pillar chunk [y * chunk, (y + 1) * chunk).
__global__ __launch_bounds__(kThreads) void pair_tiles(const float4* __restrict__ pil,
                                                        int npil, int chunk, Grid g,
                                                        float2* __restrict__ part) {
    __shared__ float4 tile[kThreads];
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    const int npix = g.n * g.n;
    const int pc = pix < npix ? pix : 0;
    const float fx = g.x0 + static_cast<float>(pc % g.n) * g.step;
    const float fy = g.y0 + static_cast<float>(pc / g.n) * g.step;
    const int q0 = blockIdx.y * chunk, q1 = min(npil, q0 + chunk);
    float er = 0.0f, ei = 0.0f;
    for (int base = q0; base < q1; base += kThreads) {
        const int n = min(kThreads, q1 - base);
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < n) tile[threadIdx.x] = pil[base + threadIdx.x];
        __syncthreads();
        for (int q = 0; q < n; ++q) {
            const float4 Q = tile[q];
            const float dx = fx - Q.x, dy = fy - Q.y;
            const float R = sqrtf(dx * dx + dy * dy + g.z * g.z);
            float s, c;
            sincosf(g.k * R, &s, &c);
            const float inv = 1.0f / R;
            er += (Q.z * c - Q.w * s) * inv;
            ei += (Q.z * s + Q.w * c) * inv;
        }
    }
    if (pix < npix) part[static_cast<int64_t>(blockIdx.y) * npix + pix] = make_float2(er, ei);
}

template <class Launch>
double median_ms(Launch launch, int reps) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    launch();  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    std::vector<float> ms;
    for (int r = 0; r < reps; ++r) {
        CHECK(hipEventRecord(a, nullptr));
        launch();
        CHECK(hipEventRecord(b, nullptr));
        CHECK(hipEventSynchronize(b));
        float t = 0.0f;
        CHECK(hipEventElapsedTime(&t, a, b));
        ms.push_back(t);
    }
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}

}  // namespace

int main(int argc, char** argv) {
    const int side = argc > 1 ? std::atoi(argv[1]) : 303;
    const int n = argc > 2 ? std::atoi(argv[2]) : 161;
    const int chunk = argc > 3 ? std::atoi(argv[3]) : 2048;
    const int reps = argc > 4 ? std::atoi(argv[4]) : 10;
    if (side <= 0 || n <= 1 || chunk <= 0 || reps <= 0 || side > 4096 || n > 8192) {
        std::fprintf(stderr, "reread_probe: bad arguments\n");
        return 2;
    }
    // A focusing lens of side x side pillars at 0.35 um (lambda 0.633, f 100), the grid
    // over +-4 lambda / NA at z = f: the headline shape by default.
    const float lam = 0.633f, p = 0.35f, f = 100.0f;
    const float k = 6.28318530717958648f / lam;
    const int npil = side * side;
    std::vector<float4> pil(npil);
    for (int j = 0; j < side; ++j)
        for (int i = 0; i < side; ++i) {
            const float x = (i - 0.5f * (side - 1)) * p, y = (j - 0.5f * (side - 1)) * p;
            const float phi = -k * (std::sqrt(x * x + y * y + f * f) - f);
            pil[j * side + i] = make_float4(x, y, std::cos(phi), std::sin(phi));
        }
    const float half = 0.5f * side * p;
    const float na = half / std::sqrt(half * half + f * f);
    const float W = 4.0f * lam / na;
    const Grid g{-W, -W, 2.0f * W / (n - 1), f, k, n};
    const int npix = n * n;
    const int nchunk = (npil + chunk - 1) / chunk;
    if (nchunk > 65535) {
        std::fprintf(stderr, "reread_probe: too many chunks\n");
        return 2;
    }
    float4* d_pil = nullptr;
    float2* d_part = nullptr;
    CHECK(hipMalloc(&d_pil, sizeof(float4) * npil));
    CHECK(hipMalloc(&d_part, sizeof(float2) * static_cast<size_t>(npix) * nchunk));
    CHECK(hipMemcpy(d_pil, pil.data(), sizeof(float4) * npil, hipMemcpyHostToDevice));
    const int gx = (npix + kThreads - 1) / kThreads;
    const double t_whole = median_ms([&] {
        pair_tiles<<<dim3(gx, 1), kThreads>>>(d_pil, npil, npil, g, d_part);
    }, reps);
    std::vector<float2> whole(npix), split(static_cast<size_t>(npix) * nchunk);
    CHECK(hipMemcpy(whole.data(), d_part, sizeof(float2) * npix, hipMemcpyDeviceToHost));
    const double t_split = median_ms([&] {
        pair_tiles<<<dim3(gx, nchunk), kThreads>>>(d_pil, npil, chunk, g, d_part);
    }, reps);
    CHECK(hipMemcpy(split.data(), d_part, sizeof(float2) * split.size(), hipMemcpyDeviceToHost));
    CHECK(hipFree(d_pil));
    CHECK(hipFree(d_part));
    // the chunk sums (added in FP64) against the whole sums: float rounding only
    double num = 0.0, den = 0.0;
    for (int i = 0; i < npix; ++i) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            sr += split[static_cast<size_t>(c) * npix + i].x;
            si += split[static_cast<size_t>(c) * npix + i].y;
        }
        const double dr = sr - whole[i].x, di = si - whole[i].y;
        num += dr * dr + di * di;
        den += sr * sr + si * si;
    }
    const double pairs = static_cast<double>(npil) * npix;
    std::printf("{\"pillars\": %d, \"pixels\": %d, \"whole\": {\"work_groups\": %d, \"ms\": %.4f, "
                "\"G_pairs_s\": %.1f, \"pillar_MB_read\": %.1f}, \"split\": {\"chunk\": %d, "
                "\"work_groups\": %d, \"ms\": %.4f, \"G_pairs_s\": %.1f, \"pillar_MB_read\": %.1f}, "
                "\"rel_l2_split_vs_whole\": %.2e}\n",
                npil, npix, gx, t_whole, pairs / t_whole / 1e6, gx * npil * 16e-6, chunk,
                gx * nchunk, t_split, pairs / t_split / 1e6, gx * npil * 16e-6,
                std::sqrt(num / den));
    return 0;
}
