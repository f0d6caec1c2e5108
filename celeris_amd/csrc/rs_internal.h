// rs_internal.h — device-side descriptors shared by the HIP kernels (rs_kernels.hip)
// and the host planner (rs_api.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crs {

// Work-group size of every pair kernel: 4 wave64s, one per SIMD of a CU.
constexpr int kThreads = 256;
// FP32C forward: each thread owns a kPX x kPY block of pixels (columns x rows);
// the pillar's dx is shared down a column and its (fy-y)^2+z^2 along a row.
constexpr int kPX = 4;
constexpr int kPY = 4;
// FP32C adjoint: each thread owns kPQ pillars (q0 + tid + k*kThreads).
constexpr int kPQ = 8;

// One plane as the kernels see it (host computes the derived fields).
struct DevPlane {
    double k;       // rad/um
    double kturn;   // k / (2 pi), turns per um
    double x0, sx;  // fx = x0 + i*sx
    double y0, sy;  // fy = y0 + j*sy
    double z0, sz;  // z  = z0 + j*sz
    int32_t nx, ny;
    int32_t tx;     // thread tiles across a row: ceil(nx / px_per_thread)
    int32_t tset;
    int64_t offset; // first output element of the plane
    int64_t part;   // first partial element of slot 0 (forward)
    int32_t nslots; // pillar chunks the forward sum is split into
    int32_t ty;     // thread tiles down a column: ceil(ny / py_per_thread)
};

// Forward work-group: thread tiles [tile0, tile0+kThreads) of `plane`, summed over
// pillars [q0, q1); results go to partial[part + pixel].
struct FwdBlock {
    int32_t plane, tile0, q0, q1;
    int64_t part;
};

// A run of <= kThreads consecutive pixels of one plane (staging and reductions).
struct PixBlock {
    int32_t plane, pix0;
};

// One pixel row in the adjoint's per-set row list.
struct AdjRow {
    int32_t plane, j;
};

// Adjoint work-group: pillars q0 + tid + k*kThreads (k < pillars per thread) of
// transmission set `tset`, summed over rows [r0, r1) of the row list; results go
// to partial[part + q].
struct AdjBlock {
    int32_t tset, q0, r0, r1;
    int64_t part;
};

}  // namespace crs
