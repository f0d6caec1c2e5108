// rs_internal.h — device-side descriptors shared by the HIP kernels (rs_kernels.hip)
// and the host planner (rs_api.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crs {

// Work-group size of every pair kernel: 4 wave64s, one per SIMD of a CU.
constexpr int kThreads = 256;
// Waves (wave64) per work-group.
constexpr int kWaves = kThreads / 64;
// FP32C: pixels are handled in blocks of 4 x 4 (4 columns x 4 rows) that share one
// reference distance, or 1 x 1 where a plane's pixel pitch is too coarse for the
// block expansion.  The forward gives each lane one block ("thread tile") and a
// work-group kFwdTiles tiles, its 4 waves splitting the pillar chunk; the adjoint
// walks a plane in strips of kQuad rows ("quads") and each quad in 4-column groups.
constexpr int kBlk = 4;
constexpr int kQuad = 4;
constexpr int kFwdTiles = 64;
// Outputs per work-group of the partial-sum reductions (each wave adds a quarter of
// the slots for kReduceCols outputs).
constexpr int kReduceCols = 64;
// Longest pillar chunk one FP32C forward work-group sums (each wave a quarter of it,
// in float, folded every 64 pillars): bounds the float accumulation error.
constexpr int kMaxChunk = 16384;
// Shortest pillar chunk of a forward split (small problems only; 16 pillars per wave).
constexpr int kMinChunk = 64;
// Plan order code of the 1 x 1 (unexpanded) FP32C kernels, used when every plane of a
// call needs each pair at its own reference distance (fp32c_order); codes 0..3 select
// the 4 x 4 kernels, which choose their series order per work-group; planes flagged
// exact in such a call are evaluated pair by pair (forward: 1 x 1 tiles; adjoint: a
// second launch of the 1 x 1 kernel).
constexpr int kExactOrder = 4;
// FP32C adjoint: pillars per thread (q0 + tid + k*kThreads, k < kPQ).  One: four per
// thread (sharing each block's broadcast pixel records) measured 5-6% slower at 91,809
// pillars x 161^2 and 321^2 and at 1,002,001 x 1024^2.
constexpr int kPQ = 1;
// FP32C adjoint: columns per LDS tile of a quad (kAdjCols x 4 pixel records).
constexpr int kAdjCols = kThreads / kQuad;

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

// FP32C staged pillar buffer (forward): positions (double2, npil) and {|t|, arg t / 2 pi}
// (float2, nsets x npil).
struct StagedLayout {
    size_t pos, ap, bytes;
};
inline StagedLayout fp32c_staged_layout(int64_t npil, int64_t nsets) {
    StagedLayout L;
    L.pos = 0;
    L.ap = static_cast<size_t>(npil) * 16;
    L.bytes = L.ap + static_cast<size_t>(npil * nsets) * 8;
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
    int32_t tx;     // forward thread tiles across a row (FP32C: floor(nx / 4) 4 x 4 tiles)
    int32_t ty;     // forward thread tiles down a column (FP32C: floor(ny / 4))
    int32_t tset;
    int32_t nslots; // pillar chunks the forward sum is split into
    int64_t offset; // first output element of the plane
    int64_t part;   // first forward partial element of slot 0
    int64_t sbase;  // FP32C adjoint: first staged pixel record of the plane
    int32_t nxp;    // FP32C adjoint: nx rounded up to a multiple of 4
    int32_t nquads; // ceil(ny / 4)
};

// Forward work-group: thread tiles [tile0, tile0 + kFwdTiles) of class `shape`
// (FP32C) or pixels [tile0, tile0 + kThreads) (FP32, FP64) of `plane`, summed over
// pillars [q0, q1); results go to partial[part + pixel].
struct FwdBlock {
    int32_t plane, tile0, q0, q1;
    int64_t part;
    int32_t shape;  // FP32C tile class: 0 = 4 x 4, 1 = 1 x 4, 2 = 4 x 1, 3 = 1 x 1
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

// Adjoint work-group: pillars q0 + tid + k*kThreads (k < pillars per thread) of
// transmission set `tset`, summed over quads [u0, u1) of the quad list; results go
// to partial[part + q].
struct AdjBlock {
    int32_t tset, q0, u0, u1;
    int64_t part;
};

}  // namespace crs
