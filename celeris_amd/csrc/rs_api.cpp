// rs_api.cpp — host side of the C ABI (include/celeris_rs.h): argument checks, launch
// planning, scratch management and the reference-compatible drop-in entry points.
//
// Reference behaviour kept by the drop-in layer (src/celeris/cuda/propagate.cu):
//   - false when no device is visible (:137-138) or npil <= 0 || n <= 0 (:139);
//     propagate_zx also rejects nx <= 1 || nz <= 1 (:201)
//   - false on any allocation (:164-171) or launch/sync failure (:184), with `out`
//     left untouched in every false case
//   - grid step 2W/(n-1) (:143); n == 1 gives step = inf and NaN coordinates, which
//     the reference propagates into NaN intensities; so does this library
//   - synchronous; callable from several host threads at once (the reference's
//     callers do so from std::async trials and GUI threads)
#include "../../include/celeris_rs.h"
#include "rs_host.h"
#include "rs_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

namespace crs {
hipError_t launch_stage_pillars(int mode, const double* px, const double* py, const double* t,
                                int npil, int nsets, void* out, hipStream_t s);
hipError_t launch_forward(int mode, int order, const void* pil, int npil, int nsets,
                          const DevPlane* planes, const FwdBlock* blocks, int nblocks,
                          void* partial, hipStream_t s);
hipError_t launch_fwd_reduce(int mode, const DevPlane* planes, const PixBlock* pblocks,
                             int nblocks, const void* partial, double2* field,
                             double* intensity, hipStream_t s);
hipError_t launch_stage_pixels(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                               const double2* u, float4* pix, hipStream_t s);
hipError_t launch_adjoint(int mode, int order, const void* pil, int npil,
                          const DevPlane* planes, const AdjQuad* quads, const AdjBlock* blocks,
                          int nblocks, const void* pix_or_u, void* partial, hipStream_t s);
hipError_t launch_adj_reduce(int mode, const void* partial, int nslots, int64_t n, double2* g,
                             hipStream_t s);
hipError_t launch_mtf(const double2* P, const double* P0, int n, int ns, double p, double f,
                      double lam, double2* acc, double* acc0, float* freq, float* mtf,
                      float* ideal, hipStream_t s);
}  // namespace crs

using namespace crs;

namespace {
thread_local std::string g_last_error;
}  // namespace

void crs::set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void crs::clear_error() { g_last_error.clear(); }

namespace {

// Fault injection for tests of the failure paths: CELERIS_RS_FAIL=alloc makes every
// scratch allocation fail, CELERIS_RS_FAIL=launch makes every launch fail; in both
// cases before any device work is enqueued.
bool inject(const char* what) {
    const char* v = std::getenv("CELERIS_RS_FAIL");
    return v && std::strcmp(v, what) == 0;
}

constexpr double kInv2Pi = 0.15915494309189533577;

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct DevBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
};

bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st != hipStreamCaptureStatusNone;
}

}  // namespace

namespace {

// Host-built descriptor tables travel in one pinned buffer and one async copy.
struct TableBuilder {
    std::vector<unsigned char> bytes;
    size_t add(const void* p, size_t n) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + n);
        if (n) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};

struct FwdPlan {
    int mode = 0;
    int order = 0;
    int npil = 0, nsets = 0;
    int64_t part_elems = 0;
    int nblocks = 0, npblocks = 0;
    TableBuilder tb;
    size_t o_planes = 0, o_blocks = 0, o_pblocks = 0;
};

struct AdjPlan {
    int mode = 0;
    int order = 1;
    int npil = 0, nsets = 0;
    int64_t nsplit = 1, n_out = 0, staged_records = 0;  // nsplit: partial slots, all launches
    int nblocks = 0, npblocks = 0;
    // FP32C calls mixing planes flagged exact with blocked ones: a second launch, of the
    // 1 x 1 kernel over the flagged planes' quads, into partial slots after the first's.
    int nblocks_exact = 0;
    TableBuilder tb;
    size_t o_planes = 0, o_quads = 0, o_blocks = 0, o_blocks_exact = 0, o_pblocks = 0;
};

// A context remembers the last plan of each direction; a call with the same shape
// (mode, sizes, planes, tuning) reuses it and its device tables without host work.
struct PlanSlot {
    std::vector<unsigned char> key;
    DevBuffer tables;
    bool valid = false;
};

constexpr int kPinnedRing = 4;

}  // namespace

struct crs_context {
    int device = 0;
    DevBuffer pil, pix, partial;
    PlanSlot slot[2];  // [0] forward, [1] adjoint
    FwdPlan fwd_plan;
    AdjPlan adj_plan;
    // Pinned staging for table uploads: a ring, so a call waits only for the copy
    // issued kPinnedRing calls earlier.
    void* host_tables[kPinnedRing] = {nullptr, nullptr, nullptr, nullptr};
    size_t host_cap[kPinnedRing] = {0, 0, 0, 0};
    hipEvent_t tables_ev[kPinnedRing] = {nullptr, nullptr, nullptr, nullptr};
    bool tables_pending[kPinnedRing] = {false, false, false, false};
    int ring_next = 0;
    int fwd_target[3] = {0, 0, 0};
    int adj_target[3] = {0, 0, 0};
    bool timing = false;
    hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [adjoint][start/stop]
    bool ev_valid[2] = {false, false};
    const char* last_kernel[2] = {nullptr, nullptr};  // [adjoint]: last pair kernel launched
    int last_groups[2] = {0, 0};
};

namespace {

// Built-in launch targets: work-groups per forward/adjoint launch.  256 CUs x a few
// resident work-groups, several waves deep so the tail of uneven splits is short.  On
// the 91809-pillar, 161 x 161 case both FP32C pair kernels measure within 2% of their
// best between 3072 and 6144.
int default_target(int mode, bool adjoint) {
    (void)mode;
    (void)adjoint;
    return 4096;
}

crs_status ensure(DevBuffer& b, size_t bytes, hipStream_t stream, const char* what) {
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return CRS_OK;
    if (stream_capturing(stream)) {
        set_error("%s: %zu bytes of scratch needed while the stream is being captured; "
                  "make one uncaptured call of the same shape first", what, bytes);
        return CRS_ERR_CAPTURE;
    }
    if (inject("alloc")) {
        set_error("%s: injected allocation failure", what);
        return CRS_ERR_ALLOC;
    }
    if (b.ptr) {
        // The old buffer may still be read by queued work on this stream.
        hipStreamSynchronize(stream);
        hipFree(b.ptr);
        b.ptr = nullptr;
        b.cap = 0;
    }
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        hipGetLastError();
        b.ptr = nullptr;
        set_error("%s: hipMalloc of %zu bytes failed", what, want);
        return CRS_ERR_ALLOC;
    }
    b.cap = want;
    return CRS_OK;
}


crs_status upload_tables(crs_context* ctx, DevBuffer& dst, const TableBuilder& tb,
                         hipStream_t stream) {
    const size_t n = tb.bytes.size();
    crs_status st = ensure(dst, n, stream, "descriptor tables");
    if (st != CRS_OK) return st;
    const int r = ctx->ring_next;
    ctx->ring_next = (r + 1) % kPinnedRing;
    // The copy issued kPinnedRing calls ago may still be reading this pinned buffer.
    if (ctx->tables_pending[r]) { hipEventSynchronize(ctx->tables_ev[r]); ctx->tables_pending[r] = false; }
    if (ctx->host_cap[r] < n) {
        if (stream_capturing(stream)) {
            set_error("descriptor tables: pinned staging must grow while capturing");
            return CRS_ERR_CAPTURE;
        }
        if (ctx->host_tables[r]) hipHostFree(ctx->host_tables[r]);
        ctx->host_tables[r] = nullptr;
        ctx->host_cap[r] = 0;
        const size_t want = n + n / 4 + 4096;
        if (hipHostMalloc(&ctx->host_tables[r], want, hipHostMallocDefault) != hipSuccess) {
            hipGetLastError();
            ctx->host_tables[r] = nullptr;
            set_error("descriptor tables: hipHostMalloc of %zu bytes failed", want);
            return CRS_ERR_ALLOC;
        }
        ctx->host_cap[r] = want;
    }
    std::memcpy(ctx->host_tables[r], tb.bytes.data(), n);
    if (hipMemcpyAsync(dst.ptr, ctx->host_tables[r], n, hipMemcpyHostToDevice, stream) !=
        hipSuccess) {
        hipGetLastError();
        set_error("descriptor tables: hipMemcpyAsync failed");
        return CRS_ERR_LAUNCH;
    }
    if (!stream_capturing(stream)) {
        hipEventRecord(ctx->tables_ev[r], stream);
        ctx->tables_pending[r] = true;
    }
    return CRS_OK;
}

// Key of a call's plan: everything the tables depend on (not the data pointers).
std::vector<unsigned char> plan_key(int mode, int npil, int nsets, const crs_plane* planes,
                                    int nplanes, int target) {
    std::vector<unsigned char> key(5 * sizeof(int) + static_cast<size_t>(nplanes) * sizeof(crs_plane));
    const int head[5] = {mode, npil, nsets, nplanes, target};
    std::memcpy(key.data(), head, sizeof head);
    if (nplanes) std::memcpy(key.data() + sizeof head, planes, static_cast<size_t>(nplanes) * sizeof(crs_plane));
    return key;
}


// Arguments every batched entry point shares, checked before any device work.
crs_status check_args(int mode, int npil, int nsets, const crs_plane* planes, int nplanes) {
    if (mode < 0 || mode > 2) { set_error("unknown mode %d", mode); return CRS_ERR_INVALID; }
    if (npil <= 0) { set_error("npil must be positive (got %d)", npil); return CRS_ERR_INVALID; }
    if (nsets <= 0) { set_error("nsets must be positive (got %d)", nsets); return CRS_ERR_INVALID; }
    if (nplanes < 0 || (nplanes > 0 && !planes)) {
        set_error("bad plane list");
        return CRS_ERR_INVALID;
    }
    if (static_cast<int64_t>(npil) * nsets > (int64_t(1) << 31) - 1) {
        set_error("nsets*npil exceeds 2^31-1");
        return CRS_ERR_INVALID;
    }
    for (int p = 0; p < nplanes; ++p) {
        const crs_plane& P = planes[p];
        if (P.nx <= 0 || P.ny <= 0 || static_cast<int64_t>(P.nx) * P.ny > (int64_t(1) << 30)) {
            set_error("plane %d: bad pixel counts %d x %d", p, P.nx, P.ny);
            return CRS_ERR_INVALID;
        }
        if (P.tset < 0 || P.tset >= nsets) {
            set_error("plane %d: tset %d outside [0, %d)", p, P.tset, nsets);
            return CRS_ERR_INVALID;
        }
        if (P.offset < 0 || P.reserved != 0) {
            set_error("plane %d: negative offset or nonzero reserved field", p);
            return CRS_ERR_INVALID;
        }
    }
    return CRS_OK;
}

crs_status check_common(crs_context* ctx, int mode, int npil, int nsets, const crs_plane* planes,
                        int nplanes) {
    if (!ctx) { set_error("null context"); return CRS_ERR_INVALID; }
    const crs_status st = check_args(mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        hipGetLastError();
        set_error("hipSetDevice(%d) failed", ctx->device);
        return CRS_ERR_NO_DEVICE;
    }
    return CRS_OK;
}

DevPlane make_dev_plane(const crs_plane& P, int px_per_thread, int py_per_thread) {
    DevPlane d{};
    d.k = P.k;
    d.kturn = P.k * kInv2Pi;
    d.x0 = P.x0; d.sx = P.sx;
    d.y0 = P.y0; d.sy = P.sy;
    d.z0 = P.z0; d.sz = P.sz;
    d.zref = d.kturn * P.z0;
    d.nx = P.nx; d.ny = P.ny;
    d.tx = static_cast<int32_t>(ceil_div(P.nx, px_per_thread));
    d.ty = static_cast<int32_t>(ceil_div(P.ny, py_per_thread));
    d.tset = P.tset;
    d.nslots = 1;
    d.offset = P.offset;
    d.part = 0;
    d.sbase = 0;
    d.nxp = static_cast<int32_t>(ceil_div(P.nx, kQuad) * kQuad);
    d.nquads = static_cast<int32_t>(ceil_div(P.ny, kQuad));
    return d;
}

// Series order 4 x 4 blocks need on one plane, with no knowledge of where the pillars
// are (the kernels refine it per work-group from the pillars' extent).  A block spans at
// most e turns around its centre; at distance >= z' the expansion parameter is
// r = e / z' (y <= r/2).  Dropped terms, at worst (a pillar seen at grazing angle): phase
// order 1 leaves 2 x1 y^2 <= e r^2 / 2 turns, order 2 leaves ~5 x1 y^3 <= 0.625 e r^3;
// amplitude order 1 leaves 6 y^2 <= 1.5 r^2, order 2 leaves ~20 y^3 <= 2.5 r^3, against
// budgets of 1.6e-7 turns (1e-6 rad) / 4e-6 of amplitude.  Returns kExactOrder when even
// order 2 would exceed them (the plane's pairs are then evaluated one by one at their
// own reference distances), else 2 * (phase order - 1) + (amplitude order - 1).  Planes
// with non-finite geometry (n == 1 windows) give NaN either way and return 0.
int fp32c_plane_order(const DevPlane& d) {
    const double ex = 1.5 * std::fabs(d.sx) * d.kturn;
    const double ey = 1.5 * std::fabs(d.sy) * d.kturn;
    const double ez = 1.5 * std::fabs(d.sz) * d.kturn;
    const double e = std::sqrt(ex * ex + ey * ey) + ez;
    const double zmin = d.kturn * plane_min_abs_z(d.z0, d.sz, d.ny);
    if (!std::isfinite(e)) return 0;
    if (!(zmin > 0.0)) return kExactOrder;  // a z = 0 row: no distance bound
    int po = 1, ao = 1;
    const double r = e / zmin;
    if (0.5 * e * r * r > 1.6e-7) {
        if (0.625 * e * r * r * r > 1.6e-7) return kExactOrder;
        po = 2;
    }
    if (1.5 * r * r > 4e-6) {
        if (2.5 * r * r * r > 4e-6) return kExactOrder;
        ao = 2;
    }
    return 2 * (po - 1) + (ao - 1);
}

// Flags (exact[p]) each plane that needs pair-by-pair evaluation and returns the call's
// order code: kExactOrder when every plane does (the 1 x 1 kernels), else the highest
// order of the others (informational; any code below kExactOrder selects the 4 x 4
// kernels, and the flagged planes of such a call are taken pair by pair).
int fp32c_order(const std::vector<DevPlane>& dp, std::vector<char>& exact) {
    int order = 0;
    bool all_exact = !dp.empty();
    exact.assign(dp.size(), 0);
    for (size_t p = 0; p < dp.size(); ++p) {
        const int o = fp32c_plane_order(dp[p]);
        exact[p] = o == kExactOrder;
        if (exact[p]) continue;
        all_exact = false;
        order = std::max(order, o);
    }
    return all_exact ? kExactOrder : order;
}

// Bytes of staged pillar data: FP32C keeps FP64 positions once plus {|t|, arg t} per
// set; FP32 and FP64 keep one {x, y, Re t, Im t} record per pillar and set.
size_t staged_pillar_bytes(int mode, int64_t npil, int64_t nsets) {
    if (mode == CRS_MODE_FP32C) return fp32c_staged_layout(npil, nsets).bytes;
    return npil * nsets * (mode == CRS_MODE_FP64 ? sizeof(double4) : sizeof(float4));
}

crs_status launch_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return CRS_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return CRS_ERR_LAUNCH;
}

}  // namespace

// ------------------------------------------------------------------------------------
// Batched layer.

extern "C" crs_status crs_context_create(int device, crs_context** out) {
    if (!out) { set_error("null output pointer"); return CRS_ERR_INVALID; }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        hipGetLastError();
        set_error("no HIP device visible");
        return CRS_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        set_error("device %d outside [0, %d)", device, count);
        return CRS_ERR_INVALID;
    }
    if (hipSetDevice(device) != hipSuccess) {
        hipGetLastError();
        set_error("hipSetDevice(%d) failed", device);
        return CRS_ERR_NO_DEVICE;
    }
    crs_context* ctx = new crs_context();
    ctx->device = device;
    for (hipEvent_t& e : ctx->tables_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            crs_context_destroy(ctx);
            set_error("hipEventCreate failed");
            return CRS_ERR_LAUNCH;
        }
    }
    *out = ctx;
    return CRS_OK;
}

extern "C" void crs_context_destroy(crs_context* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    for (DevBuffer* b : {&ctx->pil, &ctx->pix, &ctx->partial, &ctx->slot[0].tables,
                         &ctx->slot[1].tables})
        if (b->ptr) hipFree(b->ptr);
    for (auto& pair : ctx->ev)
        for (hipEvent_t e : pair)
            if (e) hipEventDestroy(e);
    for (int r = 0; r < kPinnedRing; ++r) {
        if (ctx->host_tables[r]) hipHostFree(ctx->host_tables[r]);
        if (ctx->tables_ev[r]) hipEventDestroy(ctx->tables_ev[r]);
    }
    delete ctx;
}

extern "C" const char* crs_last_error(void) { return g_last_error.c_str(); }

extern "C" crs_status crs_set_tuning(crs_context* ctx, int mode, int fwd_target_blocks,
                                     int adj_target_blocks) {
    if (!ctx || mode < 0 || mode > 2 || fwd_target_blocks < 0 || adj_target_blocks < 0) {
        set_error("bad tuning arguments");
        return CRS_ERR_INVALID;
    }
    ctx->fwd_target[mode] = fwd_target_blocks;
    ctx->adj_target[mode] = adj_target_blocks;
    return CRS_OK;
}

extern "C" crs_status crs_set_timing(crs_context* ctx, int enable) {
    if (!ctx) { set_error("null context"); return CRS_ERR_INVALID; }
    if (enable && !ctx->ev[0][0]) {
        hipSetDevice(ctx->device);
        for (auto& pair : ctx->ev)
            for (hipEvent_t& e : pair)
                if (hipEventCreate(&e) != hipSuccess) {
                    hipGetLastError();
                    set_error("hipEventCreate failed");
                    return CRS_ERR_LAUNCH;
                }
    }
    ctx->timing = enable != 0;
    ctx->ev_valid[0] = ctx->ev_valid[1] = false;
    return CRS_OK;
}

extern "C" crs_status crs_last_kernel_ms(crs_context* ctx, int adjoint, float* ms) {
    if (!ctx || !ms || adjoint < 0 || adjoint > 1) { set_error("bad arguments"); return CRS_ERR_INVALID; }
    if (!ctx->ev_valid[adjoint]) { set_error("no timed launch recorded"); return CRS_ERR_INVALID; }
    if (hipEventSynchronize(ctx->ev[adjoint][1]) != hipSuccess ||
        hipEventElapsedTime(ms, ctx->ev[adjoint][0], ctx->ev[adjoint][1]) != hipSuccess) {
        hipGetLastError();
        set_error("hipEventElapsedTime failed");
        return CRS_ERR_LAUNCH;
    }
    return CRS_OK;
}

namespace {
// Profiler name of the pair kernel a plan launches (launch_forward / launch_adjoint).
const char* pair_kernel(int mode, int order, bool adjoint) {
    if (mode == CRS_MODE_FP32C) {
        if (order == kExactOrder) return adjoint ? "crs::adj_fp32c<1, 1>" : "crs::fwd_fp32c<1>";
        return adjoint ? "crs::adj_fp32c<4, 4>" : "crs::fwd_fp32c<4>";
    }
    if (mode == CRS_MODE_FP64) return adjoint ? "crs::adj_fp64" : "crs::fwd_fp64";
    return adjoint ? "crs::adj_fp32" : "crs::fwd_fp32";
}
}  // namespace

extern "C" crs_status crs_last_launch(crs_context* ctx, int adjoint, const char** kernel,
                                      int* workgroups) {
    if (!ctx || !kernel || !workgroups || adjoint < 0 || adjoint > 1) {
        set_error("bad arguments");
        return CRS_ERR_INVALID;
    }
    if (!ctx->last_kernel[adjoint]) { set_error("no launch recorded"); return CRS_ERR_INVALID; }
    *kernel = ctx->last_kernel[adjoint];
    *workgroups = ctx->last_groups[adjoint];
    return CRS_OK;
}

extern "C" const char* crs_kernel_name(int mode, int adjoint) {
    switch (mode) {
        case CRS_MODE_FP32C: return adjoint ? "adj_fp32c" : "fwd_fp32c";
        case CRS_MODE_FP64: return adjoint ? "adj_fp64" : "fwd_fp64";
        case CRS_MODE_FP32: return adjoint ? "adj_fp32" : "fwd_fp32";
        default: return "";
    }
}

// ------------------------------------------------------------------------------------
// Planning: host-built descriptor tables for one forward / adjoint call.

namespace {

struct Scratch {  // where one call's device data lives (a context's or a graph's)
    DevBuffer* pil;
    DevBuffer* pix;
    DevBuffer* partial;
    unsigned char* tables;  // device copy of the plan's table bytes
};


crs_status plan_forward(const crs_context* ctx, int mode, int npil, int nsets,
                        const crs_plane* planes, int nplanes, FwdPlan& plan) {
    // Output ranges [offset, offset + nx*ny) must not overlap: each element has one writer.
    {
        std::vector<std::pair<int64_t, int64_t>> r(nplanes);
        for (int p = 0; p < nplanes; ++p)
            r[p] = {planes[p].offset, planes[p].offset + static_cast<int64_t>(planes[p].nx) * planes[p].ny};
        std::sort(r.begin(), r.end());
        for (int p = 1; p < nplanes; ++p)
            if (r[p].first < r[p - 1].second) {
                set_error("planes' output ranges overlap at element %lld", (long long)r[p].first);
                return CRS_ERR_INVALID;
            }
    }
    plan.mode = mode;
    plan.npil = npil;
    plan.nsets = nsets;
    std::vector<DevPlane> dp(nplanes);
    for (int p = 0; p < nplanes; ++p) dp[p] = make_dev_plane(planes[p], 1, 1);
    int blk = 1;
    std::vector<char> exact(nplanes, 0);
    if (mode == CRS_MODE_FP32C) {
        plan.order = fp32c_order(dp, exact);
        blk = plan.order == kExactOrder ? 1 : kBlk;
    }
    // Split over pillar chunks: enough work-groups to fill the chip, chunks of whole
    // LDS tiles (FP32C: no longer than kMaxChunk).  FP32C work-groups take kFwdTiles
    // thread tiles of a plane, FP32 / FP64 ones kThreads pixels.
    // FP32C tiles come in four classes (kernel: tile_count / tile_origin): 4 x 4 tiles
    // over the largest multiple-of-4 part of the plane, then 1 x 4, 4 x 1 and 1 x 1
    // tiles for the leftover columns, rows and corner, so no lane computes a pixel
    // outside the plane.  With 1 x 1 blocks (tx = ty = 0: the 1 x 1 kernels, or a plane
    // flagged exact in a call of the 4 x 4 kernels) the last class is the plane.
    const int per_wg = mode == CRS_MODE_FP32C ? kFwdTiles : kThreads;
    const int nclass = mode == CRS_MODE_FP32C ? 4 : 1;
    std::vector<std::array<int64_t, 4>> counts(nplanes);
    int64_t base_blocks = 0;
    for (int p = 0; p < nplanes; ++p) {
        std::array<int64_t, 4>& c = counts[p];
        if (mode == CRS_MODE_FP32C) {
            const bool blocked = blk > 1 && !exact[p];
            dp[p].tx = blocked ? dp[p].nx / blk : 0;
            dp[p].ty = blocked ? dp[p].ny / blk : 0;
            const int64_t tx = dp[p].tx, ty = dp[p].ty;
            const int64_t rx = dp[p].nx - 4 * tx, ry = dp[p].ny - 4 * ty;
            c = {tx * ty, rx * ty, tx * ry, rx * ry};
        } else {
            c = {static_cast<int64_t>(dp[p].nx) * dp[p].ny, 0, 0, 0};
        }
        for (int k = 0; k < nclass; ++k) base_blocks += ceil_div(c[k], per_wg);
    }
    const int target = ctx->fwd_target[mode] > 0 ? ctx->fwd_target[mode] : default_target(mode, false);
    int64_t nsplit = std::max<int64_t>(1, ceil_div(target, base_blocks));
    if (mode == CRS_MODE_FP32C) nsplit = std::max<int64_t>(nsplit, ceil_div(npil, kMaxChunk));
    // Chunks are whole LDS tiles of kThreads pillars, or, when a problem is too small to
    // fill the chip that way, multiples of kMinChunk (each wave then takes kMinChunk / 4).
    nsplit = std::min<int64_t>(nsplit, std::max<int64_t>(1, npil / kMinChunk));
    int64_t chunk = ceil_div(npil, nsplit);
    chunk = chunk >= kThreads ? ceil_div(chunk, kThreads) * kThreads : ceil_div(chunk, kMinChunk) * kMinChunk;
    nsplit = ceil_div(npil, chunk);
    if (base_blocks * nsplit > (int64_t(1) << 31) - 1) {
        set_error("launch too large (%lld work-groups)", (long long)(base_blocks * nsplit));
        return CRS_ERR_INVALID;
    }
    int64_t part = 0;
    for (int p = 0; p < nplanes; ++p) {
        dp[p].part = part;
        dp[p].nslots = static_cast<int32_t>(nsplit);
        part += nsplit * static_cast<int64_t>(dp[p].nx) * dp[p].ny;
    }
    std::vector<FwdBlock> blocks;
    blocks.reserve(static_cast<size_t>(base_blocks * nsplit));
    for (int64_t s = 0; s < nsplit; ++s) {
        const int32_t q0 = static_cast<int32_t>(s * chunk);
        const int32_t q1 = static_cast<int32_t>(std::min<int64_t>(npil, (s + 1) * chunk));
        for (int p = 0; p < nplanes; ++p) {
            const int64_t npix = static_cast<int64_t>(dp[p].nx) * dp[p].ny;
            for (int k = 0; k < nclass; ++k)
                for (int64_t b = 0; b < counts[p][k]; b += per_wg)
                    blocks.push_back({p, static_cast<int32_t>(b), q0, q1, dp[p].part + s * npix,
                                      k, 0});
        }
    }
    // Heaviest work-groups first.  The dispatcher hands work-groups to the 8 XCDs round
    // robin in blockIdx order, so a run of equal heavy groups spreads evenly over the
    // XCDs whatever the groups per split, and the light edge and short-chunk groups
    // fill the tail instead of landing on the same XCDs split after split.  Each block
    // names its own output slot, so the order does not change any result.
    {
        const int64_t pix_per_tile[4] = {int64_t(blk) * blk, blk, blk, 1};
        auto cost = [&](const FwdBlock& B) {
            const int64_t tiles = std::min<int64_t>(per_wg, counts[B.plane][B.shape] - B.tile0);
            const int64_t pix = mode == CRS_MODE_FP32C ? pix_per_tile[B.shape] : 1;
            return static_cast<int64_t>(B.q1 - B.q0) * tiles * (pix + 2);
        };
        std::stable_sort(blocks.begin(), blocks.end(), [&](const FwdBlock& a, const FwdBlock& b) {
            return cost(a) > cost(b);
        });
    }
    std::vector<PixBlock> pblocks;
    for (int p = 0; p < nplanes; ++p) {
        const int64_t npix = static_cast<int64_t>(dp[p].nx) * dp[p].ny;
        for (int64_t b = 0; b < npix; b += kReduceCols) pblocks.push_back({p, static_cast<int32_t>(b)});
    }
    plan.part_elems = part;
    plan.nblocks = static_cast<int>(blocks.size());
    plan.npblocks = static_cast<int>(pblocks.size());
    plan.o_planes = plan.tb.add(dp.data(), dp.size() * sizeof(DevPlane));
    plan.o_blocks = plan.tb.add(blocks.data(), blocks.size() * sizeof(FwdBlock));
    plan.o_pblocks = plan.tb.add(pblocks.data(), pblocks.size() * sizeof(PixBlock));
    return CRS_OK;
}

// Bytes of one split-sum partial: FP32C keeps each chunk's (or quad run's) sum as a
// float pair, rounded once from its FP64 wave combine; FP64 and FP32 keep doubles.
size_t partial_bytes(int mode) {
    return mode == CRS_MODE_FP32C ? sizeof(float2) : sizeof(double2);
}

crs_status size_forward(const FwdPlan& plan, const Scratch& sc, hipStream_t stream) {
    crs_status st = ensure(*sc.pil, staged_pillar_bytes(plan.mode, plan.npil, plan.nsets), stream,
                           "staged pillars");
    if (st != CRS_OK) return st;
    return ensure(*sc.partial, static_cast<size_t>(plan.part_elems) * partial_bytes(plan.mode),
                  stream, "forward partials");
}

crs_status enqueue_forward(const FwdPlan& plan, const Scratch& sc, const double* px,
                           const double* py, const double* t, double* field, double* intensity,
                           hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
    crs_status st;
    const int mode = plan.mode;
    const DevPlane* d_planes = reinterpret_cast<const DevPlane*>(sc.tables + plan.o_planes);
    if ((st = launch_status(launch_stage_pillars(mode, px, py, t, plan.npil, plan.nsets,
                                                 sc.pil->ptr, stream),
                            "stage pillars")) != CRS_OK)
        return st;
    if (ev0) hipEventRecord(ev0, stream);
    if ((st = launch_status(launch_forward(mode, plan.order, sc.pil->ptr, plan.npil, plan.nsets, d_planes,
                                           reinterpret_cast<const FwdBlock*>(sc.tables + plan.o_blocks),
                                           plan.nblocks, sc.partial->ptr, stream),
                            "forward")) != CRS_OK)
        return st;
    if (ev1) hipEventRecord(ev1, stream);
    return launch_status(launch_fwd_reduce(mode, d_planes,
                                           reinterpret_cast<const PixBlock*>(sc.tables + plan.o_pblocks),
                                           plan.npblocks, sc.partial->ptr,
                                           reinterpret_cast<double2*>(field), intensity, stream),
                         "forward reduce");
}


crs_status plan_adjoint(const crs_context* ctx, int mode, int npil, int nsets,
                        const crs_plane* planes, int nplanes, AdjPlan& plan) {
    plan.mode = mode;
    plan.npil = npil;
    plan.nsets = nsets;
    std::vector<DevPlane> dp(nplanes);
    int64_t srec = 0;
    for (int p = 0; p < nplanes; ++p) {
        dp[p] = make_dev_plane(planes[p], 1, 1);
        dp[p].sbase = srec;
        srec += static_cast<int64_t>(dp[p].nquads) * dp[p].nxp * kQuad;
    }
    if (srec > (int64_t(1) << 31)) { set_error("too many pixels"); return CRS_ERR_INVALID; }
    std::vector<char> exact(nplanes, 0);
    if (mode == CRS_MODE_FP32C) plan.order = fp32c_order(dp, exact);
    // A mixed FP32C call (some planes flagged exact, the order below kExactOrder) gets
    // two launches: the 4 x 4 kernel over the other planes, then the 1 x 1 kernel over
    // the flagged ones, so the 4 x 4 kernel carries no pair-by-pair path.
    bool mixed = false;
    if (mode == CRS_MODE_FP32C && plan.order != kExactOrder)
        for (char e : exact) mixed = mixed || e;

    const int target = ctx->adj_target[mode] > 0 ? ctx->adj_target[mode] : default_target(mode, true);
    const int PQ = mode == CRS_MODE_FP32C ? kPQ : 1;
    const int64_t groups = ceil_div(npil, static_cast<int64_t>(kThreads) * PQ);
    const int64_t base_blocks = groups * nsets;
    const int64_t n_out = static_cast<int64_t>(nsets) * npil;
    std::vector<AdjQuad> quads;

    // One launch's work-groups over the planes `take` selects: per-set quad lists (planes
    // in order, quads in order, appended to `quads`), each set's quads cut into nsplit
    // runs of about equal work (in pixel records: FP32C computes the padded 4 x nxp
    // records of a quad), results to partial slots slot0, slot0 + 1, ...  Returns nsplit.
    auto add_launch = [&](auto take, int64_t slot0, std::vector<AdjBlock>& blocks) -> int64_t {
        std::vector<int64_t> set_q0(nsets + 1, 0);
        std::vector<std::vector<int64_t>> work_prefix(nsets);
        for (int s = 0; s < nsets; ++s) {
            set_q0[s] = static_cast<int64_t>(quads.size());
            std::vector<int64_t>& pre = work_prefix[s];
            pre.push_back(0);
            for (int p = 0; p < nplanes; ++p) {
                if (dp[p].tset != s || !take(p)) continue;
                for (int j = 0; j < dp[p].ny; j += kQuad) {
                    quads.push_back({p, j});
                    const int64_t w = mode == CRS_MODE_FP32C
                                          ? static_cast<int64_t>(kQuad) * dp[p].nxp
                                          : static_cast<int64_t>(std::min(kQuad, dp[p].ny - j)) * dp[p].nx;
                    pre.push_back(pre.back() + w);
                }
            }
        }
        set_q0[nsets] = static_cast<int64_t>(quads.size());
        int64_t max_quads = 1;
        for (int s = 0; s < nsets; ++s) max_quads = std::max<int64_t>(max_quads, set_q0[s + 1] - set_q0[s]);
        int64_t nsplit = std::max<int64_t>(1, ceil_div(target, base_blocks));
        nsplit = std::min<int64_t>(nsplit, max_quads);
        std::vector<int64_t> block_cost;
        const size_t first = blocks.size();
        for (int64_t c = 0; c < nsplit; ++c) {
            for (int s = 0; s < nsets; ++s) {
                const std::vector<int64_t>& pre = work_prefix[s];
                const int64_t nq = static_cast<int64_t>(pre.size()) - 1;
                const int64_t total = pre.back();
                auto cut = [&](int64_t k) -> int64_t {  // first quad whose prefix >= k*total/nsplit
                    if (k >= nsplit) return nq;
                    const int64_t want = (total * k) / nsplit;
                    return static_cast<int64_t>(std::lower_bound(pre.begin(), pre.end(), want) - pre.begin());
                };
                const int64_t u0 = std::min(cut(c), nq), u1 = std::min(cut(c + 1), nq);
                for (int64_t gq = 0; gq < groups; ++gq) {
                    blocks.push_back({s, static_cast<int32_t>(gq * kThreads * PQ),
                                      static_cast<int32_t>(set_q0[s] + u0),
                                      static_cast<int32_t>(set_q0[s] + std::max(u0, u1)),
                                      (slot0 + c) * n_out + static_cast<int64_t>(s) * npil});
                    const int64_t pil = std::min<int64_t>(int64_t(kThreads) * PQ, npil - gq * kThreads * PQ);
                    block_cost.push_back(pil * (pre[std::max(u0, u1)] - pre[u0]));
                }
            }
        }
        // Heaviest work-groups first (quad runs hold whole quads, so runs differ by up to
        // one quad): the XCDs take work-groups round robin in blockIdx order, and the
        // light ones then fill the tail.  Each block names its own output slot.
        std::vector<int32_t> order(block_cost.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int32_t>(i);
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return block_cost[a] > block_cost[b]; });
        std::vector<AdjBlock> sorted(order.size());
        for (size_t i = 0; i < order.size(); ++i) sorted[i] = blocks[first + order[i]];
        std::copy(sorted.begin(), sorted.end(), blocks.begin() + static_cast<std::ptrdiff_t>(first));
        return nsplit;
    };

    std::vector<AdjBlock> blocks, blocks_exact;
    int64_t nsplit = mixed ? add_launch([&](int p) { return !exact[p]; }, 0, blocks)
                           : add_launch([](int) { return true; }, 0, blocks);
    if (mixed) nsplit += add_launch([&](int p) { return exact[p] != 0; }, nsplit, blocks_exact);
    if (quads.size() > static_cast<size_t>(INT32_MAX) ||
        blocks.size() + blocks_exact.size() > static_cast<size_t>(INT32_MAX)) {
        set_error("launch too large");
        return CRS_ERR_INVALID;
    }
    // Staging blocks walk each plane's staged records.
    std::vector<PixBlock> pblocks;
    for (int p = 0; p < nplanes; ++p) {
        const int64_t nrec = static_cast<int64_t>(dp[p].nquads) * dp[p].nxp * kQuad;
        for (int64_t b = 0; b < nrec; b += kThreads) pblocks.push_back({p, static_cast<int32_t>(b)});
    }
    plan.nsplit = nsplit;
    plan.n_out = n_out;
    plan.staged_records = srec;
    plan.nblocks = static_cast<int>(blocks.size());
    plan.nblocks_exact = static_cast<int>(blocks_exact.size());
    plan.npblocks = static_cast<int>(pblocks.size());
    plan.o_planes = plan.tb.add(dp.data(), dp.size() * sizeof(DevPlane));
    plan.o_quads = plan.tb.add(quads.data(), quads.size() * sizeof(AdjQuad));
    plan.o_blocks = plan.tb.add(blocks.data(), blocks.size() * sizeof(AdjBlock));
    plan.o_blocks_exact = plan.tb.add(blocks_exact.data(), blocks_exact.size() * sizeof(AdjBlock));
    plan.o_pblocks = plan.tb.add(pblocks.data(), pblocks.size() * sizeof(PixBlock));
    return CRS_OK;
}

crs_status size_adjoint(const AdjPlan& plan, const Scratch& sc, hipStream_t stream) {
    crs_status st;
    // Positions only (FP32 / FP64; FP32C reads the caller's px, py directly).
    if (plan.mode != CRS_MODE_FP32C &&
        (st = ensure(*sc.pil, staged_pillar_bytes(plan.mode, plan.npil, 1), stream,
                     "staged pillars")) != CRS_OK)
        return st;
    if (plan.mode == CRS_MODE_FP32C &&
        (st = ensure(*sc.pix, static_cast<size_t>(plan.staged_records) * sizeof(float4), stream,
                     "staged pixels")) != CRS_OK)
        return st;
    return ensure(*sc.partial,
                  static_cast<size_t>(plan.nsplit * plan.n_out) * partial_bytes(plan.mode), stream,
                  "adjoint partials");
}

crs_status enqueue_adjoint(const AdjPlan& plan, const Scratch& sc, const double* px,
                           const double* py, const double* u, double* g, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1) {
    crs_status st;
    const int mode = plan.mode;
    const DevPlane* d_planes = reinterpret_cast<const DevPlane*>(sc.tables + plan.o_planes);
    // FP32 / FP64: t = NULL stages positions with zero transmissions.
    const double* const xy[2] = {px, py};
    const void* pil = xy;
    if (mode != CRS_MODE_FP32C) {
        if ((st = launch_status(launch_stage_pillars(mode, px, py, nullptr, plan.npil, 1,
                                                     sc.pil->ptr, stream),
                                "stage positions")) != CRS_OK)
            return st;
        pil = sc.pil->ptr;
    }
    const void* pix_or_u = u;
    if (mode == CRS_MODE_FP32C) {
        if ((st = launch_status(
                 launch_stage_pixels(d_planes,
                                     reinterpret_cast<const PixBlock*>(sc.tables + plan.o_pblocks),
                                     plan.npblocks, reinterpret_cast<const double2*>(u),
                                     static_cast<float4*>(sc.pix->ptr), stream),
                 "stage pixels")) != CRS_OK)
            return st;
        pix_or_u = sc.pix->ptr;
    }
    if (ev0) hipEventRecord(ev0, stream);
    const AdjQuad* d_quads = reinterpret_cast<const AdjQuad*>(sc.tables + plan.o_quads);
    if ((st = launch_status(launch_adjoint(mode, plan.order, pil, plan.npil, d_planes, d_quads,
                                           reinterpret_cast<const AdjBlock*>(sc.tables + plan.o_blocks),
                                           plan.nblocks, pix_or_u, sc.partial->ptr, stream),
                            "adjoint")) != CRS_OK)
        return st;
    if (plan.nblocks_exact > 0 &&
        (st = launch_status(launch_adjoint(mode, kExactOrder, pil, plan.npil, d_planes, d_quads,
                                           reinterpret_cast<const AdjBlock*>(sc.tables + plan.o_blocks_exact),
                                           plan.nblocks_exact, pix_or_u, sc.partial->ptr, stream),
                            "adjoint (planes evaluated pair by pair)")) != CRS_OK)
        return st;
    if (ev1) hipEventRecord(ev1, stream);
    return launch_status(launch_adj_reduce(mode, sc.partial->ptr,
                                           static_cast<int>(plan.nsplit), plan.n_out,
                                           reinterpret_cast<double2*>(g), stream),
                         "adjoint reduce");
}

}  // namespace

extern "C" crs_status crs_forward(crs_context* ctx, void* stream_, int mode, const double* px,
                                  const double* py, const double* t, int npil, int nsets,
                                  const crs_plane* planes, int nplanes, double* field,
                                  double* intensity) {
    g_last_error.clear();
    crs_status st = check_common(ctx, mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!px || !py || !t) { set_error("null pillar array"); return CRS_ERR_INVALID; }
    if (nplanes == 0 || (!field && !intensity)) return CRS_OK;
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    PlanSlot& slot = ctx->slot[0];
    std::vector<unsigned char> key = plan_key(mode, npil, nsets, planes, nplanes, ctx->fwd_target[mode]);
    const bool reuse = slot.valid && slot.key == key;
    if (!reuse) {
        slot.valid = false;
        ctx->fwd_plan = FwdPlan();
        if ((st = plan_forward(ctx, mode, npil, nsets, planes, nplanes, ctx->fwd_plan)) != CRS_OK)
            return st;
    }
    const FwdPlan& plan = ctx->fwd_plan;
    Scratch sc{&ctx->pil, &ctx->pix, &ctx->partial, nullptr};
    if ((st = size_forward(plan, sc, stream)) != CRS_OK) return st;
    if (!reuse) {
        if ((st = upload_tables(ctx, slot.tables, plan.tb, stream)) != CRS_OK) return st;
        slot.key = std::move(key);
        slot.valid = true;
    }
    if (inject("launch")) { set_error("injected launch failure"); return CRS_ERR_LAUNCH; }
    sc.tables = static_cast<unsigned char*>(slot.tables.ptr);
    const bool timed = ctx->timing && !stream_capturing(stream);
    st = enqueue_forward(plan, sc, px, py, t, field, intensity, stream,
                         timed ? ctx->ev[0][0] : nullptr, timed ? ctx->ev[0][1] : nullptr);
    if (st == CRS_OK && timed) ctx->ev_valid[0] = true;
    if (st == CRS_OK) {
        ctx->last_kernel[0] = pair_kernel(mode, plan.order, false);
        ctx->last_groups[0] = plan.nblocks;
    }
    return st;
}

extern "C" crs_status crs_adjoint(crs_context* ctx, void* stream_, int mode, const double* px,
                                  const double* py, int npil, int nsets, const crs_plane* planes,
                                  int nplanes, const double* u, double* g) {
    g_last_error.clear();
    crs_status st = check_common(ctx, mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!px || !py || !g || (nplanes > 0 && !u)) { set_error("null array argument"); return CRS_ERR_INVALID; }
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    PlanSlot& slot = ctx->slot[1];
    std::vector<unsigned char> key = plan_key(mode, npil, nsets, planes, nplanes, ctx->adj_target[mode]);
    const bool reuse = slot.valid && slot.key == key;
    if (!reuse) {
        slot.valid = false;
        ctx->adj_plan = AdjPlan();
        if ((st = plan_adjoint(ctx, mode, npil, nsets, planes, nplanes, ctx->adj_plan)) != CRS_OK)
            return st;
    }
    const AdjPlan& plan = ctx->adj_plan;
    Scratch sc{&ctx->pil, &ctx->pix, &ctx->partial, nullptr};
    if ((st = size_adjoint(plan, sc, stream)) != CRS_OK) return st;
    if (!reuse) {
        if ((st = upload_tables(ctx, slot.tables, plan.tb, stream)) != CRS_OK) return st;
        slot.key = std::move(key);
        slot.valid = true;
    }
    if (inject("launch")) { set_error("injected launch failure"); return CRS_ERR_LAUNCH; }
    sc.tables = static_cast<unsigned char*>(slot.tables.ptr);
    const bool timed = ctx->timing && !stream_capturing(stream);
    st = enqueue_adjoint(plan, sc, px, py, u, g, stream, timed ? ctx->ev[1][0] : nullptr,
                         timed ? ctx->ev[1][1] : nullptr);
    if (st == CRS_OK && timed) ctx->ev_valid[1] = true;
    if (st == CRS_OK) {
        ctx->last_kernel[1] = pair_kernel(mode, plan.order, true);
        ctx->last_groups[1] = plan.nblocks;
    }
    return st;
}

// ------------------------------------------------------------------------------------
// HIP graphs: one forward or adjoint, captured with scratch the graph owns.

struct crs_graph {
    int device = 0;
    DevBuffer pil, pix, tables, partial;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <class Plan, class SizeFn, class EnqueueFn>
crs_status build_graph(crs_context* ctx, Plan& plan, SizeFn size_fn, EnqueueFn enqueue_fn,
                       crs_graph** out) {
    crs_graph* gr = new crs_graph();
    gr->device = ctx->device;
    Scratch sc{&gr->pil, &gr->pix, &gr->partial, nullptr};
    hipStream_t cap = nullptr;
    crs_status st = CRS_OK;
    auto fail = [&](crs_status s) {
        if (cap) hipStreamDestroy(cap);
        if (gr->exec) hipGraphExecDestroy(gr->exec);
        if (gr->graph) hipGraphDestroy(gr->graph);
        for (DevBuffer* b : {&gr->pil, &gr->pix, &gr->tables, &gr->partial})
            if (b->ptr) hipFree(b->ptr);
        delete gr;
        return s;
    };
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
        hipGetLastError();
        set_error("hipStreamCreate failed");
        return fail(CRS_ERR_LAUNCH);
    }
    if ((st = size_fn(plan, sc, cap)) != CRS_OK) return fail(st);
    const size_t nt = plan.tb.bytes.size();
    if ((st = ensure(gr->tables, nt, cap, "graph tables")) != CRS_OK) return fail(st);
    // Tables are copied once, before capture: the graph holds kernel nodes only.
    if (hipMemcpy(gr->tables.ptr, plan.tb.bytes.data(), nt, hipMemcpyHostToDevice) != hipSuccess) {
        hipGetLastError();
        set_error("graph tables: hipMemcpy failed");
        return fail(CRS_ERR_LAUNCH);
    }
    sc.tables = static_cast<unsigned char*>(gr->tables.ptr);
    if (inject("launch")) { set_error("injected launch failure"); return fail(CRS_ERR_LAUNCH); }
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        hipGetLastError();
        set_error("hipStreamBeginCapture failed");
        return fail(CRS_ERR_LAUNCH);
    }
    st = enqueue_fn(plan, sc, cap);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cap, &g);
    gr->graph = g;
    if (st != CRS_OK) return fail(st);
    if (ec != hipSuccess || !g) {
        hipGetLastError();
        set_error("hipStreamEndCapture failed");
        return fail(CRS_ERR_LAUNCH);
    }
    if (hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        hipGetLastError();
        set_error("hipGraphInstantiate failed");
        return fail(CRS_ERR_LAUNCH);
    }
    hipStreamDestroy(cap);
    *out = gr;
    return CRS_OK;
}

}  // namespace

extern "C" crs_status crs_graph_forward(crs_context* ctx, int mode, const double* px,
                                        const double* py, const double* t, int npil, int nsets,
                                        const crs_plane* planes, int nplanes, double* field,
                                        double* intensity, crs_graph** out) {
    g_last_error.clear();
    if (!out) { set_error("null output pointer"); return CRS_ERR_INVALID; }
    *out = nullptr;
    crs_status st = check_common(ctx, mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!px || !py || !t || nplanes == 0 || (!field && !intensity)) {
        set_error("graph needs pillars, at least one plane and an output");
        return CRS_ERR_INVALID;
    }
    FwdPlan plan;
    if ((st = plan_forward(ctx, mode, npil, nsets, planes, nplanes, plan)) != CRS_OK) return st;
    return build_graph(
        ctx, plan, size_forward,
        [&](const FwdPlan& p, const Scratch& sc, hipStream_t s) {
            return enqueue_forward(p, sc, px, py, t, field, intensity, s, nullptr, nullptr);
        },
        out);
}

extern "C" crs_status crs_graph_adjoint(crs_context* ctx, int mode, const double* px,
                                        const double* py, int npil, int nsets,
                                        const crs_plane* planes, int nplanes, const double* u,
                                        double* g, crs_graph** out) {
    g_last_error.clear();
    if (!out) { set_error("null output pointer"); return CRS_ERR_INVALID; }
    *out = nullptr;
    crs_status st = check_common(ctx, mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!px || !py || !g || !u || nplanes == 0) {
        set_error("graph needs pillars, at least one plane, u and g");
        return CRS_ERR_INVALID;
    }
    AdjPlan plan;
    if ((st = plan_adjoint(ctx, mode, npil, nsets, planes, nplanes, plan)) != CRS_OK) return st;
    return build_graph(
        ctx, plan, size_adjoint,
        [&](const AdjPlan& p, const Scratch& sc, hipStream_t s) {
            return enqueue_adjoint(p, sc, px, py, u, g, s, nullptr, nullptr);
        },
        out);
}

extern "C" crs_status crs_graph_launch(crs_graph* g, void* stream) {
    if (!g || !g->exec) { set_error("null graph"); return CRS_ERR_INVALID; }
    if (hipSetDevice(g->device) != hipSuccess ||
        hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) != hipSuccess) {
        hipGetLastError();
        set_error("hipGraphLaunch failed");
        return CRS_ERR_LAUNCH;
    }
    return CRS_OK;
}

extern "C" void crs_graph_destroy(crs_graph* g) {
    if (!g) return;
    hipSetDevice(g->device);
    hipDeviceSynchronize();
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    for (DevBuffer* b : {&g->pil, &g->pix, &g->tables, &g->partial})
        if (b->ptr) hipFree(b->ptr);
    delete g;
}

// ------------------------------------------------------------------------------------
// Drop-in layer: the reference's propagate.hpp semantics on top of crs_forward.

namespace {

std::atomic<int> g_default_mode{CRS_MODE_FP32C};

// One drop-in call's device state.  States live in a process-wide pool: a call takes
// one, uses it on its own stream and returns it, so concurrent callers never share
// buffers and nothing is torn down at thread or process exit.
struct DropinState {
    crs_context* ctx = nullptr;
    hipStream_t stream = nullptr;
    DevBuffer in, out;
    std::vector<double> host_out;
};

std::mutex g_pool_mu;
std::vector<DropinState*> g_pool;

DropinState* acquire_state() {
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (!g_pool.empty()) {
            DropinState* s = g_pool.back();
            g_pool.pop_back();
            return s;
        }
    }
    DropinState* s = new DropinState();
    if (crs_context_create(0, &s->ctx) != CRS_OK) { delete s; return nullptr; }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        hipGetLastError();
        crs_context_destroy(s->ctx);
        delete s;
        set_error("hipStreamCreate failed");
        return nullptr;
    }
    return s;
}

void release_state(DropinState* s) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(s);
}

// Runs planes through crs_forward with host inputs and host outputs (complex field
// and/or |E|^2, each `total` elements).  Outputs are written only when every step
// succeeded.
crs_status host_run(int mode, const double* px, const double* py, const double* t, int npil,
                    int nsets, const crs_plane* planes, int nplanes, double* field,
                    double* intensity) {
    int64_t total = 0;
    for (int p = 0; p < nplanes; ++p)
        total = std::max<int64_t>(total, planes[p].offset + static_cast<int64_t>(planes[p].nx) * planes[p].ny);
    DropinState* s = acquire_state();
    if (!s) return CRS_ERR_NO_DEVICE;
    const size_t pil_bytes = static_cast<size_t>(npil) * sizeof(double);
    const size_t t_bytes = 2 * pil_bytes * static_cast<size_t>(nsets);
    const size_t f_elems = field ? 2 * static_cast<size_t>(total) : 0;
    const size_t i_elems = intensity ? static_cast<size_t>(total) : 0;
    crs_status st = ensure(s->in, 2 * pil_bytes + t_bytes, s->stream, "host inputs");
    if (st == CRS_OK)
        st = ensure(s->out, (f_elems + i_elems) * sizeof(double), s->stream, "host outputs");
    if (st == CRS_OK) {
        double* d_px = static_cast<double*>(s->in.ptr);
        double* d_py = d_px + npil;
        double* d_t = d_py + npil;
        double* d_f = field ? static_cast<double*>(s->out.ptr) : nullptr;
        double* d_i = intensity ? static_cast<double*>(s->out.ptr) + f_elems : nullptr;
        if (hipMemcpyAsync(d_px, px, pil_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(d_py, py, pil_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
            hipMemcpyAsync(d_t, t, t_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
            hipGetLastError();
            set_error("host inputs: hipMemcpyAsync failed");
            st = CRS_ERR_LAUNCH;
        }
        if (st == CRS_OK)
            st = crs_forward(s->ctx, s->stream, mode, d_px, d_py, d_t, npil, nsets, planes, nplanes,
                             d_f, d_i);
        if (st == CRS_OK) {
            s->host_out.resize(f_elems + i_elems);
            if (hipMemcpyAsync(s->host_out.data(), s->out.ptr, (f_elems + i_elems) * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                hipStreamSynchronize(s->stream) != hipSuccess) {
                hipGetLastError();
                set_error("host outputs: copy or synchronize failed");
                st = CRS_ERR_LAUNCH;
            } else {
                if (field) std::memcpy(field, s->host_out.data(), f_elems * sizeof(double));
                if (intensity)
                    std::memcpy(intensity, s->host_out.data() + f_elems, i_elems * sizeof(double));
            }
        }
    }
    if (st != CRS_OK) {
        hipGetLastError();
        hipStreamSynchronize(s->stream);
    }
    release_state(s);
    return st;
}

// One plane of the drop-in layer: |E|^2 into `out`, nonzero on success.
int dropin_run(const double* px, const double* py, const double* pt, int npil,
               const crs_plane& plane, double* out) {
    return host_run(g_default_mode.load(), px, py, pt, npil, 1, &plane, 1, nullptr, out) == CRS_OK;
}

}  // namespace

extern "C" int crs_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        hipGetLastError();
        return 0;
    }
    return count > 0 ? 1 : 0;
}

extern "C" const char* crs_device_name(void) {
    static std::once_flag once;
    static char name[256] = {0};
    std::call_once(once, [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) { hipGetLastError(); return; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) { hipGetLastError(); return; }
        std::snprintf(name, sizeof name, "%s", prop.name);
    });
    return name;
}

extern "C" crs_status crs_forward_host(int mode, const double* px, const double* py,
                                       const double* t, int npil, int nsets,
                                       const crs_plane* planes, int nplanes, double* field,
                                       double* intensity) {
    g_last_error.clear();
    if (!px || !py || !t || (!field && !intensity)) { set_error("null array argument"); return CRS_ERR_INVALID; }
    if (nplanes <= 0 || !planes) { set_error("no planes"); return CRS_ERR_INVALID; }
    // Sizes and planes are checked here, before host_run sizes buffers from them.
    const crs_status st = check_args(mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!crs_available()) { set_error("no HIP device visible"); return CRS_ERR_NO_DEVICE; }
    return host_run(mode, px, py, t, npil, nsets, planes, nplanes, field, intensity);
}

extern "C" crs_status crs_analyze_mtf(const double* pupil, const double* pupil0, int n,
                                      double period_um, double focal_um, double wavelength_um,
                                      double diameter_um, int cap, float* freq_cyc_mm,
                                      float* mtf, float* mtf_ideal, double* cutoff_cyc_mm,
                                      int* npoints) {
    g_last_error.clear();
    if (!pupil || !pupil0 || !freq_cyc_mm || !mtf || !mtf_ideal || !cutoff_cyc_mm || !npoints ||
        n <= 0 || !(period_um > 0.0)) {
        set_error("bad MTF arguments");
        return CRS_ERR_INVALID;
    }
    // Shifts s = 0, 1, ... while s <= int(D/p) + 1 and s < n  (mtf.cpp:53-54).
    const double q = diameter_um / period_um;
    if (std::isnan(q)) { set_error("bad MTF arguments"); return CRS_ERR_INVALID; }
    const int64_t max_s = q >= n ? n : static_cast<int64_t>(q) + 1;  // int() truncates
    const int ns = static_cast<int>(std::max<int64_t>(0, std::min<int64_t>(max_s, n - 1) + 1));
    if (ns == 0) {
        *cutoff_cyc_mm = diameter_um / (wavelength_um * focal_um) * 1000.0;
        *npoints = 0;
        return CRS_OK;
    }
    if (ns > cap) {
        set_error("MTF needs %d points, capacity %d", ns, cap);
        return CRS_ERR_INVALID;
    }
    if (!crs_available()) { set_error("no HIP device visible"); return CRS_ERR_NO_DEVICE; }
    DropinState* st_ = acquire_state();
    if (!st_) return CRS_ERR_NO_DEVICE;
    const size_t nn = static_cast<size_t>(n) * n;
    crs_status st = ensure(st_->in, nn * 3 * sizeof(double), st_->stream, "MTF pupil");
    if (st == CRS_OK)
        st = ensure(st_->out, static_cast<size_t>(ns) * (3 * sizeof(double) + 3 * sizeof(float)),
                    st_->stream, "MTF sums");
    if (st == CRS_OK) {
        double2* d_P = static_cast<double2*>(st_->in.ptr);
        double* d_P0 = reinterpret_cast<double*>(d_P + nn);
        double2* d_acc = static_cast<double2*>(st_->out.ptr);
        double* d_acc0 = reinterpret_cast<double*>(d_acc + ns);
        float* d_f = reinterpret_cast<float*>(d_acc0 + ns);
        std::vector<float> h(3 * static_cast<size_t>(ns));
        if (hipMemcpyAsync(d_P, pupil, nn * sizeof(double2), hipMemcpyHostToDevice, st_->stream) != hipSuccess ||
            hipMemcpyAsync(d_P0, pupil0, nn * sizeof(double), hipMemcpyHostToDevice, st_->stream) != hipSuccess) {
            hipGetLastError();
            set_error("MTF pupil: hipMemcpyAsync failed");
            st = CRS_ERR_LAUNCH;
        }
        if (st == CRS_OK &&
            (st = launch_status(launch_mtf(d_P, d_P0, n, ns, period_um, focal_um, wavelength_um,
                                           d_acc, d_acc0, d_f, d_f + ns, d_f + 2 * ns, st_->stream),
                                "MTF")) == CRS_OK) {
            if (hipMemcpyAsync(h.data(), d_f, h.size() * sizeof(float), hipMemcpyDeviceToHost,
                               st_->stream) != hipSuccess ||
                hipStreamSynchronize(st_->stream) != hipSuccess) {
                hipGetLastError();
                set_error("MTF: copy or synchronize failed");
                st = CRS_ERR_LAUNCH;
            } else {
                std::memcpy(freq_cyc_mm, h.data(), ns * sizeof(float));
                std::memcpy(mtf, h.data() + ns, ns * sizeof(float));
                std::memcpy(mtf_ideal, h.data() + 2 * ns, ns * sizeof(float));
                *cutoff_cyc_mm = diameter_um / (wavelength_um * focal_um) * 1000.0;
                *npoints = ns;
            }
        }
    }
    if (st != CRS_OK) {
        hipGetLastError();
        hipStreamSynchronize(st_->stream);
    }
    release_state(st_);
    return st;
}

extern "C" int crs_set_default_mode(int mode) {
    if (mode < 0 || mode > 2) return -1;
    return g_default_mode.exchange(mode);
}

extern "C" int crs_propagate_psf(const double* px, const double* py, const double* pt, int npil,
                                 double cx, double cy, double z, double k, int n,
                                 double half_window, double* out) {
    if (!crs_available()) return 0;
    if (npil <= 0 || n <= 0) return 0;
    if (!px || !py || !pt || !out) return 0;
    const double step = 2.0 * half_window / (n - 1);
    crs_plane plane{};
    plane.k = k;
    plane.x0 = cx - half_window;
    plane.sx = step;
    plane.y0 = cy - half_window;
    plane.sy = step;
    plane.z0 = z;
    plane.sz = 0.0;
    plane.nx = n;
    plane.ny = n;
    return dropin_run(px, py, pt, npil, plane, out);
}

extern "C" int crs_propagate_zx(const double* px, const double* py, const double* pt, int npil,
                                double xmin, double xmax, int nx, double zmin, double zmax, int nz,
                                double k, double* out) {
    if (!crs_available()) return 0;
    if (npil <= 0 || nx <= 1 || nz <= 1) return 0;
    if (!px || !py || !pt || !out) return 0;
    crs_plane plane{};
    plane.k = k;
    plane.x0 = xmin;
    plane.sx = (xmax - xmin) / (nx - 1);
    plane.y0 = 0.0;
    plane.sy = 0.0;
    plane.z0 = zmin;
    plane.sz = (zmax - zmin) / (nz - 1);
    plane.nx = nx;
    plane.ny = nz;
    return dropin_run(px, py, pt, npil, plane, out);
}
