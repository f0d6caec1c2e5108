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
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace crs {
hipError_t launch_stage_pillars(int mode, const double* px, const double* py, const double* t,
                                int npil, int nsets, void* out, hipStream_t s);
hipError_t launch_forward_fp32c(int bx, int by, bool rsi, const Fp32cFwdArgs& a, hipStream_t s);
hipError_t launch_fwd_reduce_fp32c(const DevPlane* planes, const FwdTile* tiles,
                                   const FwdRed* reds, int nred, const float2* partial,
                                   int64_t slot_elems, double2* field, double* intensity,
                                   hipStream_t s);
hipError_t launch_adjoint_fp32c(int bx, int by, bool rsi, const Fp32cAdjArgs& a, hipStream_t s);
hipError_t launch_adj_reduce_fp32c(const float2* partial, const int32_t* slot_off,
                                   const int32_t* slots, int npil, int ngroups, int64_t n,
                                   double2* g, hipStream_t s);
int fp32c_resident_per_cu(bool adjoint, int bx, int by, bool rsi);
const char* fp32c_kernel_name(bool adjoint, int bx, int by, bool rsi);
hipError_t launch_forward(int mode, const void* pil, int npil, const DevPlane* planes,
                          const FwdBlock* blocks, int nblocks, void* partial, hipStream_t s);
hipError_t launch_fwd_reduce(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                             const void* partial, double2* field, double* intensity,
                             hipStream_t s);
hipError_t launch_stage_pixels(const DevPlane* planes, const PixBlock* pblocks, int nblocks,
                               const double2* u, float4* pix, hipStream_t s);
hipError_t launch_adjoint(int mode, const void* pil, int npil, const DevPlane* planes,
                          const AdjQuad* quads, const AdjBlock* blocks, int nblocks,
                          const void* u, void* partial, hipStream_t s);
hipError_t launch_adj_reduce(const void* partial, int nslots, int64_t n, double2* g,
                             hipStream_t s);
hipError_t launch_mtf(const double2* P, const double* P0, int n, int ns, double p, double f,
                      double lam, double2* acc, double* acc0, float* freq, float* mtf,
                      float* ideal, hipStream_t s);
hipError_t noalloc_scratch(void** ptr);
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

// One FP32C pair-kernel launch of a plan: block shape and its work-group table.
struct Fp32cLaunch {
    int bx = 1, by = 1;
    bool rsi = false;  // RS-I planes in blocks (their own kernel)
    int nwg = 0;
    size_t o_wg = 0;  // int32 wg_seg[nwg + 1]
};

struct FwdPlan {
    int mode = 0;
    int npil = 0, nsets = 0;
    TableBuilder tb;
    size_t o_planes = 0;
    // FP32 / FP64: uniform split.
    int64_t part_elems = 0;
    int nblocks = 0, npblocks = 0;
    size_t o_blocks = 0, o_pblocks = 0;
    // FP32C: stream-K segments.
    std::vector<Fp32cLaunch> launches;
    int64_t nseg = 0, slot_elems = 0;
    int nred = 0;
    size_t o_tiles = 0, o_segs = 0, o_reds = 0;
};

struct AdjPlan {
    int mode = 0;
    int npil = 0, nsets = 0;
    int64_t n_out = 0, staged_records = 0;
    TableBuilder tb;
    size_t o_planes = 0, o_pblocks = 0;
    int npblocks = 0;
    // FP32 / FP64: uniform split over quad runs.
    int64_t nsplit = 1;
    int nblocks = 0;
    size_t o_quads = 0, o_blocks = 0;
    // FP32C: stream-K segments with per-(set, group) slot lists.
    std::vector<Fp32cLaunch> launches;
    int64_t nseg = 0;
    int ngroups = 0;
    size_t o_strips = 0, o_segs = 0, o_slot_off = 0, o_slots = 0;
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
    int cus = 0;
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
    int fwd_shape[2] = {0, 0};  // FP32C block shape (0, 0: built-in default)
    int adj_shape[2] = {0, 0};
    bool timing = false;
    hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [adjoint][start/stop]
    bool ev_valid[2] = {false, false};
    const char* last_kernel[2] = {nullptr, nullptr};  // [adjoint]: last pair kernel launched
    int last_groups[2] = {0, 0};
    // Completion of the last call's work and its stream: a call on another stream first
    // waits for it (the scratch and tables are shared), and scratch is freed only after it.
    // A call on a created stream records done_ev as it ends; a call on the null stream,
    // which lives as long as the process, leaves null_pending set instead, and the event
    // is recorded on the null stream only when a later call or a scratch release needs it
    // (an event record costs ~4 us of GPU time per call, 12% of a 4,096-pillar x 81^2
    // forward + adjoint step).
    hipEvent_t done_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool done_valid = false;
    bool null_pending = false;
};

namespace {

// Built-in launch targets of the FP32 / FP64 kernels: work-groups per launch.  256 CUs
// x a few resident work-groups, several waves deep so the tail of uneven splits is short.
int default_target() { return 4096; }

// Built-in FP32C block shapes: 4 x 4 pixels per forward lane (4 x 8 blocks, 3 waves per
// SIMD, run as fast at 161^2 and 3% faster at 321^2, but their second-order loops reload
// from scratch, so the autotuner, not the default, picks them), 8 x 8 per adjoint block
// (the separable form where its bound holds: 36.8 against 43.1 cycles per 64 pairs for
// the 4 x 8 series at 91,809 pillars x 321^2).
constexpr int kDefFwdBX = 4, kDefFwdBY = 4;
constexpr int kDefAdjBX = 8, kDefAdjBY = 8;
// Built-in FP32C launch sizes, in multiples of the resident work-groups (4 per CU for
// both default shapes).  Several rounds let the CUs that finish early take more work:
// at 91,809 pillars x 161^2 (whole call, pair kernel and reduction) the forward takes
// 0.795 / 0.745 / 0.720 / 0.715 ms at 1 / 2 / 4 / 8 rounds, the adjoint 0.751 / 0.711 /
// 0.693 / 0.687 ms.
constexpr int kFwdRounds = 8;
constexpr int kAdjRounds = 8;
// Fewest pairs per FP32C work-group piece: below this a piece's fixed costs (pillar
// loads, state, the partial slot and its reduction) outweigh its pairs, so small calls
// launch fewer, larger pieces (measured at 4,096 pillars x 81^2: 512 adjoint pieces take
// 0.6x the time of 3,696).
constexpr int64_t kMinPiecePairs = 65536;

// Records done_ev for a previous call on the null stream that has not recorded it yet.
void settle_null_pending(crs_context* ctx) {
    if (!ctx || !ctx->null_pending) return;
    ctx->null_pending = false;
    if (hipEventRecord(ctx->done_ev, nullptr) == hipSuccess) ctx->done_valid = true;
    else hipGetLastError();
}

// The previous call's work must be finished before its scratch is freed.
void wait_previous(crs_context* ctx) {
    settle_null_pending(ctx);
    if (ctx && ctx->done_valid) hipEventSynchronize(ctx->done_ev);
}

crs_status ensure(crs_context* ctx, DevBuffer& b, size_t bytes, hipStream_t stream,
                  const char* what) {
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
        // The old buffer may still be read by queued work on this stream or, through the
        // context, on the stream of its previous call.
        hipStreamSynchronize(stream);
        wait_previous(ctx);
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
    crs_status st = ensure(ctx, dst, n, stream, "descriptor tables");
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
                                    int nplanes, int target, const int shape[2]) {
    std::vector<unsigned char> key(7 * sizeof(int) + static_cast<size_t>(nplanes) * sizeof(crs_plane));
    const int head[7] = {mode, npil, nsets, nplanes, target, shape[0], shape[1]};
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
        if (P.offset < 0) {
            set_error("plane %d: negative offset", p);
            return CRS_ERR_INVALID;
        }
        if (P.kernel != CRS_KERNEL_REF && P.kernel != CRS_KERNEL_RSI) {
            set_error("plane %d: unknown kernel %d", p, P.kernel);
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

// Orders a call on `stream` after the context's previous call when that ran on another
// stream: the scratch buffers and descriptor tables it reads are shared by the calls.
void order_after_previous(crs_context* ctx, hipStream_t stream) {
    if (ctx->last_stream == stream || stream_capturing(stream)) return;
    settle_null_pending(ctx);
    if (ctx->done_valid) hipStreamWaitEvent(stream, ctx->done_ev, 0);
}

// Marks the end of a call's device work on `stream` (outside capture).
void record_done(crs_context* ctx, hipStream_t stream) {
    if (stream_capturing(stream)) return;
    ctx->last_stream = stream;
    if (stream == nullptr) {
        ctx->null_pending = true;
        ctx->done_valid = false;
        return;
    }
    ctx->null_pending = false;
    if (hipEventRecord(ctx->done_ev, stream) == hipSuccess) {
        ctx->done_valid = true;
    } else {
        ctx->done_valid = false;
        hipGetLastError();
    }
}

DevPlane make_dev_plane(const crs_plane& P) {
    DevPlane d{};
    d.k = P.k;
    d.kturn = P.k * kInv2Pi;
    d.x0 = P.x0; d.sx = P.sx;
    d.y0 = P.y0; d.sy = P.sy;
    d.z0 = P.z0; d.sz = P.sz;
    d.zref = d.kturn * P.z0;
    d.nx = P.nx; d.ny = P.ny;
    d.tx = 0;
    d.ty = 0;
    d.tset = P.tset;
    d.nslots = 1;
    d.offset = P.offset;
    d.part = 0;
    d.sbase = 0;
    d.nxp = P.nx;
    d.nquads = static_cast<int32_t>(ceil_div(P.ny, kQuad));
    d.bx = 1;
    d.by = 1;
    d.obl1 = 1.0;
    d.obl2 = 1.0;
    d.kernel = P.kernel;
    return d;
}

// Whether bx x by blocks of plane d meet the FP32C pair budget at the highest series
// order for pillars at any obliquity (the host does not know where the pillars are;
// the kernels then pick the cheapest order each work item's pillars admit).  ny_rows:
// rows whose depth counts (the adjoint's padded strips reach beyond ny).
bool fp32c_block_admitted(const DevPlane& d, int bx, int by, int ny_rows) {
    DevPlane q = d;
    q.ny = ny_rows;
    const SeriesGeom g = series_geom(q, bx, by);
    if (!std::isfinite(g.a) || !std::isfinite(g.c) || !(g.z > 0.0)) return false;
    return fp32c_pair_error(g, 1.0, kP2A2, d.kernel == CRS_KERNEL_RSI) <= kPairBudget;
}

// FP32C relative cost of one pillar against one lane of a thread tile of nel pixels
// (pair ops plus the per-pillar FP64 distance and state), for the stream-K cuts.
double fwd_tile_rate(int nel) { return nel == 1 ? 48.0 : 7.0 * nel + 40.0; }

// Work-groups of an FP32C launch resident on the whole chip at once.
int64_t fp32c_slots(const crs_context* ctx, bool adjoint, int bx, int by, bool rsi = false) {
    int per_cu = fp32c_resident_per_cu(adjoint, bx, by, rsi);
    if (per_cu <= 0) per_cu = 2;
    return static_cast<int64_t>(per_cu) * std::max(1, ctx->cus);
}

// Work-groups of an FP32C launch: the tuned target, else `rounds` x the resident slots.
int64_t fp32c_target(const crs_context* ctx, int tuned, bool adjoint, int bx, int by, int rounds,
                     bool rsi = false) {
    if (tuned > 0) return tuned;
    return fp32c_slots(ctx, adjoint, bx, by, rsi) * rounds;
}

// Bytes of staged pillar data: FP32C keeps FP64 positions once plus {|t|, arg t} per
// set and per-tile extents; FP32 and FP64 keep one {x, y, Re t, Im t} record per
// pillar and set.
size_t staged_pillar_bytes(int mode, int64_t npil, int64_t nsets) {
    if (mode == CRS_MODE_FP32C) return fp32c_staged_layout(npil, nsets).bytes;
    return npil * nsets * (mode == CRS_MODE_FP64 ? sizeof(double4) : sizeof(float4));
}

crs_status launch_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return CRS_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return CRS_ERR_LAUNCH;
}

const char* plain_kernel(int mode, bool adjoint) {
    if (mode == CRS_MODE_FP64) return adjoint ? "crs::adj_fp64" : "crs::fwd_fp64";
    return adjoint ? "crs::adj_fp32" : "crs::fwd_fp32";
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
    if (hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess) {
        hipGetLastError();
        ctx->cus = 256;
    }
    for (hipEvent_t& e : ctx->tables_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            crs_context_destroy(ctx);
            set_error("hipEventCreate failed");
            return CRS_ERR_LAUNCH;
        }
    }
    if (hipEventCreateWithFlags(&ctx->done_ev, hipEventDisableTiming) != hipSuccess) {
        hipGetLastError();
        crs_context_destroy(ctx);
        set_error("hipEventCreate failed");
        return CRS_ERR_LAUNCH;
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
    if (ctx->done_ev) hipEventDestroy(ctx->done_ev);
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

namespace {
bool fwd_shape_ok(int bx, int by) {
    return (bx == 0 && by == 0) || (bx == 4 && by == 4) || (bx == 4 && by == 8) ||
           (bx == 8 && by == 4);
}
bool adj_shape_ok(int bx, int by) {
    return (bx == 0 && by == 0) || ((bx == 4 || bx == 8) && (by == 4 || by == 8));
}
}  // namespace

extern "C" crs_status crs_set_block_shape(crs_context* ctx, int fwd_bx, int fwd_by, int adj_bx,
                                          int adj_by) {
    if (!ctx || !fwd_shape_ok(fwd_bx, fwd_by) || !adj_shape_ok(adj_bx, adj_by)) {
        set_error("bad block shape (forward %d x %d, adjoint %d x %d)", fwd_bx, fwd_by, adj_bx,
                  adj_by);
        return CRS_ERR_INVALID;
    }
    ctx->fwd_shape[0] = fwd_bx;
    ctx->fwd_shape[1] = fwd_by;
    ctx->adj_shape[0] = adj_bx;
    ctx->adj_shape[1] = adj_by;
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
    crs_context* owner;     // the context whose previous call may still use the buffers
};

// Piece costs of a launch of nwg work-groups with `slots` resident at a time.  The
// hardware hands out work-groups in index order as slots free up, so a launch of equal
// pieces ends with a ragged last round: at 91,809 pillars x 161^2, 4 x 4 blocks in 8,192
// equal pieces, the slots held a work-group for 0.92 of the kernel's span (per-work-group
// clock stamps), the gap spread over the last ~100 us.  With nwg >= 2 slots the costs
// fall geometrically instead, c_i ~ exp(-x i / nwg), x fixed by making the last piece a
// quarter of the mean (the first is 2.6 times the mean): the final round is short (0.96
// busy, a 25 us tail; an eighth gave 0.95), and the piece count, hence the partial-sum
// traffic, stays nwg.
// Returns the cumulative cost at each of the nwg - 1 cuts.
std::vector<double> piece_bounds(double total, int64_t nwg, int64_t slots) {
    std::vector<double> b;
    if (nwg <= 1) return b;
    b.resize(static_cast<size_t>(nwg - 1));
    if (slots <= 0 || nwg < 2 * slots) {
        for (int64_t i = 1; i < nwg; ++i)
            b[static_cast<size_t>(i - 1)] = total * static_cast<double>(i) / static_cast<double>(nwg);
        return b;
    }
    // x / (e^x - 1) = 1/4: the last piece's cost over the mean.
    const double x = 2.3366629822630536;
    const double q = std::exp(-x / static_cast<double>(nwg));  // ratio of successive pieces
    // cumulative share after i pieces: (1 - q^i) / (1 - q^nwg)
    const double denom = 1.0 - std::pow(q, static_cast<double>(nwg));
    for (int64_t i = 1; i < nwg; ++i)
        b[static_cast<size_t>(i - 1)] =
            total * (1.0 - std::pow(q, static_cast<double>(i))) / denom;
    return b;
}

// Stream-K cut of a work line into work-groups.  The line is a sequence of runs (run r
// has len[r] units of cost rate[r] each); `nwg` pieces of the costs piece_bounds gives
// (equal ones below 2 slots' worth) are cut at unit positions that are multiples of
// `gran` within a run (no empty segment; a piece that would be empty merges into its
// neighbour, so fewer pieces may result).  emit(run, u0, u1) is called for each segment
// in order; wg_start receives the index of each piece's first segment followed by the
// total segment count.
template <class Emit>
void streamk_cut(const std::vector<int64_t>& len, const std::vector<double>& rate, int64_t nwg,
                 int64_t slots, int64_t gran, Emit emit, std::vector<int32_t>& wg_start) {
    double total = 0.0;
    for (size_t r = 0; r < len.size(); ++r) total += rate[r] * static_cast<double>(len[r]);
    nwg = std::max<int64_t>(1, nwg);
    const std::vector<double> bnd = piece_bounds(total, nwg, slots);
    auto bound = [&](int64_t k) { return bnd[static_cast<size_t>(k - 1)]; };  // 1 <= k < nwg
    int64_t nseg = 0;
    wg_start.clear();
    wg_start.push_back(0);
    int64_t next = 1;  // index of the next piece boundary, at cost per * next
    double cum = 0.0;
    for (size_t r = 0; r < len.size(); ++r) {
        const int64_t n = len[r];
        if (n <= 0) continue;
        const double rt = rate[r];
        const double end_cost = cum + rt * static_cast<double>(n);
        int64_t u = 0;
        while (u < n) {
            int64_t uend = n;
            bool cut = false;
            if (next < nwg) {
                const double b = bound(next);
                if (b < end_cost) {
                    int64_t uc = static_cast<int64_t>(std::llround((b - cum) / rt / gran)) * gran;
                    if (uc <= u) uc = u + gran;
                    if (uc < n) {
                        uend = uc;
                        cut = true;
                    }
                }
            }
            emit(static_cast<int>(r), u, uend);
            ++nseg;
            u = uend;
            if (cut) {
                wg_start.push_back(static_cast<int32_t>(nseg));
                ++next;
            }
        }
        cum = end_cost;
        // Boundaries at or before this run's end that fell on its last unit start the
        // next piece at the next run.
        while (next < nwg && bound(next) <= cum * (1.0 + 1e-12)) {
            if (wg_start.back() != nseg) wg_start.push_back(static_cast<int32_t>(nseg));
            ++next;
        }
    }
    if (wg_start.back() == nseg && wg_start.size() > 1) wg_start.pop_back();
    wg_start.push_back(static_cast<int32_t>(nseg));
}

// FP32C forward: block shape per plane (the tuned shape where it meets the pair budget,
// else 4 x 4, else exact), tiles per launch, stream-K segments per launch.
crs_status plan_forward_fp32c(const crs_context* ctx, int npil, std::vector<DevPlane>& dp,
                              FwdPlan& plan) {
    const int nplanes = static_cast<int>(dp.size());
    int want_bx = ctx->fwd_shape[0] ? ctx->fwd_shape[0] : kDefFwdBX;
    int want_by = ctx->fwd_shape[1] ? ctx->fwd_shape[1] : kDefFwdBY;
    // Launches: the tuned shape, 4 x 4, RS-I planes in 4 x 4 blocks, exact planes.
    const int shapes[4][2] = {{want_bx, want_by}, {4, 4}, {4, 4}, {1, 1}};
    std::vector<int> pick(nplanes, 3);
    for (int p = 0; p < nplanes; ++p) {
        if (dp[p].kernel == CRS_KERNEL_RSI) {
            if (fp32c_block_admitted(dp[p], 4, 4, dp[p].ny)) pick[p] = 2;
            continue;
        }
        if (fp32c_block_admitted(dp[p], want_bx, want_by, dp[p].ny)) pick[p] = 0;
        else if (fp32c_block_admitted(dp[p], 4, 4, dp[p].ny)) pick[p] = 1;
    }
    if (want_bx == 4 && want_by == 4)
        for (int& k : pick) k = k == 1 ? 0 : k;
    std::vector<FwdTile> tiles;
    std::vector<FwdSeg> segs;
    std::vector<int32_t> wg_all;
    int64_t slot_elems = 64;
    for (int L = 0; L < 4; ++L) {
        const int kbx = shapes[L][0], kby = shapes[L][1];
        const bool rsi = L == 2;
        std::vector<int64_t> len;
        std::vector<double> rate;
        const size_t tile0 = tiles.size();
        for (int p = 0; p < nplanes; ++p) {
            if (pick[p] != L) continue;
            DevPlane& d = dp[p];
            d.bx = kbx;
            d.by = kby;
            d.tx = kbx > 1 ? d.nx / kbx : 0;
            d.ty = kby > 1 ? d.ny / kby : 0;
            d.obl1 = kbx > 1 ? fp32c_order_obliquity(d, kbx, kby, kP1A1) : 1.0;
            d.obl2 = kbx > 1 ? std::max(d.obl1, fp32c_order_obliquity(d, kbx, kby, kP1A2)) : 1.0;
            for (int k = 0; k < 4; ++k) {
                const int64_t cnt = tile_count(d, k, kbx, kby);
                const int nel = k == 0 ? kbx * kby : k == 1 ? kby : k == 2 ? kbx : 1;
                for (int64_t b = 0; b < cnt; b += kFwdTiles) {
                    tiles.push_back({p, k, static_cast<int32_t>(b), nel, 0, 0});
                    len.push_back(npil);
                    rate.push_back(fwd_tile_rate(nel));
                }
            }
        }
        if (tiles.size() == tile0) continue;
        slot_elems = std::max<int64_t>(slot_elems, 64 * kbx * kby);
        // Enough pieces to fill the chip, each at least 64 pillars and kMinPiecePairs pairs.
        int64_t units = 0, pairs = 0;
        for (size_t r = 0; r < len.size(); ++r) {
            units += ceil_div(len[r], 64);
            pairs += len[r] * 64 * tiles[tile0 + r].nel;
        }
        const int64_t nwg = std::min<int64_t>(
            fp32c_target(ctx, ctx->fwd_target[CRS_MODE_FP32C], false, kbx, kby, kFwdRounds, rsi),
            std::max<int64_t>(1, std::min(units, pairs / kMinPiecePairs)));
        std::vector<int32_t> wg_start;
        const int64_t seg_base = static_cast<int64_t>(segs.size());
        int last_tile = -1;
        streamk_cut(len, rate, nwg, fp32c_slots(ctx, false, kbx, kby, rsi), 16,
                    [&](int r, int64_t u0, int64_t u1) {
                        const int t = static_cast<int>(tile0) + r;
                        if (t != last_tile) {
                            tiles[t].seg0 = static_cast<int32_t>(segs.size());
                            last_tile = t;
                        }
                        ++tiles[t].nseg;
                        segs.push_back({t, static_cast<int32_t>(u0), static_cast<int32_t>(u1), 0});
                    },
                    wg_start);
        Fp32cLaunch La;
        La.bx = kbx;
        La.by = kby;
        La.rsi = rsi;
        La.nwg = static_cast<int>(wg_start.size()) - 1;
        for (int32_t& w : wg_start) w += static_cast<int32_t>(seg_base);
        La.o_wg = wg_all.size();
        wg_all.insert(wg_all.end(), wg_start.begin(), wg_start.end());
        plan.launches.push_back(La);
    }
    if (segs.size() > static_cast<size_t>(INT32_MAX) ||
        static_cast<double>(segs.size()) * slot_elems * 8.0 > 64e9) {
        set_error("launch too large (%zu segments)", segs.size());
        return CRS_ERR_INVALID;
    }
    std::vector<FwdRed> reds;
    for (size_t t = 0; t < tiles.size(); ++t)
        for (int el = 0; el < tiles[t].nel * 64; el += 64)
            reds.push_back({static_cast<int32_t>(t), el});
    plan.nseg = static_cast<int64_t>(segs.size());
    plan.slot_elems = slot_elems;
    plan.nred = static_cast<int>(reds.size());
    plan.o_planes = plan.tb.add(dp.data(), dp.size() * sizeof(DevPlane));
    plan.o_tiles = plan.tb.add(tiles.data(), tiles.size() * sizeof(FwdTile));
    plan.o_segs = plan.tb.add(segs.data(), segs.size() * sizeof(FwdSeg));
    plan.o_reds = plan.tb.add(reds.data(), reds.size() * sizeof(FwdRed));
    const size_t o_wg = plan.tb.add(wg_all.data(), wg_all.size() * sizeof(int32_t));
    for (Fp32cLaunch& La : plan.launches) La.o_wg = o_wg + La.o_wg * sizeof(int32_t);
    return CRS_OK;
}

// FP32 / FP64 forward: pillar chunks split uniformly; work-groups of kThreads pixels.
crs_status plan_forward_plain(const crs_context* ctx, int mode, int npil,
                              std::vector<DevPlane>& dp, FwdPlan& plan) {
    const int nplanes = static_cast<int>(dp.size());
    int64_t base_blocks = 0;
    for (int p = 0; p < nplanes; ++p)
        base_blocks += ceil_div(static_cast<int64_t>(dp[p].nx) * dp[p].ny, kThreads);
    const int target = ctx->fwd_target[mode] > 0 ? ctx->fwd_target[mode] : default_target();
    int64_t nsplit = std::max<int64_t>(1, ceil_div(target, base_blocks));
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
            for (int64_t b = 0; b < npix; b += kThreads)
                blocks.push_back({p, static_cast<int32_t>(b), q0, q1, dp[p].part + s * npix, 0, 0});
        }
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
    for (int p = 0; p < nplanes; ++p) dp[p] = make_dev_plane(planes[p]);
    if (mode == CRS_MODE_FP32C) return plan_forward_fp32c(ctx, npil, dp, plan);
    return plan_forward_plain(ctx, mode, npil, dp, plan);
}

// Bytes of one split-sum partial: FP32C keeps each work item's sum as a float pair,
// rounded once from its FP64 combine; FP64 and FP32 keep doubles.
size_t partial_bytes(int mode) {
    return mode == CRS_MODE_FP32C ? sizeof(float2) : sizeof(double2);
}

crs_status size_forward(const FwdPlan& plan, const Scratch& sc, hipStream_t stream) {
    crs_status st = ensure(sc.owner, *sc.pil, staged_pillar_bytes(plan.mode, plan.npil, plan.nsets),
                           stream, "staged pillars");
    if (st != CRS_OK) return st;
    const int64_t elems = plan.mode == CRS_MODE_FP32C ? plan.nseg * plan.slot_elems : plan.part_elems;
    return ensure(sc.owner, *sc.partial, static_cast<size_t>(elems) * partial_bytes(plan.mode),
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
    if (mode == CRS_MODE_FP32C) {
        const StagedLayout L = fp32c_staged_layout(plan.npil, plan.nsets);
        unsigned char* base = static_cast<unsigned char*>(sc.pil->ptr);
        Fp32cFwdArgs a;
        a.pos = reinterpret_cast<const double2*>(base + L.pos);
        a.ap = reinterpret_cast<const float2*>(base + L.ap);
        a.ext = reinterpret_cast<const double2*>(base + L.ext);
        a.npil = plan.npil;
        a.planes = d_planes;
        a.tiles = reinterpret_cast<const FwdTile*>(sc.tables + plan.o_tiles);
        a.segs = reinterpret_cast<const FwdSeg*>(sc.tables + plan.o_segs);
        a.partial = static_cast<float2*>(sc.partial->ptr);
        a.slot_elems = plan.slot_elems;
        for (const Fp32cLaunch& La : plan.launches) {
            a.wg_seg = reinterpret_cast<const int32_t*>(sc.tables + La.o_wg);
            a.nwg = La.nwg;
            if ((st = launch_status(launch_forward_fp32c(La.bx, La.by, La.rsi, a, stream), "forward")) != CRS_OK)
                return st;
        }
        if (ev1) hipEventRecord(ev1, stream);
        return launch_status(
            launch_fwd_reduce_fp32c(d_planes, a.tiles,
                                    reinterpret_cast<const FwdRed*>(sc.tables + plan.o_reds),
                                    plan.nred, a.partial, plan.slot_elems,
                                    reinterpret_cast<double2*>(field), intensity, stream),
            "forward reduce");
    }
    if ((st = launch_status(launch_forward(mode, sc.pil->ptr, plan.npil, d_planes,
                                           reinterpret_cast<const FwdBlock*>(sc.tables + plan.o_blocks),
                                           plan.nblocks, sc.partial->ptr, stream),
                            "forward")) != CRS_OK)
        return st;
    if (ev1) hipEventRecord(ev1, stream);
    return launch_status(launch_fwd_reduce(d_planes,
                                           reinterpret_cast<const PixBlock*>(sc.tables + plan.o_pblocks),
                                           plan.npblocks, sc.partial->ptr,
                                           reinterpret_cast<double2*>(field), intensity, stream),
                         "forward reduce");
}

// FP32C adjoint: strips of kStrip rows; block shape per plane as for the forward (with
// the strips' padded rows counted for depth); per launch a strip list per set and
// stream-K segments over the (set, pillar group, strip columns) work line.
crs_status plan_adjoint_fp32c(const crs_context* ctx, int npil, int nsets,
                              std::vector<DevPlane>& dp, AdjPlan& plan) {
    const int nplanes = static_cast<int>(dp.size());
    int64_t srec = 0;
    for (int p = 0; p < nplanes; ++p) {
        dp[p].nxp = static_cast<int32_t>(ceil_div(dp[p].nx, kColAlign) * kColAlign);
        dp[p].nquads = static_cast<int32_t>(ceil_div(dp[p].ny, kStrip));
        dp[p].sbase = srec;
        srec += static_cast<int64_t>(dp[p].nquads) * dp[p].nxp * kStrip;
    }
    if (srec > (int64_t(1) << 31)) { set_error("too many pixels"); return CRS_ERR_INVALID; }
    const int want_bx = ctx->adj_shape[0] ? ctx->adj_shape[0] : kDefAdjBX;
    const int want_by = ctx->adj_shape[1] ? ctx->adj_shape[1] : kDefAdjBY;
    // Launches: the tuned shape, 4 x 4, RS-I planes in 4 x 4 blocks, exact planes.
    const int shapes[4][2] = {{want_bx, want_by}, {4, 4}, {4, 4}, {1, 1}};
    std::vector<int> pick(nplanes, 3);
    for (int p = 0; p < nplanes; ++p) {
        const int rows = dp[p].nquads * kStrip;
        if (dp[p].kernel == CRS_KERNEL_RSI) {
            if (fp32c_block_admitted(dp[p], 4, 4, rows)) pick[p] = 2;
            continue;
        }
        DevPlane q = dp[p];
        q.ny = rows;
        // 8 x 8 blocks take the separable form where its bound holds at every obliquity
        // (the staging then writes the plane's records in its layout), else the series.
        if (want_bx * want_by == 64 && sep_plane_admitted(q, want_bx, want_by)) {
            pick[p] = 0;
            dp[p].sep = 1;
        } else if (fp32c_block_admitted(dp[p], want_bx, want_by, rows)) {
            pick[p] = 0;
        } else if (fp32c_block_admitted(dp[p], 4, 4, rows)) {
            pick[p] = 1;
        }
    }
    if (want_bx == 4 && want_by == 4)
        for (int& k : pick) k = k == 1 ? 0 : k;
    for (int p = 0; p < nplanes; ++p) {
        DevPlane& d = dp[p];
        d.bx = shapes[pick[p]][0];
        d.by = shapes[pick[p]][1];
        DevPlane q = d;
        q.ny = d.nquads * kStrip;  // the strips' padded rows count for the nearest depth
        d.obl1 = d.bx > 1 ? fp32c_order_obliquity(q, d.bx, d.by, kP1A1) : 1.0;
        d.obl2 = d.obl1;
    }
    const int ngroups = static_cast<int>(ceil_div(npil, kThreads));
    std::vector<AdjStrip> strips;
    std::vector<AdjSeg> segs;
    std::vector<int32_t> wg_all;
    // Slots of each (set, group), in launch order then segment order.
    std::vector<std::vector<int32_t>> gslots(static_cast<size_t>(nsets) * ngroups);
    for (int L = 0; L < 4; ++L) {
        const int bx = shapes[L][0], by = shapes[L][1];
        const bool rsi = L == 2;
        // Per set: its strips (planes in order, strips in order) and their block columns.
        std::vector<int64_t> set_u0(nsets), set_nb(nsets, 0);
        std::vector<std::vector<int64_t>> prefix(nsets);  // block-column prefix per set's strips
        bool any = false;
        for (int s = 0; s < nsets; ++s) {
            set_u0[s] = static_cast<int64_t>(strips.size());
            prefix[s].push_back(0);
            for (int p = 0; p < nplanes; ++p) {
                if (pick[p] != L || dp[p].tset != s) continue;
                any = true;
                const int64_t nb = ceil_div(dp[p].nx, bx);
                for (int j = 0; j < dp[p].ny; j += kStrip) {
                    strips.push_back({p, j});
                    prefix[s].push_back(prefix[s].back() + nb);
                }
            }
            set_nb[s] = prefix[s].back();
        }
        if (!any) continue;
        // Runs: (set, group) in order, each the set's block columns; cost per block column:
        // its pairs and per-block state (1 x 1: the rows a strip really has).
        std::vector<int64_t> len;
        std::vector<double> rate;
        std::vector<std::pair<int, int>> run_sg;
        const double r_blk = bx * by == 1 ? 8.0 * 48.0 : 8.0 * bx * 7.0 + (kStrip / by) * 60.0;
        for (int s = 0; s < nsets; ++s)
            for (int g = 0; g < ngroups; ++g) {
                if (set_nb[s] == 0) continue;
                len.push_back(set_nb[s]);
                rate.push_back(r_blk);
                run_sg.push_back({s, g});
            }
        int64_t units = 0;
        for (int64_t l : len) units += l;
        const int64_t pairs = units * bx * kStrip * kThreads;
        const int64_t nwg = std::min<int64_t>(
            fp32c_target(ctx, ctx->adj_target[CRS_MODE_FP32C], true, bx, by, kAdjRounds, rsi),
            std::max<int64_t>(1, std::min(units, pairs / kMinPiecePairs)));
        std::vector<int32_t> wg_start;
        const int64_t seg_base = static_cast<int64_t>(segs.size());
        streamk_cut(len, rate, nwg, fp32c_slots(ctx, true, bx, by, rsi), 1,
                    [&](int r, int64_t b0, int64_t b1) {
                        const int s = run_sg[r].first, g = run_sg[r].second;
                        const std::vector<int64_t>& pre = prefix[s];
                        // strip holding block column b: last k with pre[k] <= b
                        auto strip_of = [&](int64_t b) {
                            return static_cast<int64_t>(std::upper_bound(pre.begin(), pre.end(), b) - pre.begin()) - 1;
                        };
                        const int64_t k0 = strip_of(b0), k1 = strip_of(b1 - 1);
                        AdjSeg S;
                        S.tset = s;
                        S.q0 = g * kThreads;
                        S.u0 = static_cast<int32_t>(set_u0[s] + k0);
                        S.c0 = static_cast<int32_t>((b0 - pre[k0]) * bx);
                        S.u1 = static_cast<int32_t>(set_u0[s] + k1);
                        S.c1 = static_cast<int32_t>((b1 - pre[k1]) * bx);
                        S.slot = static_cast<int32_t>(segs.size());
                        S.pad = 0;
                        gslots[static_cast<size_t>(s) * ngroups + g].push_back(S.slot);
                        segs.push_back(S);
                    },
                    wg_start);
        Fp32cLaunch La;
        La.bx = bx;
        La.by = by;
        La.rsi = rsi;
        La.nwg = static_cast<int>(wg_start.size()) - 1;
        for (int32_t& w : wg_start) w += static_cast<int32_t>(seg_base);
        La.o_wg = wg_all.size();
        wg_all.insert(wg_all.end(), wg_start.begin(), wg_start.end());
        plan.launches.push_back(La);
    }
    if (segs.size() > static_cast<size_t>(INT32_MAX) || strips.size() > static_cast<size_t>(INT32_MAX)) {
        set_error("launch too large");
        return CRS_ERR_INVALID;
    }
    std::vector<int32_t> slot_off(1, 0), slots;
    for (const std::vector<int32_t>& v : gslots) {
        slots.insert(slots.end(), v.begin(), v.end());
        slot_off.push_back(static_cast<int32_t>(slots.size()));
    }
    std::vector<PixBlock> pblocks;
    for (int p = 0; p < nplanes; ++p) {
        const int64_t nrec = static_cast<int64_t>(dp[p].nquads) * dp[p].nxp * kStrip;
        for (int64_t b = 0; b < nrec; b += kThreads) pblocks.push_back({p, static_cast<int32_t>(b)});
    }
    plan.staged_records = srec;
    plan.nseg = static_cast<int64_t>(segs.size());
    plan.ngroups = ngroups;
    plan.npblocks = static_cast<int>(pblocks.size());
    plan.o_planes = plan.tb.add(dp.data(), dp.size() * sizeof(DevPlane));
    plan.o_strips = plan.tb.add(strips.data(), strips.size() * sizeof(AdjStrip));
    plan.o_segs = plan.tb.add(segs.data(), segs.size() * sizeof(AdjSeg));
    plan.o_slot_off = plan.tb.add(slot_off.data(), slot_off.size() * sizeof(int32_t));
    plan.o_slots = plan.tb.add(slots.data(), std::max<size_t>(1, slots.size()) * sizeof(int32_t));
    plan.o_pblocks = plan.tb.add(pblocks.data(), pblocks.size() * sizeof(PixBlock));
    const size_t o_wg = plan.tb.add(wg_all.data(), wg_all.size() * sizeof(int32_t));
    for (Fp32cLaunch& La : plan.launches) La.o_wg = o_wg + La.o_wg * sizeof(int32_t);
    return CRS_OK;
}

// FP32 / FP64 adjoint: quads of 4 rows, runs split uniformly.
crs_status plan_adjoint_plain(const crs_context* ctx, int mode, int npil, int nsets,
                              std::vector<DevPlane>& dp, AdjPlan& plan) {
    const int nplanes = static_cast<int>(dp.size());
    const int target = ctx->adj_target[mode] > 0 ? ctx->adj_target[mode] : default_target();
    const int64_t groups = ceil_div(npil, kThreads);
    const int64_t base_blocks = groups * nsets;
    const int64_t n_out = static_cast<int64_t>(nsets) * npil;
    std::vector<AdjQuad> quads;
    std::vector<int64_t> set_q0(nsets + 1, 0);
    std::vector<std::vector<int64_t>> work_prefix(nsets);
    for (int s = 0; s < nsets; ++s) {
        set_q0[s] = static_cast<int64_t>(quads.size());
        std::vector<int64_t>& pre = work_prefix[s];
        pre.push_back(0);
        for (int p = 0; p < nplanes; ++p) {
            if (dp[p].tset != s) continue;
            for (int j = 0; j < dp[p].ny; j += kQuad) {
                quads.push_back({p, j});
                pre.push_back(pre.back() + static_cast<int64_t>(std::min(kQuad, dp[p].ny - j)) * dp[p].nx);
            }
        }
    }
    set_q0[nsets] = static_cast<int64_t>(quads.size());
    int64_t max_quads = 1;
    for (int s = 0; s < nsets; ++s) max_quads = std::max<int64_t>(max_quads, set_q0[s + 1] - set_q0[s]);
    int64_t nsplit = std::max<int64_t>(1, ceil_div(target, base_blocks));
    nsplit = std::min<int64_t>(nsplit, max_quads);
    std::vector<AdjBlock> blocks;
    for (int64_t c = 0; c < nsplit; ++c) {
        for (int s = 0; s < nsets; ++s) {
            const std::vector<int64_t>& pre = work_prefix[s];
            const int64_t nq = static_cast<int64_t>(pre.size()) - 1;
            const int64_t total = pre.back();
            auto cut = [&](int64_t k) -> int64_t {
                if (k >= nsplit) return nq;
                const int64_t want = (total * k) / nsplit;
                return static_cast<int64_t>(std::lower_bound(pre.begin(), pre.end(), want) - pre.begin());
            };
            const int64_t u0 = std::min(cut(c), nq), u1 = std::min(cut(c + 1), nq);
            for (int64_t gq = 0; gq < groups; ++gq)
                blocks.push_back({s, static_cast<int32_t>(gq * kThreads),
                                  static_cast<int32_t>(set_q0[s] + u0),
                                  static_cast<int32_t>(set_q0[s] + std::max(u0, u1)),
                                  c * n_out + static_cast<int64_t>(s) * npil});
        }
    }
    if (quads.size() > static_cast<size_t>(INT32_MAX) || blocks.size() > static_cast<size_t>(INT32_MAX)) {
        set_error("launch too large");
        return CRS_ERR_INVALID;
    }
    plan.nsplit = nsplit;
    plan.nblocks = static_cast<int>(blocks.size());
    plan.o_planes = plan.tb.add(dp.data(), dp.size() * sizeof(DevPlane));
    plan.o_quads = plan.tb.add(quads.data(), quads.size() * sizeof(AdjQuad));
    plan.o_blocks = plan.tb.add(blocks.data(), blocks.size() * sizeof(AdjBlock));
    return CRS_OK;
}

crs_status plan_adjoint(const crs_context* ctx, int mode, int npil, int nsets,
                        const crs_plane* planes, int nplanes, AdjPlan& plan) {
    plan.mode = mode;
    plan.npil = npil;
    plan.nsets = nsets;
    plan.n_out = static_cast<int64_t>(nsets) * npil;
    std::vector<DevPlane> dp(nplanes);
    for (int p = 0; p < nplanes; ++p) dp[p] = make_dev_plane(planes[p]);
    if (mode == CRS_MODE_FP32C) return plan_adjoint_fp32c(ctx, npil, nsets, dp, plan);
    return plan_adjoint_plain(ctx, mode, npil, nsets, dp, plan);
}

crs_status size_adjoint(const AdjPlan& plan, const Scratch& sc, hipStream_t stream) {
    crs_status st;
    if (plan.mode != CRS_MODE_FP32C) {
        // Positions only (FP32C reads the caller's px, py directly).
        if ((st = ensure(sc.owner, *sc.pil, staged_pillar_bytes(plan.mode, plan.npil, 1), stream,
                         "staged pillars")) != CRS_OK)
            return st;
        return ensure(sc.owner, *sc.partial,
                      static_cast<size_t>(plan.nsplit * plan.n_out) * partial_bytes(plan.mode),
                      stream, "adjoint partials");
    }
    if ((st = ensure(sc.owner, *sc.pix, static_cast<size_t>(plan.staged_records) * sizeof(float4),
                     stream, "staged pixels")) != CRS_OK)
        return st;
    return ensure(sc.owner, *sc.partial,
                  static_cast<size_t>(plan.nseg) * kThreads * partial_bytes(plan.mode), stream,
                  "adjoint partials");
}

crs_status enqueue_adjoint(const AdjPlan& plan, const Scratch& sc, const double* px,
                           const double* py, const double* u, double* g, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1) {
    crs_status st;
    const int mode = plan.mode;
    const DevPlane* d_planes = reinterpret_cast<const DevPlane*>(sc.tables + plan.o_planes);
    if (mode == CRS_MODE_FP32C) {
        if ((st = launch_status(
                 launch_stage_pixels(d_planes,
                                     reinterpret_cast<const PixBlock*>(sc.tables + plan.o_pblocks),
                                     plan.npblocks, reinterpret_cast<const double2*>(u),
                                     static_cast<float4*>(sc.pix->ptr), stream),
                 "stage pixels")) != CRS_OK)
            return st;
        if (ev0) hipEventRecord(ev0, stream);
        Fp32cAdjArgs a;
        a.px = px;
        a.py = py;
        a.npil = plan.npil;
        a.planes = d_planes;
        a.strips = reinterpret_cast<const AdjStrip*>(sc.tables + plan.o_strips);
        a.segs = reinterpret_cast<const AdjSeg*>(sc.tables + plan.o_segs);
        a.pix = static_cast<const float4*>(sc.pix->ptr);
        a.partial = static_cast<float2*>(sc.partial->ptr);
        for (const Fp32cLaunch& La : plan.launches) {
            a.wg_seg = reinterpret_cast<const int32_t*>(sc.tables + La.o_wg);
            a.nwg = La.nwg;
            if ((st = launch_status(launch_adjoint_fp32c(La.bx, La.by, La.rsi, a, stream), "adjoint")) != CRS_OK)
                return st;
        }
        if (ev1) hipEventRecord(ev1, stream);
        return launch_status(
            launch_adj_reduce_fp32c(a.partial,
                                    reinterpret_cast<const int32_t*>(sc.tables + plan.o_slot_off),
                                    reinterpret_cast<const int32_t*>(sc.tables + plan.o_slots),
                                    plan.npil, plan.ngroups, plan.n_out,
                                    reinterpret_cast<double2*>(g), stream),
            "adjoint reduce");
    }
    // FP32 / FP64: t = NULL stages positions with zero transmissions.
    if ((st = launch_status(launch_stage_pillars(mode, px, py, nullptr, plan.npil, 1, sc.pil->ptr,
                                                 stream),
                            "stage positions")) != CRS_OK)
        return st;
    if (ev0) hipEventRecord(ev0, stream);
    if ((st = launch_status(launch_adjoint(mode, sc.pil->ptr, plan.npil, d_planes,
                                           reinterpret_cast<const AdjQuad*>(sc.tables + plan.o_quads),
                                           reinterpret_cast<const AdjBlock*>(sc.tables + plan.o_blocks),
                                           plan.nblocks, u, sc.partial->ptr, stream),
                            "adjoint")) != CRS_OK)
        return st;
    if (ev1) hipEventRecord(ev1, stream);
    return launch_status(launch_adj_reduce(sc.partial->ptr, static_cast<int>(plan.nsplit),
                                           plan.n_out, reinterpret_cast<double2*>(g), stream),
                         "adjoint reduce");
}

// Profiler name and work-group count of a plan's (first) pair-kernel launch.
template <class Plan>
void note_launch(crs_context* ctx, const Plan& plan, bool adjoint) {
    const int a = adjoint ? 1 : 0;
    if (plan.mode == CRS_MODE_FP32C) {
        if (plan.launches.empty()) return;
        ctx->last_kernel[a] = fp32c_kernel_name(adjoint, plan.launches[0].bx, plan.launches[0].by,
                                                plan.launches[0].rsi);
        ctx->last_groups[a] = plan.launches[0].nwg;
    } else {
        ctx->last_kernel[a] = plain_kernel(plan.mode, adjoint);
        ctx->last_groups[a] = plan.nblocks;
    }
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
    order_after_previous(ctx, stream);
    PlanSlot& slot = ctx->slot[0];
    std::vector<unsigned char> key =
        plan_key(mode, npil, nsets, planes, nplanes, ctx->fwd_target[mode], ctx->fwd_shape);
    const bool reuse = slot.valid && slot.key == key;
    if (!reuse) {
        slot.valid = false;
        ctx->fwd_plan = FwdPlan();
        if ((st = plan_forward(ctx, mode, npil, nsets, planes, nplanes, ctx->fwd_plan)) != CRS_OK)
            return st;
    }
    const FwdPlan& plan = ctx->fwd_plan;
    Scratch sc{&ctx->pil, &ctx->pix, &ctx->partial, nullptr, ctx};
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
    record_done(ctx, stream);
    if (st == CRS_OK && timed) ctx->ev_valid[0] = true;
    if (st == CRS_OK) note_launch(ctx, plan, false);
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
    order_after_previous(ctx, stream);
    PlanSlot& slot = ctx->slot[1];
    std::vector<unsigned char> key =
        plan_key(mode, npil, nsets, planes, nplanes, ctx->adj_target[mode], ctx->adj_shape);
    const bool reuse = slot.valid && slot.key == key;
    if (!reuse) {
        slot.valid = false;
        ctx->adj_plan = AdjPlan();
        if ((st = plan_adjoint(ctx, mode, npil, nsets, planes, nplanes, ctx->adj_plan)) != CRS_OK)
            return st;
    }
    const AdjPlan& plan = ctx->adj_plan;
    Scratch sc{&ctx->pil, &ctx->pix, &ctx->partial, nullptr, ctx};
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
    record_done(ctx, stream);
    if (st == CRS_OK && timed) ctx->ev_valid[1] = true;
    if (st == CRS_OK) note_launch(ctx, plan, true);
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
    Scratch sc{&gr->pil, &gr->pix, &gr->partial, nullptr, nullptr};
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
    if ((st = ensure(nullptr, gr->tables, nt, cap, "graph tables")) != CRS_OK) return fail(st);
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
// Allocation-free forward: crs_forward with its scratch carved from the static device
// array `scratch` (rs_kernels.hip) instead of device allocations.

namespace {

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Per-device state of the allocation-free entry point, set up on its first call on a
// device: the scratch address, the device's CU count, a host staging buffer for the
// descriptor tables and an event marking when the last upload from it was read.
struct NoAllocDevice {
    std::mutex mu;
    bool ready = false;
    unsigned char* base = nullptr;
    int cus = 0;
    std::vector<unsigned char> host_tables;
    hipEvent_t uploaded = nullptr;
    bool pending = false;
};

NoAllocDevice* noalloc_device(int dev) {
    static NoAllocDevice devs[64];
    return dev >= 0 && dev < 64 ? &devs[dev] : nullptr;
}

}  // namespace

extern "C" crs_status crs_forward_noalloc(void* stream_, int mode, const double* px,
                                          const double* py, const double* t, int npil,
                                          int nsets, const crs_plane* planes, int nplanes,
                                          double* field, double* intensity) {
    g_last_error.clear();
    crs_status st = check_args(mode, npil, nsets, planes, nplanes);
    if (st != CRS_OK) return st;
    if (!px || !py || !t) { set_error("null pillar array"); return CRS_ERR_INVALID; }
    if (nplanes == 0 || (!field && !intensity)) return CRS_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        hipGetLastError();
        set_error("no HIP device");
        return CRS_ERR_NO_DEVICE;
    }
    NoAllocDevice* D = noalloc_device(dev);
    if (!D) { set_error("device %d outside the supported range", dev); return CRS_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(D->mu);
    if (!D->ready) {
        void* p = nullptr;
        if (noalloc_scratch(&p) != hipSuccess || !p) {
            hipGetLastError();
            set_error("hipGetSymbolAddress of the scratch array failed");
            return CRS_ERR_LAUNCH;
        }
        D->base = static_cast<unsigned char*>(p);
        if (hipDeviceGetAttribute(&D->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            hipGetLastError();
            D->cus = 256;
        }
        if (hipEventCreateWithFlags(&D->uploaded, hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            set_error("hipEventCreate failed");
            return CRS_ERR_LAUNCH;
        }
        D->ready = true;
    }
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    // Plan with the built-in shapes and sizes; if the partial sums do not fit the
    // scratch, cut the call into fewer, longer pieces (fewer partial slots).
    crs_context pc;
    pc.device = dev;
    pc.cus = D->cus;
    FwdPlan plan;
    size_t staged = 0, partial = 0, tables = 0;
    for (int attempt = 0;; ++attempt) {
        plan = FwdPlan();
        if ((st = plan_forward(&pc, mode, npil, nsets, planes, nplanes, plan)) != CRS_OK) return st;
        staged = align256(staged_pillar_bytes(mode, npil, nsets));
        const int64_t elems = mode == CRS_MODE_FP32C ? plan.nseg * plan.slot_elems : plan.part_elems;
        partial = align256(static_cast<size_t>(elems) * partial_bytes(mode));
        tables = align256(plan.tb.bytes.size());
        if (staged + partial + tables <= kNoAllocScratchBytes) break;
        const int groups = mode == CRS_MODE_FP32C ? (plan.launches.empty() ? 0 : plan.launches[0].nwg)
                                                  : plan.nblocks;
        if (attempt >= 12 || groups <= 64) {
            set_error("crs_forward_noalloc: the call needs %zu bytes of scratch, more than the "
                      "%zu-byte allocation-free maximum (use crs_forward)",
                      staged + partial + tables, kNoAllocScratchBytes);
            return CRS_ERR_INVALID;
        }
        pc.fwd_target[mode] = std::max(1, groups / 2);
    }
    DevBuffer pil_buf{D->base, staged}, part_buf{D->base + staged, partial};
    unsigned char* d_tables = D->base + staged + partial;
    // The host staging buffer may still be read by the previous call's upload.
    if (D->pending) {
        hipEventSynchronize(D->uploaded);
        D->pending = false;
    }
    D->host_tables.assign(plan.tb.bytes.begin(), plan.tb.bytes.end());
    if (hipMemcpyAsync(d_tables, D->host_tables.data(), D->host_tables.size(), hipMemcpyHostToDevice,
                       stream) != hipSuccess) {
        hipGetLastError();
        set_error("descriptor tables: hipMemcpyAsync failed");
        return CRS_ERR_LAUNCH;
    }
    if (!stream_capturing(stream) && hipEventRecord(D->uploaded, stream) == hipSuccess)
        D->pending = true;
    if (inject("launch")) { set_error("injected launch failure"); return CRS_ERR_LAUNCH; }
    Scratch sc{&pil_buf, nullptr, &part_buf, d_tables, nullptr};
    return enqueue_forward(plan, sc, px, py, t, field, intensity, stream, nullptr, nullptr);
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
    crs_status st = ensure(nullptr, s->in, 2 * pil_bytes + t_bytes, s->stream, "host inputs");
    if (st == CRS_OK)
        st = ensure(nullptr, s->out, (f_elems + i_elems) * sizeof(double), s->stream, "host outputs");
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
    crs_status st = ensure(nullptr, st_->in, nn * 3 * sizeof(double), st_->stream, "MTF pupil");
    if (st == CRS_OK)
        st = ensure(nullptr, st_->out, static_cast<size_t>(ns) * (3 * sizeof(double) + 3 * sizeof(float)),
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
