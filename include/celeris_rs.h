/*
 * celeris_rs.h — C ABI of the MI355X (gfx950) Rayleigh–Sommerfeld far-field library.
 *
 * Computes, for metalens pillars q at (x_q, y_q) with complex transmission t_q and
 * focal-plane pixels p at (fx_p, fy_p, z_p),
 *
 *     E_p = sum_q t_q * exp(i k R_pq) / R_pq,   R_pq = sqrt((fx_p-x_q)^2 + (fy_p-y_q)^2 + z_p^2)
 *
 * (the kernel of src/celeris/cuda/propagate.hpp:4-5 and analysis/focal.cpp:229-233 in
 * the reference) and its adjoint g_q = sum_p conj(G_pq) u_p.  Units are microns; k is
 * 2*pi/lambda in rad/um.
 *
 * Two layers:
 *   1. Drop-in entry points with the reference's exact semantics (host arrays in,
 *      host |E|^2 out, synchronous, thread-safe, nonzero = success / 0 = "fall back to
 *      CPU"): crs_available, crs_device_name, crs_propagate_psf, crs_propagate_zx.
 *      include/celeris/cuda/propagate.hpp wraps them in the reference's C++ names.
 *   2. Batched device-pointer entry points (no reference counterpart): many planes,
 *      wavelengths and transmission sets in one launch, complex-field output, an
 *      adjoint, three arithmetic modes, an explicit stream.  These return a crs_status.
 *
 * Every pointer argument is a plain pointer; no HIP or torch types appear here.
 * Complex arrays are interleaved (re, im) float64 pairs, i.e. std::complex<double>.
 */
#ifndef CELERIS_RS_H
#define CELERIS_RS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of the batched entry points. */
typedef enum crs_status {
    CRS_OK = 0,
    CRS_ERR_NO_DEVICE = 1, /* no HIP device visible */
    CRS_ERR_INVALID = 2,   /* bad argument (sizes, plane descriptors, mode) */
    CRS_ERR_ALLOC = 3,     /* device or pinned allocation failed */
    CRS_ERR_LAUNCH = 4,    /* kernel launch or HIP runtime call failed */
    CRS_ERR_CAPTURE = 5    /* the call would allocate while the stream is being captured */
} crs_status;

/* Arithmetic modes. */
typedef enum crs_mode {
    /* FP32 pair arithmetic; each pixel block's phase reference k*(R0 - z) is formed
     * and reduced mod 2*pi in FP64, and each pixel's offset from it is expanded to the
     * lowest series order whose worst-case error per pair (truncation, float rounding,
     * the hardware sin/cos) stays within 8e-6 of |t|/R; a plane too close or too
     * coarsely sampled for that is evaluated pair by pair, each pair at its own FP64
     * distance (decided per plane).  Forward: float sums over at most 1,024 pillars per
     * wave, then FP64.  Adjoint: float sums per pixel block, Kahan-compensated across
     * pixel tiles, then FP64 across work items.  Default and fastest. */
    CRS_MODE_FP32C = 0,
    /* The reference algorithm in FP64 throughout (sqrt, sincos, 1/R, complex MAC). */
    CRS_MODE_FP64 = 1,
    /* The reference GPU kernel's arithmetic: FP32 R, FP32 k*R, FP32 sincos, FP32
     * accumulation (propagate.cu:68-78). */
    CRS_MODE_FP32 = 2
} crs_mode;

/* Propagation kernels G_pq (per plane). */
typedef enum crs_kernel {
    /* exp(i k R) / R: the reference's kernel (propagate.hpp:4-5, focal.cpp:232).
     * The default, and the only one the drop-in layer uses. */
    CRS_KERNEL_REF = 0,
    /* Rayleigh-Sommerfeld I: exp(i k R) / R * (z / R) * (1 / R - i k), the z-derivative
     * form -d/dz [exp(i k R) / R] (no reference counterpart; opt-in).  In CRS_MODE_FP32C
     * such a plane runs in 4 x 4 pixel blocks where its series bound admits them (kernels
     * crs::fwd_fp32c_rsi<4, 4> / crs::adj_fp32c_rsi<4, 4>), else pair by pair, each pair
     * at its own FP64 distance. */
    CRS_KERNEL_RSI = 1
} crs_kernel;

/* One focal plane (an "instance": one wavelength, one pixel lattice).
 * Pixel (i, j), 0 <= i < nx, 0 <= j < ny, sits at
 *     fx = x0 + i*sx,   fy = y0 + j*sy,   z = z0 + j*sz
 * so a psf window is (x0 = cx-W, sx = 2W/(n-1), y0 = cy-W, sy = sx, sz = 0) and a
 * y=0 caustic slice is (x0 = xmin, sx = xstep, y0 = sy = 0, z0 = zmin, sz = zstep).
 * Its field occupies elements [offset, offset + nx*ny) of the output arrays, row
 * j at offset + j*nx.  tset selects the transmission set (0 <= tset < nsets);
 * kernel is a crs_kernel (0 for the reference's). */
typedef struct crs_plane {
    double k;
    double x0, sx;
    double y0, sy;
    double z0, sz;
    int32_t nx, ny;
    int32_t tset;
    int32_t kernel;
    int64_t offset;
} crs_plane;

typedef struct crs_context crs_context;

/* ---- drop-in layer (reference: src/celeris/cuda/propagate.hpp) ------------------ */

/* Replaces celeris::cuda::available() (propagate.hpp:20, propagate.cu:18-21).
 * Nonzero when at least one HIP device is visible. */
int crs_available(void);

/* Replaces celeris::cuda::device_name() (propagate.hpp:21, propagate.cu:23-31).
 * Name of device 0, or "" when none.  The string stays valid for the process. */
const char *crs_device_name(void);

/* Replaces celeris::cuda::propagate_psf (propagate.hpp:27-30, propagate.cu:133-193).
 * px, py: npil pillar centres (um); pt: npil complex transmissions (interleaved);
 * out: caller-owned n*n doubles, row-major (row j = y), receiving |E|^2 on the grid
 * fx = cx - W + i*2W/(n-1), fy = cy - W + j*2W/(n-1) at distance z.
 * Returns nonzero on success; 0 (out untouched) when no device is visible,
 * npil <= 0 or n <= 0, or any HIP call fails, so the caller can fall back to its
 * CPU loop.  Arithmetic: the mode from crs_set_default_mode (CRS_MODE_FP32C unless
 * changed).  Synchronous; safe to call from several host threads at once (each
 * thread gets its own stream and buffers). */
int crs_propagate_psf(const double *px, const double *py, const double *pt, int npil,
                      double cx, double cy, double z, double k, int n,
                      double half_window, double *out);

/* Replaces celeris::cuda::propagate_zx (propagate.hpp:35-38, propagate.cu:195-253).
 * |E|^2 on an nz x nx grid in the plane y = 0: fx = xmin + ix*(xmax-xmin)/(nx-1),
 * z = zmin + iz*(zmax-zmin)/(nz-1); out row-major nz*nx (row = z index).
 * Returns 0 when no device, npil <= 0, nx <= 1 or nz <= 1, or on any HIP failure. */
int crs_propagate_zx(const double *px, const double *py, const double *pt, int npil,
                     double xmin, double xmax, int nx, double zmin, double zmax, int nz,
                     double k, double *out);

/* Mode used by crs_propagate_psf / crs_propagate_zx in this process (no reference
 * counterpart; the reference always used FP32).  Returns the previous mode, or -1
 * for an unknown mode (then nothing changes). */
int crs_set_default_mode(int mode);

/* ---- batched device-pointer layer (no reference counterpart) --------------------- */

/* Create a context bound to HIP device `device`.  A context owns device scratch
 * (staged pillars, split-sum partials, descriptor tables) that grows on demand and
 * is reused; one context must not be used by two host threads at once.  Calls on one
 * context may use different streams: a call on a stream other than the previous
 * call's first makes its stream wait for the previous call's work (outside stream
 * capture), and scratch is freed only after that work has finished. */
crs_status crs_context_create(int device, crs_context **ctx);
void crs_context_destroy(crs_context *ctx);

/* Message describing the last failure on the calling host thread ("" if none). */
const char *crs_last_error(void);

/* Forward sum.  All array arguments are device pointers:
 *   px, py : npil pillar centres, float64
 *   t      : nsets*npil complex transmissions (set-major, interleaved float64)
 *   field  : complex output, total pixels (max over planes of offset + nx*ny), or NULL
 *   intensity : |E|^2 output of the same length, or NULL
 * planes is a HOST array of nplanes descriptors.  Work is enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) and the call returns without waiting.
 * Inputs must stay valid and unmodified until the stream reaches this work. */
crs_status crs_forward(crs_context *ctx, void *stream, int mode,
                       const double *px, const double *py, const double *t,
                       int npil, int nsets, const crs_plane *planes, int nplanes,
                       double *field, double *intensity);

/* Adjoint (vector-Jacobian product) of crs_forward with respect to t:
 *   g[s*npil + q] = sum over planes with tset == s, over their pixels p, of conj(G_pq) u_p
 * u: complex cotangent per pixel (device, laid out like crs_forward's field);
 * g: nsets*npil complex output (device).  With u = dL/dRe(E) + i dL/dIm(E) for a
 * real loss L, g = dL/dRe(t) + i dL/dIm(t).  Same stream semantics as crs_forward. */
crs_status crs_adjoint(crs_context *ctx, void *stream, int mode,
                       const double *px, const double *py, int npil, int nsets,
                       const crs_plane *planes, int nplanes,
                       const double *u, double *g);

/* crs_forward with no device allocation (no reference counterpart; for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and stream semantics as
 * crs_forward (built-in block shapes and launch sizes), on the current HIP device and
 * without a context: staged pillars, partial sums and descriptor tables are carved from
 * a 64 MiB device array each device holds for the life of the process (set up on the
 * first call on a device).  A call whose scratch would not fit is cut into fewer work
 * items; if it still does not fit, CRS_ERR_INVALID (use crs_forward).  The array keeps
 * the last call's contents and nothing in it needs to start at zero.  Calls that overlap
 * on one device (on different streams or host threads without ordering) share the array
 * and race: serialise them. */
crs_status crs_forward_noalloc(void *stream, int mode, const double *px, const double *py,
                               const double *t, int npil, int nsets, const crs_plane *planes,
                               int nplanes, double *field, double *intensity);

/* crs_forward on HOST arrays (no reference counterpart): inputs are copied to the
 * device, the planes run in one launch, and the outputs (field: complex, and/or
 * intensity, total-pixel length) are copied back before the call returns; they are
 * written only on success.  Synchronous and thread-safe like crs_propagate_psf.  The
 * analysis entry points below are built on it. */
crs_status crs_forward_host(int mode, const double *px, const double *py, const double *t,
                            int npil, int nsets, const crs_plane *planes, int nplanes,
                            double *field, double *intensity);

/* ---- analyses on the propagation (reference: src/celeris/analysis/) -------------------
 * Each takes the aperture pillar list the reference builds from a MetalensDesign and a
 * UnitCellLibrary (focal.cpp:42-57: cells inside the circular aperture, t from the
 * library), runs all of its sums in one batched launch, and does the reference's scalar
 * post-processing on the host.  `mode` as for crs_forward.  Synchronous. */

/* analyze_focus (focal.cpp:32-146).  phi_ideal: each cell's ideal focusing phase
 * (focal.cpp:52-54).  out5 = {strehl, phase_strehl, fwhm_um, diffraction_limit_um,
 * encircled_energy}; intensity (optional) receives the 121 x 121 |E|^2 grid. */
crs_status crs_analyze_focus(const double *px, const double *py, const double *t,
                             const double *phi_ideal, int npil, double focal_um,
                             double wavelength_um, double diameter_um, int mode, double *out5,
                             double *intensity);

/* analyze_through_focus (throughfocus.cpp:15-132): the 161-point on-axis scan and a
 * caustic_n x caustic_n y=0 slice (201 on the reference's GPU path, 81 on its CPU path).
 * z161, I161: axial positions and peak-normalised intensity (float-rounded, as stored
 * by the reference); scalars6 = {z_peak, dof, xmin, xmax, zmin, zmax}; caustic
 * (optional, caustic_n^2, row = z) peak-normalised. */
crs_status crs_analyze_through_focus(const double *px, const double *py, const double *t,
                                     int npil, double focal_um, double wavelength_um,
                                     double diameter_um, int caustic_n, int mode, double *z161,
                                     double *I161, double *scalars6, double *caustic);

/* analyze_chromatic (chromatic.cpp:9-58): n wavelengths x 120 on-axis z samples plus
 * the design-wavelength peak.  out3n = n x {lambda, best focus z, peak / design peak}. */
crs_status crs_analyze_chromatic(const double *px, const double *py, const double *t, int npil,
                                 double design_focal, double design_wavelength,
                                 double lambda_min, double lambda_max, int n, int mode,
                                 double *out3n);

/* analyze_chromatic_dispersive (chromatic.cpp:60-127): as crs_analyze_chromatic, but each
 * wavelength has its own transmissions, from the unit-cell library rebuilt at that
 * wavelength (the reference's build_lib_at(lambda); the library build itself is outside
 * this library, so the caller passes its result).  t_design: the aperture cells'
 * transmissions at design_wavelength (the design peak, chromatic.cpp:100-102); t_bands:
 * n x npil (band-major, interleaved re/im), band i at lambda_min + (lambda_max -
 * lambda_min) i / (n - 1).  n wavelengths x 120 on-axis samples plus the design peak in
 * one launch.  out3n = n x {lambda, best focus z, peak / design peak}. */
crs_status crs_analyze_chromatic_dispersive(const double *px, const double *py,
                                            const double *t_design, const double *t_bands,
                                            int npil, double design_focal,
                                            double design_wavelength, double lambda_min,
                                            double lambda_max, int n, int mode, double *out3n);

/* analyze_field_grid (field.cpp:150-233): tilted-illumination spots on a
 * (2 n_half + 1)^2 grid of field angles up to max_angle_deg, each psf_n^2, plus the
 * on-axis reference spot, all in one launch (one transmission set per field point).
 * out7 = per point {angle_x, angle_y, rel_strehl, fwhm_x, fwhm_y, cx, cy}, row-major
 * with angle_y outer. */
crs_status crs_analyze_field_grid(const double *px, const double *py, const double *t, int npil,
                                  double focal_um, double wavelength_um, double diameter_um,
                                  double max_angle_deg, int n_half, int psf_n, int mode,
                                  double *out7);

/* Field of view, replacing analysis/field.cpp:12-76 (analyze_field_of_view): for each
 * field angle, the peak |E|^2 of an 81 x 81 window at z = f around the chief ray
 * (x = f tan theta) under illumination exp(i k sin(theta) x), relative to the on-axis
 * peak, and the x of that peak.  All windows (and the on-axis reference) in one launch.
 * out3 = per angle {angle_deg, rel_strehl, spot_shift_um}. */
crs_status crs_analyze_field_of_view(const double *px, const double *py, const double *t,
                                     int npil, double focal_um, double wavelength_um,
                                     double diameter_um, const double *angles_deg, int nangles,
                                     int mode, double *out3);

/* Wide field of view behind a remote stop, replacing analysis/field.cpp:78-148
 * (analyze_wide_fov): as crs_analyze_field_of_view, but only the pillars inside the stop
 * patch of diameter stop_diameter_um centred at x = stop_distance_um tan(theta) are lit,
 * and the window sits between f sin(theta) and f tan(theta).  The pillar list is the
 * lens aperture's (lens_diameter_um). */
crs_status crs_analyze_wide_fov(const double *px, const double *py, const double *t, int npil,
                                double focal_um, double wavelength_um, double lens_diameter_um,
                                double stop_diameter_um, double stop_distance_um,
                                const double *angles_deg, int nangles, int mode, double *out3);

/* MTF along x, replacing analysis/mtf.cpp:10-63 (analyze_mtf).  pupil is the exit pupil
 * P[iy*n + ix] = t e^{-i phi_ideal} of the n x n cell grid (complex, interleaved re/im,
 * zero outside the aperture) and pupil0 its aperture mask (1 inside, 0 outside), as
 * mtf.cpp:20-36 builds them.  For shifts s = 0 .. min(int(D/p) + 1, n - 1) writes
 * freq_cyc_mm[s] = s p / (lambda f) * 1000, mtf[s] = |sum P conj(P shifted by s)| /
 * (its s = 0 value) and mtf_ideal[s] (the same for pupil0), as floats like the
 * reference's MtfCurve, plus cutoff_cyc_mm = D / (lambda f) * 1000 and npoints.  The
 * sums run on the GPU in FP64.  CRS_ERR_INVALID if cap < npoints. */
crs_status crs_analyze_mtf(const double *pupil, const double *pupil0, int n, double period_um,
                           double focal_um, double wavelength_um, double diameter_um, int cap,
                           float *freq_cyc_mm, float *mtf, float *mtf_ideal,
                           double *cutoff_cyc_mm, int *npoints);

/* Fabrication tolerance, replacing analysis/tolerance.cpp:12-48 (analyze_tolerance).
 * For each sigma (nm of CD error) and each of max(1, trials) trials: every fill of the
 * n_cells^2 map (row-major) gets N(0, (sigma/1000)/period) added from a std::mt19937
 * seeded seed + trial * 7919 + unsigned(sigma * 13) (one std::normal_distribution draw
 * per cell, in map order) and is clamped to [0.05, 0.95]; the cells inside the aperture
 * take the nearest library fill's transmission (metalens.cpp:112-120), and the trial's
 * Strehl ratio is the peak of analyze_focus's 121 x 121 grid at z = f over the ideal
 * lens's peak (focal.cpp:32-146).  All trials run as transmission sets of batched
 * launches.  out4 = per sigma {sigma_nm, mean, standard deviation (population), worst};
 * strehl (optional) = every trial's ratio, sigma-major.  lib_* : the unit-cell library's
 * nlib sampled fills, phases (rad) and amplitudes.  Synchronous. */
crs_status crs_analyze_tolerance(const double *fill_map, int n_cells, double period_um,
                                 const double *lib_fill, const double *lib_phase,
                                 const double *lib_amp, int nlib, double focal_um,
                                 double wavelength_um, double diameter_um,
                                 const double *sigmas_nm, int nsigmas, int trials,
                                 unsigned seed, int mode, double *out4, double *strehl);

/* ---- on-axis focus scans -------------------------------------------------------------
 * The reference's design checks find a lens's focus by scanning the on-axis intensity
 * I(z) = |sum_q t_q exp(i k R_q) / R_q|^2, R_q = sqrt(x_q^2 + y_q^2 + z^2), over z and
 * taking the first maximum, each in an FP64 CPU loop: verify_achromatic_focus
 * (design/achromatic.cpp:288-338), verify_pb_achromatic_focus
 * (design/pb_achromatic.cpp:255-301), the polarisation check (cli/cmd_polar.cpp:85-126,
 * gui/workers.cpp:181-197) and the profile proof (cli/cli_common.cpp:112-133).  Here
 * every scan of a call is one plane of one batched launch. */

/* One scan: transmission set `tset` (0 .. nsets-1) at wavelength_um (k = 2 pi /
 * wavelength_um), nz >= 1 samples z_j = zlo + zspan * j / (nz - 1) (z_0 = zlo when
 * nz == 1), the callers' own expression: zspan = zhi - zlo for the verify and polarisation
 * scans, zspan = f for the profile proof.  best_z is the first sample whose I beats every
 * earlier one and -1 (the reference's `I > best_I` from best_I = -1); z_default when
 * none does (every I is NaN). */
typedef struct crs_axis_scan_desc {
    int tset;
    int nz;
    double wavelength_um;
    double zlo, zspan;
    double z_default;
} crs_axis_scan_desc;

/* Batched on-axis scans (no single reference counterpart: the loop the callers above
 * share).  t: nsets x npil complex transmissions (interleaved re/im, set-major).
 * best_z[s], best_I[s] per scan; curves (optional) receives every I, scan after scan
 * (sum of nz).  Synchronous. */
crs_status crs_axis_scan(const double *px, const double *py, const double *t, int npil,
                         int nsets, const crs_axis_scan_desc *scans, int nscans, int mode,
                         double *best_z, double *best_I, double *curves);

/* verify_achromatic_focus (design/achromatic.cpp:288-338) and
 * verify_pb_achromatic_focus (design/pb_achromatic.cpp:255-301), which differ only in
 * how each band's transmissions are built (the caller builds them: polar(amplitude,
 * phase) of the site's atom, or its cross transmission times exp(i phi_geo)).
 * t_bands: nb x npil complex (the aperture cells, band-major).  For each band j: 200
 * on-axis samples over [0.4 f, 1.8 f] at wavelengths_um[j]; out3nb = nb x {lambda, best
 * z, peak / design peak}, the design peak being I(f) of band center_index at
 * center_wavelength_um (rel = 0 when that peak is not positive). */
crs_status crs_verify_achromatic_focus(const double *px, const double *py,
                                       const double *t_bands, int npil, int nb,
                                       const double *wavelengths_um, int center_index,
                                       double center_wavelength_um, double focal_um, int mode,
                                       double *out3nb);

/* Polarisation-multiplexed focus check (cli/cmd_polar.cpp:85-126 with nz = 240,
 * gui/workers.cpp:181-197 with nz = 200): tx, ty the X and Y channel transmissions of
 * the aperture cells; each channel's on-axis focus over [0.5 min(fx, fy), 1.5 max(fx,
 * fy)], then the focal isolation 10 log10(I_own(z) / max(I_other(z), 1e-30)) at each
 * channel's focus.  out4 = {z_x, z_y, iso_x_db, iso_y_db}; the isolations are 0 when
 * |fx - fy| <= 1e-6 (the reference prints none then). */
crs_status crs_polar_focus_check(const double *px, const double *py, const double *tx,
                                 const double *ty, int npil, double wavelength_um,
                                 double focal_x_um, double focal_y_um, int nz, int mode,
                                 double *out4);

/* Launch-shape tuning (no reference counterpart).  target_blocks: work-groups a
 * forward/adjoint launch aims for; 0 restores the built-in default.  FP32C cuts each
 * call's work into that many equal-cost pieces (default: the pair kernel's resident
 * work-groups on every CU); FP32 / FP64 split their sums over pillar chunks (forward)
 * or pixel rows (adjoint) to reach it.  Applies to the context's later calls. */
crs_status crs_set_tuning(crs_context *ctx, int mode, int fwd_target_blocks,
                          int adj_target_blocks);

/* FP32C block shape (no reference counterpart): pixels per lane of the forward
 * (fwd_bx columns x fwd_by rows: 4 x 4, 4 x 8 or 8 x 4) and per pixel block of the
 * adjoint (adj_bx x adj_by, each 4 or 8); 0, 0 restores the built-in forward 4 x 4 and
 * adjoint 8 x 8.  Adjoint 8 x 8 blocks take the separable form (rank-2 column x row
 * factors) on planes whose separable pair-error bound holds, else the series orders.  A
 * plane whose pair-error bound does not admit the shape falls back to 4 x 4, then to pair
 * by pair.  CRS_ERR_INVALID for other shapes.  Applies to the context's later calls. */
crs_status crs_set_block_shape(crs_context *ctx, int fwd_bx, int fwd_by, int adj_bx,
                               int adj_by);

/* HIP graphs (no reference counterpart).  Capture one crs_forward / crs_adjoint of
 * this exact shape and these exact device pointers into an instantiated graph.  The
 * graph owns its scratch and descriptor tables (copied to the device before capture),
 * so it holds kernel nodes only and later calls on `ctx` do not disturb it.  Replays
 * with crs_graph_launch recompute from the current contents of the bound inputs. */
typedef struct crs_graph crs_graph;
crs_status crs_graph_forward(crs_context *ctx, int mode, const double *px, const double *py,
                             const double *t, int npil, int nsets, const crs_plane *planes,
                             int nplanes, double *field, double *intensity, crs_graph **graph);
crs_status crs_graph_adjoint(crs_context *ctx, int mode, const double *px, const double *py,
                             int npil, int nsets, const crs_plane *planes, int nplanes,
                             const double *u, double *g, crs_graph **graph);
crs_status crs_graph_launch(crs_graph *graph, void *stream);
void crs_graph_destroy(crs_graph *graph);

/* Kernel timing (no reference counterpart).  With timing enabled, each crs_forward /
 * crs_adjoint records HIP events on its stream immediately before and after the
 * kernels that do the pair sums (a call whose planes need several block shapes runs
 * one such kernel per shape, and the events bracket them all);
 * crs_last_kernel_ms waits for the last such event and returns that duration in
 * milliseconds (adjoint = 0: forward). */
crs_status crs_set_timing(crs_context *ctx, int enable);
crs_status crs_last_kernel_ms(crs_context *ctx, int adjoint, float *ms);

/* Name (as a profiler lists it, demangled prefix) of the kernel that does the pair
 * sums for `mode`, forward (adjoint = 0) or adjoint (adjoint = 1); "" if unknown. */
const char *crs_kernel_name(int mode, int adjoint);

/* The pair-sum kernel the last successful crs_forward (adjoint = 0) or crs_adjoint
 * (adjoint = 1) on ctx launched: its name as a profiler lists it (demangled, without
 * the argument list, e.g. "crs::fwd_fp32c<4, 8>") and its number of work-groups (for a
 * call whose planes need several block shapes, the first of its launches).
 * CRS_ERR_INVALID before the first such call. */
crs_status crs_last_launch(crs_context *ctx, int adjoint, const char **kernel,
                           int *workgroups);

#ifdef __cplusplus
}
#endif

#endif /* CELERIS_RS_H */
