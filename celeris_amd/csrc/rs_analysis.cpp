// rs_analysis.cpp — the reference's focal-plane analyses on the MI355X propagation,
// each as ONE batched launch (crs_forward_host) instead of a GPU grid plus CPU loops.
//
//   crs_analyze_focus          analysis/focal.cpp:32-146   (analyze_focus)
//   crs_analyze_through_focus  analysis/throughfocus.cpp:15-132
//   crs_analyze_chromatic      analysis/chromatic.cpp:9-58
//   crs_analyze_chromatic_dispersive  analysis/chromatic.cpp:60-127
//   crs_analyze_field_grid     analysis/field.cpp:150-233
//   crs_analyze_field_of_view  analysis/field.cpp:12-76
//   crs_analyze_wide_fov       analysis/field.cpp:78-148
//   crs_axis_scan, crs_verify_achromatic_focus, crs_polar_focus_check: the on-axis
//                              focus scans of design/achromatic.cpp, design/pb_achromatic.cpp,
//                              cli/cmd_polar.cpp, cli/cli_common.cpp, gui/workers.cpp
//
// Inputs are the aperture pillar list the reference builds from a MetalensDesign and
// a UnitCellLibrary (focal.cpp:42-57): positions, complex transmissions and, for the
// focus metrics, each cell's ideal focusing phase.  The scalar post-processing
// (peaks, FWHM, encircled energy, normalisation, argmax) follows the reference line
// by line; only the sums move to the GPU.  In the reference the on-axis scans of
// through-focus and chromatic, and analyze_focus's ideal peak, were FP64 CPU loops.
#include "../../include/celeris_rs.h"
#include "rs_host.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

crs_plane psf_plane(double cx, double cy, double z, double k, int n, double W, int tset,
                    int64_t offset) {
    crs_plane p{};
    p.k = k;
    p.x0 = cx - W;
    p.sx = 2.0 * W / (n - 1);
    p.y0 = cy - W;
    p.sy = p.sx;
    p.z0 = z;
    p.sz = 0.0;
    p.nx = n;
    p.ny = n;
    p.tset = tset;
    p.offset = offset;
    return p;
}

crs_plane axis_plane(double zlo, double zhi, int nz, double k, int tset, int64_t offset) {
    crs_plane p{};
    p.k = k;
    p.z0 = zlo;
    p.sz = nz > 1 ? (zhi - zlo) / (nz - 1) : 0.0;
    p.nx = 1;
    p.ny = nz;
    p.tset = tset;
    p.offset = offset;
    return p;
}

// FWHM along a line of samples: (hi - lo) * step over the samples >= half, or 0
// (focal.cpp:117-124).
double fwhm_line(const double* v, int n, int stride, double half, double step) {
    int lo = -1, hi = -1;
    for (int s = 0; s < n; ++s) {
        if (v[static_cast<int64_t>(s) * stride] >= half) {
            if (lo < 0) lo = s;
            hi = s;
        }
    }
    return (hi > lo && lo >= 0) ? (hi - lo) * step : 0.0;
}

}  // namespace

extern "C" crs_status crs_analyze_focus(const double* px, const double* py, const double* t,
                                        const double* phi_ideal, int npil, double focal_um,
                                        double wavelength_um, double diameter_um, int mode,
                                        double* out5, double* intensity) {
    crs::clear_error();
    if (!px || !py || !t || !phi_ideal || !out5 || npil <= 0) {
        crs::set_error("crs_analyze_focus: null array or npil <= 0");
        return CRS_ERR_INVALID;
    }
    const double f = focal_um;
    const double k = 2.0 * kPi / wavelength_um;
    const double dl = wavelength_um * focal_um / diameter_um;
    const double W = std::max(6.0 * dl, 4.0);  // focal.cpp:61-64
    const int NG = 121;
    const double step = 2.0 * W / (NG - 1);
    const int64_t ng2 = static_cast<int64_t>(NG) * NG;

    // Set 0: the design; set 1: the ideal lens, polar(1, phi_ideal) (focal.cpp:72).
    std::vector<double> tt(4 * static_cast<size_t>(npil));
    for (int q = 0; q < npil; ++q) {
        tt[2 * q] = t[2 * q];
        tt[2 * q + 1] = t[2 * q + 1];
        tt[2 * (npil + q)] = std::cos(phi_ideal[q]);
        tt[2 * (npil + q) + 1] = std::sin(phi_ideal[q]);
    }
    crs_plane planes[2] = {psf_plane(0.0, 0.0, f, k, NG, W, 0, 0),
                           axis_plane(f, f, 1, k, 1, ng2)};
    std::vector<double> field(2 * (ng2 + 1)), I(ng2 + 1);
    crs_status st = crs_forward_host(mode, px, py, tt.data(), npil, 2, planes, 2, field.data(),
                                     I.data());
    if (st != CRS_OK) return st;

    const double peak = *std::max_element(I.begin(), I.begin() + ng2);
    const double peak_ideal = I[ng2];
    double amp_inphase = 0.0;  // focal.cpp:109-115
    for (int q = 0; q < npil; ++q) {
        const double Rr = std::sqrt(px[q] * px[q] + py[q] * py[q] + f * f);
        amp_inphase += std::abs(std::complex<double>(t[2 * q], t[2 * q + 1])) / Rr;
    }
    const double peak_same_amp = amp_inphase * amp_inphase;
    const int center_row = NG / 2;
    const double fwhm = fwhm_line(I.data() + static_cast<int64_t>(center_row) * NG, NG, 1,
                                  peak / 2.0, step);
    const double r_null = 1.22 * dl;  // focal.cpp:126-137
    double in = 0.0, total = 0.0;
    for (int j = 0; j < NG; ++j) {
        const double fy = -W + j * step;
        for (int i = 0; i < NG; ++i) {
            const double fx = -W + i * step;
            const double v = I[static_cast<int64_t>(j) * NG + i];
            total += v;
            if (std::sqrt(fx * fx + fy * fy) <= r_null) in += v;
        }
    }
    out5[0] = peak_ideal > 0 ? peak / peak_ideal : 0.0;
    out5[1] = peak_same_amp > 0 ? peak / peak_same_amp : 0.0;
    out5[2] = fwhm;
    out5[3] = dl;
    out5[4] = total > 0 ? in / total : 0.0;
    if (intensity) std::copy(I.begin(), I.begin() + ng2, intensity);
    return CRS_OK;
}

extern "C" crs_status crs_analyze_through_focus(const double* px, const double* py,
                                                const double* t, int npil, double focal_um,
                                                double wavelength_um, double diameter_um,
                                                int caustic_n, int mode, double* z161,
                                                double* I161, double* scalars6,
                                                double* caustic) {
    crs::clear_error();
    if (!px || !py || !t || !z161 || !I161 || !scalars6 || npil <= 0 || caustic_n < 2) {
        crs::set_error("crs_analyze_through_focus: null array, npil <= 0 or caustic_n < 2");
        return CRS_ERR_INVALID;
    }
    const double f = focal_um;
    const double k = 2.0 * kPi / wavelength_um;
    const double R = diameter_um / 2.0;
    // throughfocus.cpp:78-82
    const double na = std::sin(std::atan(R / f));
    const double half = std::max(2.0 * wavelength_um / (na * na + 1e-6), 0.25 * f);
    const double zlo = std::max(0.1 * f, f - half);
    const double zhi = f + half;
    const int N = 161;
    // throughfocus.cpp:113-116
    const double dl = wavelength_um * f / std::max(diameter_um, 1e-6);
    const double Wx = std::max(8.0 * dl, 3.0);
    const int cn = caustic_n;

    crs_plane planes[2] = {axis_plane(zlo, zhi, N, k, 0, 0), crs_plane{}};
    planes[1].k = k;
    planes[1].x0 = -Wx;
    planes[1].sx = (2.0 * Wx) / (cn - 1);
    planes[1].z0 = zlo;
    planes[1].sz = (zhi - zlo) / (cn - 1);
    planes[1].nx = cn;
    planes[1].ny = cn;
    planes[1].offset = N;
    const int64_t total = N + static_cast<int64_t>(cn) * cn;
    std::vector<double> I(total);
    crs_status st = crs_forward_host(mode, px, py, t, npil, 1, planes, 2, nullptr, I.data());
    if (st != CRS_OK) return st;

    double peak = 0.0;
    for (int i = 0; i < N; ++i) {
        z161[i] = static_cast<float>(zlo + (zhi - zlo) * i / (N - 1));  // stored as float
        peak = std::max(peak, I[i]);
    }
    if (peak <= 0) peak = 1;
    int ipk = 0;
    for (int i = 0; i < N; ++i) {
        I161[i] = static_cast<float>(I[i] / peak);
        if (I[i] > I[ipk]) ipk = i;
    }
    int lo = -1, hi = -1;
    for (int i = 0; i < N; ++i)
        if (I[i] >= peak / 2.0) { if (lo < 0) lo = i; hi = i; }
    scalars6[0] = z161[ipk];
    scalars6[1] = (hi > lo && lo >= 0) ? (z161[hi] - z161[lo]) : 0.0;
    scalars6[2] = -Wx;
    scalars6[3] = Wx;
    scalars6[4] = zlo;
    scalars6[5] = zhi;
    if (caustic) {
        double cpeak = 0.0;
        for (int64_t i = N; i < total; ++i) cpeak = std::max(cpeak, I[i]);
        if (cpeak <= 0) cpeak = 1.0;
        for (int64_t i = N; i < total; ++i) caustic[i - N] = static_cast<float>(I[i] / cpeak);
    }
    return CRS_OK;
}

namespace {

// The chromatic scan of chromatic.cpp:9-58 and :60-127: the design-wavelength peak at
// the design focus (:40, :100-102) and, for each of n wavelengths, NZ = 120 on-axis
// samples over [0.55, 1.6] f0 lambda0 / lambda (:44-56, :106-124), all as planes of one
// launch.  set_of(i) is the transmission set of wavelength i (-1: the design peak).
template <class SetOf>
crs_status chromatic_scan(const double* px, const double* py, const double* t, int npil, int nsets,
                          SetOf set_of, double design_focal, double design_wavelength,
                          double lambda_min, double lambda_max, int n, int mode, double* out3n) {
    const int NZ = 120;
    std::vector<crs_plane> planes;
    planes.reserve(n + 1);
    planes.push_back(axis_plane(design_focal, design_focal, 1, 2.0 * kPi / design_wavelength,
                                set_of(-1), 0));
    std::vector<double> lam(n), zlo(n), zhi(n);
    for (int i = 0; i < n; ++i) {
        lam[i] = lambda_min + (lambda_max - lambda_min) * i / (n - 1);
        const double expected = design_focal * design_wavelength / lam[i];
        zlo[i] = 0.55 * expected;
        zhi[i] = 1.6 * expected;
        planes.push_back(axis_plane(zlo[i], zhi[i], NZ, 2.0 * kPi / lam[i], set_of(i),
                                    1 + static_cast<int64_t>(i) * NZ));
    }
    std::vector<double> I(1 + static_cast<size_t>(n) * NZ);
    crs_status st = crs_forward_host(mode, px, py, t, npil, nsets, planes.data(),
                                     static_cast<int>(planes.size()), nullptr, I.data());
    if (st != CRS_OK) return st;
    const double design_peak = I[0];
    for (int i = 0; i < n; ++i) {
        const double expected = design_focal * design_wavelength / lam[i];
        double best_z = expected, best_I = -1.0;
        for (int j = 0; j < NZ; ++j) {
            const double z = zlo[i] + (zhi[i] - zlo[i]) * j / (NZ - 1);
            const double v = I[1 + static_cast<size_t>(i) * NZ + j];
            if (v > best_I) { best_I = v; best_z = z; }
        }
        out3n[3 * i] = lam[i];
        out3n[3 * i + 1] = best_z;
        out3n[3 * i + 2] = design_peak > 0 ? best_I / design_peak : 0.0;
    }
    return CRS_OK;
}

}  // namespace

extern "C" crs_status crs_analyze_chromatic(const double* px, const double* py, const double* t,
                                            int npil, double design_focal, double design_wavelength,
                                            double lambda_min, double lambda_max, int n, int mode,
                                            double* out3n) {
    crs::clear_error();
    if (!px || !py || !t || !out3n || npil <= 0 || n < 2) {
        crs::set_error("crs_analyze_chromatic: null array, npil <= 0 or n < 2");
        return CRS_ERR_INVALID;
    }
    return chromatic_scan(px, py, t, npil, 1, [](int) { return 0; }, design_focal,
                          design_wavelength, lambda_min, lambda_max, n, mode, out3n);
}

extern "C" crs_status crs_analyze_chromatic_dispersive(
    const double* px, const double* py, const double* t_design, const double* t_bands, int npil,
    double design_focal, double design_wavelength, double lambda_min, double lambda_max, int n,
    int mode, double* out3n) {
    crs::clear_error();
    if (!px || !py || !t_design || !t_bands || !out3n || npil <= 0 || n < 2) {
        crs::set_error("crs_analyze_chromatic_dispersive: null array, npil <= 0 or n < 2");
        return CRS_ERR_INVALID;
    }
    if (static_cast<int64_t>(npil) * (n + 1) > (int64_t(1) << 31) - 1) {
        crs::set_error("crs_analyze_chromatic_dispersive: (n + 1) * npil exceeds 2^31 - 1");
        return CRS_ERR_INVALID;
    }
    // Set 0: the design wavelength's transmissions; set 1 + i: wavelength i's.
    const size_t per = 2 * static_cast<size_t>(npil);
    std::vector<double> t(per * (n + 1));
    std::copy(t_design, t_design + per, t.begin());
    std::copy(t_bands, t_bands + per * n, t.begin() + per);
    return chromatic_scan(px, py, t.data(), npil, n + 1, [](int i) { return i + 1; },
                          design_focal, design_wavelength, lambda_min, lambda_max, n, mode, out3n);
}

extern "C" crs_status crs_analyze_field_grid(const double* px, const double* py, const double* t,
                                             int npil, double focal_um, double wavelength_um,
                                             double diameter_um, double max_angle_deg, int n_half,
                                             int psf_n, int mode, double* out7) {
    crs::clear_error();
    if (!px || !py || !t || !out7 || npil <= 0 || n_half < 0 || psf_n < 2) {
        crs::set_error("crs_analyze_field_grid: null array, npil <= 0, n_half < 0 or psf_n < 2");
        return CRS_ERR_INVALID;
    }
    const double k = 2.0 * kPi / wavelength_um;
    const double dl = wavelength_um * focal_um / diameter_um;  // field.cpp:176-178
    const double W = std::max(6.0 * dl, 4.0);
    const double step = 2.0 * W / (psf_n - 1);
    const int g = 2 * n_half + 1;
    const int npts = 1 + g * g;  // the on-axis reference spot, then the grid
    const int64_t npix = static_cast<int64_t>(psf_n) * psf_n;

    // One transmission set per field point: t * polar(1, k (sx x + sy y)) (field.cpp:186-187).
    std::vector<double> tt(2 * static_cast<size_t>(npil) * npts);
    std::vector<crs_plane> planes(npts);
    std::vector<double> ang_x(npts), ang_y(npts), cxs(npts), cys(npts);
    for (int p = 0; p < npts; ++p) {
        double ax = 0.0, ay = 0.0;
        if (p > 0) {
            const int jy = (p - 1) / g - n_half, jx = (p - 1) % g - n_half;
            ay = n_half > 0 ? max_angle_deg * jy / n_half : 0.0;
            ax = n_half > 0 ? max_angle_deg * jx / n_half : 0.0;
        }
        const double tx = ax * kPi / 180.0, ty = ay * kPi / 180.0;
        const double cx = focal_um * std::tan(tx), cy = focal_um * std::tan(ty);
        const double sx = std::sin(tx), sy = std::sin(ty);
        double* ts = tt.data() + 2 * static_cast<size_t>(npil) * p;
        for (int q = 0; q < npil; ++q) {
            const std::complex<double> v = std::complex<double>(t[2 * q], t[2 * q + 1]) *
                                           std::polar(1.0, k * (sx * px[q] + sy * py[q]));
            ts[2 * q] = v.real();
            ts[2 * q + 1] = v.imag();
        }
        planes[p] = psf_plane(cx, cy, focal_um, k, psf_n, W, p, npix * p);
        ang_x[p] = ax;
        ang_y[p] = ay;
        cxs[p] = cx;
        cys[p] = cy;
    }
    std::vector<double> I(npix * npts);
    crs_status st = crs_forward_host(mode, px, py, tt.data(), npil, npts, planes.data(), npts,
                                     nullptr, I.data());
    if (st != CRS_OK) return st;

    // Peak pixel and tangential / sagittal FWHM of each spot (field.cpp:191-208).
    std::vector<double> peak(npts), fwx(npts), fwy(npts);
    for (int p = 0; p < npts; ++p) {
        const double* m = I.data() + npix * p;
        int pix = 0, piy = 0;
        double best = -1.0;
        for (int j = 0; j < psf_n; ++j)
            for (int i = 0; i < psf_n; ++i) {
                const double v = m[static_cast<int64_t>(j) * psf_n + i];
                if (v > best) { best = v; pix = i; piy = j; }
            }
        peak[p] = best;
        fwx[p] = fwhm_line(m + static_cast<int64_t>(piy) * psf_n, psf_n, 1, best / 2.0, step);
        fwy[p] = fwhm_line(m + pix, psf_n, psf_n, best / 2.0, step);
    }
    const double ref = peak[0] > 0 ? peak[0] : 1.0;  // field.cpp:210-211
    for (int p = 1; p < npts; ++p) {
        double* o = out7 + 7 * static_cast<size_t>(p - 1);
        o[0] = ang_x[p];
        o[1] = ang_y[p];
        o[2] = peak[p] / ref;
        o[3] = fwx[p];
        o[4] = fwy[p];
        o[5] = cxs[p];
        o[6] = cys[p];
    }
    return CRS_OK;
}

namespace {

// The tilted-illumination spot search shared by analyze_field_of_view and
// analyze_wide_fov (field.cpp:34-62, :105-135): for each field angle an 81 x 81 window
// at z = f, centred at (xc, 0) with half-width W, of the cells under a stop patch
// (radius^2 r2 around (x_patch, 0); r2 < 0: every cell) lit by exp(i k sin(theta) x);
// returns each window's peak |E|^2 and the x of the first pixel reaching it (rows outer,
// columns inner, strict >).  All windows run as one launch, one transmission set each.
struct TiltSpot {
    double theta, xc, W, x_patch;
};

crs_status tilt_spots(const double* px, const double* py, const double* t, int npil,
                      double focal_um, double k, double r2, const std::vector<TiltSpot>& spots,
                      int mode, std::vector<double>& peak, std::vector<double>& peak_x) {
    const int NG = 81;
    const int nsp = static_cast<int>(spots.size());
    const int64_t npix = static_cast<int64_t>(NG) * NG;
    std::vector<double> tt(2 * static_cast<size_t>(npil) * nsp);
    std::vector<crs_plane> planes(nsp);
    for (int p = 0; p < nsp; ++p) {
        const TiltSpot& sp = spots[p];
        const double sn = std::sin(sp.theta);
        double* ts = tt.data() + 2 * static_cast<size_t>(npil) * p;
        for (int q = 0; q < npil; ++q) {
            const double dx = px[q] - sp.x_patch, dy = py[q];
            if (r2 >= 0.0 && dx * dx + dy * dy > r2) {  // outside the stop patch
                ts[2 * q] = ts[2 * q + 1] = 0.0;
                continue;
            }
            const std::complex<double> v =
                std::complex<double>(t[2 * q], t[2 * q + 1]) * std::polar(1.0, k * sn * px[q]);
            ts[2 * q] = v.real();
            ts[2 * q + 1] = v.imag();
        }
        crs_plane& pl = planes[p];
        pl = crs_plane{};
        pl.k = k;
        pl.x0 = sp.xc - sp.W;
        pl.sx = 2.0 * sp.W / (NG - 1);
        pl.y0 = -sp.W;
        pl.sy = pl.sx;
        pl.z0 = focal_um;
        pl.nx = NG;
        pl.ny = NG;
        pl.tset = p;
        pl.offset = npix * p;
    }
    std::vector<double> I(npix * nsp);
    const crs_status st =
        crs_forward_host(mode, px, py, tt.data(), npil, nsp, planes.data(), nsp, nullptr, I.data());
    if (st != CRS_OK) return st;
    peak.assign(nsp, 0.0);
    peak_x.assign(nsp, 0.0);
    for (int p = 0; p < nsp; ++p) {
        const TiltSpot& sp = spots[p];
        const double* m = I.data() + npix * p;
        double best = 0.0, best_x = sp.xc;
        for (int j = 0; j < NG; ++j)
            for (int i = 0; i < NG; ++i) {
                const double v = m[static_cast<int64_t>(j) * NG + i];
                if (v > best) {
                    best = v;
                    best_x = sp.xc - sp.W + 2.0 * sp.W * i / (NG - 1);
                }
            }
        peak[p] = best;
        peak_x[p] = best_x;
    }
    return CRS_OK;
}

}  // namespace

extern "C" crs_status crs_analyze_field_of_view(const double* px, const double* py,
                                                const double* t, int npil, double focal_um,
                                                double wavelength_um, double diameter_um,
                                                const double* angles_deg, int nangles, int mode,
                                                double* out3) {
    crs::clear_error();
    if (!px || !py || !t || npil <= 0 || nangles < 0 || (nangles > 0 && (!angles_deg || !out3))) {
        crs::set_error("crs_analyze_field_of_view: null array, npil <= 0 or nangles < 0");
        return CRS_ERR_INVALID;
    }
    const double k = 2.0 * kPi / wavelength_um;
    const double dl = wavelength_um * focal_um / diameter_um;  // field.cpp:40
    auto spot = [&](double theta, double xc) {
        const double x_sin = focal_um * std::sin(theta), x_tan = focal_um * std::tan(theta);
        return TiltSpot{theta, xc, std::max(3.0 * dl, 2.0) + std::abs(x_tan - x_sin), 0.0};
    };
    std::vector<TiltSpot> spots{spot(0.0, 0.0)};  // the on-axis reference (field.cpp:64)
    for (int a = 0; a < nangles; ++a) {
        const double theta = angles_deg[a] * kPi / 180.0;
        spots.push_back(spot(theta, focal_um * std::tan(theta)));
    }
    std::vector<double> pk, pkx;
    const crs_status st = tilt_spots(px, py, t, npil, focal_um, k, -1.0, spots, mode, pk, pkx);
    if (st != CRS_OK) return st;
    for (int a = 0; a < nangles; ++a) {
        out3[3 * a] = angles_deg[a];
        out3[3 * a + 1] = pk[0] > 0 ? pk[a + 1] / pk[0] : 0.0;
        out3[3 * a + 2] = pkx[a + 1];
    }
    return CRS_OK;
}

extern "C" crs_status crs_analyze_wide_fov(const double* px, const double* py, const double* t,
                                           int npil, double focal_um, double wavelength_um,
                                           double lens_diameter_um, double stop_diameter_um,
                                           double stop_distance_um, const double* angles_deg,
                                           int nangles, int mode, double* out3) {
    (void)lens_diameter_um;  // the caller's pillar list is already cut to the lens aperture
    crs::clear_error();
    if (!px || !py || !t || npil <= 0 || nangles < 0 || (nangles > 0 && (!angles_deg || !out3))) {
        crs::set_error("crs_analyze_wide_fov: null array, npil <= 0 or nangles < 0");
        return CRS_ERR_INVALID;
    }
    const double k = 2.0 * kPi / wavelength_um;
    const double dl_stop = wavelength_um * focal_um / stop_diameter_um;  // field.cpp:91
    const double R_stop = stop_diameter_um / 2.0;
    auto spot = [&](double theta, double x_patch) {
        const double x_sin = focal_um * std::sin(theta), x_tan = focal_um * std::tan(theta);
        const double xc = 0.5 * (x_sin + x_tan);
        return TiltSpot{theta, xc, std::max(3.0 * dl_stop, 2.0) + 0.5 * std::abs(x_tan - x_sin),
                        x_patch};
    };
    std::vector<TiltSpot> spots{spot(0.0, 0.0)};  // field.cpp:137
    for (int a = 0; a < nangles; ++a) {
        const double theta = angles_deg[a] * kPi / 180.0;
        spots.push_back(spot(theta, stop_distance_um * std::tan(theta)));
    }
    std::vector<double> pk, pkx;
    const crs_status st =
        tilt_spots(px, py, t, npil, focal_um, k, R_stop * R_stop, spots, mode, pk, pkx);
    if (st != CRS_OK) return st;
    for (int a = 0; a < nangles; ++a) {
        out3[3 * a] = angles_deg[a];
        out3[3 * a + 1] = pk[0] > 0 ? pk[a + 1] / pk[0] : 0.0;
        out3[3 * a + 2] = pkx[a + 1];
    }
    return CRS_OK;
}

// ---- on-axis focus scans ---------------------------------------------------------------
//   crs_axis_scan                the scan loop of the callers below, batched
//   crs_verify_achromatic_focus  design/achromatic.cpp:288-338, design/pb_achromatic.cpp:255-301
//   crs_polar_focus_check        cli/cmd_polar.cpp:85-126, gui/workers.cpp:181-197

namespace {

// Every scan as a one-column plane of one launch; I receives the samples scan after scan.
crs_status run_axis_scans(const double* px, const double* py, const double* t, int npil, int nsets,
                          const crs_axis_scan_desc* scans, int nscans, int mode,
                          std::vector<double>& I) {
    std::vector<crs_plane> planes(nscans);
    int64_t total = 0;
    for (int s = 0; s < nscans; ++s) {
        const crs_axis_scan_desc& d = scans[s];
        if (d.nz < 1 || d.tset < 0 || d.tset >= nsets || !(d.wavelength_um > 0.0)) {
            crs::set_error("axis scan %d: nz < 1, tset outside [0, %d) or wavelength <= 0", s,
                           nsets);
            return CRS_ERR_INVALID;
        }
        planes[s] = axis_plane(d.zlo, d.zlo + d.zspan, d.nz, 2.0 * kPi / d.wavelength_um, d.tset,
                               total);
        if (d.nz > 1) planes[s].sz = d.zspan / (d.nz - 1);
        total += d.nz;
    }
    if (total > (int64_t(1) << 31)) {
        crs::set_error("axis scans: more than 2^31 samples");
        return CRS_ERR_INVALID;
    }
    I.assign(static_cast<size_t>(total), 0.0);
    return crs_forward_host(mode, px, py, t, npil, nsets, planes.data(), nscans, nullptr, I.data());
}

// The callers' argmax: the first sample beating every earlier one and -1.
void scan_argmax(const crs_axis_scan_desc& d, const double* I, double* best_z, double* best_I,
                 int* best_j) {
    double bz = d.z_default, bi = -1.0;
    int bj = -1;
    for (int j = 0; j < d.nz; ++j) {
        const double z = d.nz > 1 ? d.zlo + d.zspan * j / (d.nz - 1) : d.zlo;
        if (I[j] > bi) { bi = I[j]; bz = z; bj = j; }
    }
    *best_z = bz;
    *best_I = bi;
    if (best_j) *best_j = bj;
}

}  // namespace

extern "C" crs_status crs_axis_scan(const double* px, const double* py, const double* t, int npil,
                                    int nsets, const crs_axis_scan_desc* scans, int nscans,
                                    int mode, double* best_z, double* best_I, double* curves) {
    crs::clear_error();
    if (!px || !py || !t || !scans || !best_z || !best_I || npil <= 0 || nsets <= 0 || nscans <= 0) {
        crs::set_error("crs_axis_scan: null array or a count <= 0");
        return CRS_ERR_INVALID;
    }
    std::vector<double> I;
    crs_status st = run_axis_scans(px, py, t, npil, nsets, scans, nscans, mode, I);
    if (st != CRS_OK) return st;
    int64_t off = 0;
    for (int s = 0; s < nscans; ++s) {
        scan_argmax(scans[s], I.data() + off, best_z + s, best_I + s, nullptr);
        off += scans[s].nz;
    }
    if (curves) std::copy(I.begin(), I.end(), curves);
    return CRS_OK;
}

extern "C" crs_status crs_verify_achromatic_focus(const double* px, const double* py,
                                                  const double* t_bands, int npil, int nb,
                                                  const double* wavelengths_um, int center_index,
                                                  double center_wavelength_um, double focal_um,
                                                  int mode, double* out3nb) {
    crs::clear_error();
    if (!px || !py || !t_bands || !wavelengths_um || !out3nb || npil <= 0 || nb <= 0 ||
        center_index < 0 || center_index >= nb) {
        crs::set_error("crs_verify_achromatic_focus: null array, a count <= 0 or center_index "
                       "outside [0, nb)");
        return CRS_ERR_INVALID;
    }
    const int NZ = 200;  // achromatic.cpp:322, pb_achromatic.cpp:289
    const double zlo = 0.4 * focal_um, zhi = 1.8 * focal_um;
    // Scan 0: the design peak, I(f) of the centre band at the centre wavelength
    // (achromatic.cpp:311-313); then one 200-sample scan per band.
    std::vector<crs_axis_scan_desc> scans(1 + static_cast<size_t>(nb));
    scans[0] = {center_index, 1, center_wavelength_um, focal_um, 0.0, focal_um};
    for (int j = 0; j < nb; ++j)
        scans[1 + j] = {j, NZ, wavelengths_um[j], zlo, zhi - zlo, focal_um};
    std::vector<double> I;
    crs_status st = run_axis_scans(px, py, t_bands, npil, nb, scans.data(), 1 + nb, mode, I);
    if (st != CRS_OK) return st;
    const double design_peak = I[0];
    for (int j = 0; j < nb; ++j) {
        double bz, bi;
        scan_argmax(scans[1 + j], I.data() + 1 + static_cast<size_t>(j) * NZ, &bz, &bi, nullptr);
        out3nb[3 * j] = wavelengths_um[j];
        out3nb[3 * j + 1] = bz;
        out3nb[3 * j + 2] = design_peak > 0 ? bi / design_peak : 0.0;
    }
    return CRS_OK;
}

extern "C" crs_status crs_polar_focus_check(const double* px, const double* py, const double* tx,
                                            const double* ty, int npil, double wavelength_um,
                                            double focal_x_um, double focal_y_um, int nz, int mode,
                                            double* out4) {
    crs::clear_error();
    if (!px || !py || !tx || !ty || !out4 || npil <= 0 || nz < 1) {
        crs::set_error("crs_polar_focus_check: null array, npil <= 0 or nz < 1");
        return CRS_ERR_INVALID;
    }
    const double zlo = 0.5 * std::min(focal_x_um, focal_y_um);
    const double zhi = 1.5 * std::max(focal_x_um, focal_y_um);
    std::vector<double> t(4 * static_cast<size_t>(npil));
    std::copy(tx, tx + 2 * static_cast<size_t>(npil), t.begin());
    std::copy(ty, ty + 2 * static_cast<size_t>(npil), t.begin() + 2 * static_cast<size_t>(npil));
    const crs_axis_scan_desc scans[2] = {{0, nz, wavelength_um, zlo, zhi - zlo, zlo},
                                         {1, nz, wavelength_um, zlo, zhi - zlo, zlo}};
    std::vector<double> I;
    crs_status st = run_axis_scans(px, py, t.data(), npil, 2, scans, 2, mode, I);
    if (st != CRS_OK) return st;
    const double* Ix = I.data();
    const double* Iy = I.data() + nz;
    double bi;
    int jx, jy;
    scan_argmax(scans[0], Ix, &out4[0], &bi, &jx);
    scan_argmax(scans[1], Iy, &out4[1], &bi, &jy);
    out4[2] = out4[3] = 0.0;
    if (std::abs(focal_x_um - focal_y_um) > 1e-6 && jx >= 0 && jy >= 0) {
        // Both channels are sampled on the same z grid, so the other channel's intensity
        // at a channel's focus is its sample there (cmd_polar.cpp:113-123 recompute it).
        out4[2] = 10.0 * std::log10(Ix[jx] / std::max(Iy[jx], 1e-30));
        out4[3] = 10.0 * std::log10(Iy[jy] / std::max(Ix[jy], 1e-30));
    }
    return CRS_OK;
}
