// rs_tolerance.cpp — the fabrication-tolerance Monte Carlo of analysis/tolerance.cpp:12-48
// (analyze_tolerance) with every trial's focal grid in one batched launch.
//
// The reference runs each trial as its own analyze_focus (focal.cpp:32-146) on a host
// thread: jitter every fill of the map with N(0, sigma_fill) from a std::mt19937 seeded
// per trial, clamp to [0.05, 0.95], rebuild the aperture transmissions and take the
// Strehl ratio peak / peak_ideal of a 121 x 121 focal grid.  Here the draws are the same
// (the same standard-library engine and distribution, seeded the same way, in the same
// order), each trial is a transmission set, and up to 256 trials share one launch
// with the ideal lens's peak.

#include "../../include/celeris_rs.h"
#include "rs_host.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <random>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kNG = 121;  // focal.cpp:63

// Nearest sampled fill, the first one on ties (metalens.cpp:112-120).
int nearest_fill(const double* lib_fill, int nlib, double f) {
    int best = 0;
    double best_d = std::abs(lib_fill[0] - f);
    for (int i = 1; i < nlib; ++i) {
        const double d = std::abs(lib_fill[i] - f);
        if (d < best_d) { best_d = d; best = i; }
    }
    return best;
}

}  // namespace

extern "C" crs_status crs_analyze_tolerance(const double* fill_map, int n_cells, double period_um,
                                            const double* lib_fill, const double* lib_phase,
                                            const double* lib_amp, int nlib, double focal_um,
                                            double wavelength_um, double diameter_um,
                                            const double* sigmas_nm, int nsigmas, int trials,
                                            unsigned seed, int mode, double* out4,
                                            double* strehl) {
    crs::clear_error();
    if (!fill_map || !lib_fill || !lib_phase || !lib_amp || !sigmas_nm || !out4 || n_cells <= 0 ||
        nlib <= 0 || nsigmas <= 0 || !(period_um > 0.0) || !(wavelength_um > 0.0)) {
        crs::set_error("crs_analyze_tolerance: null array, a count <= 0, or period or "
                       "wavelength not positive");
        return CRS_ERR_INVALID;
    }
    for (int s = 0; s < nsigmas; ++s)
        if (!(sigmas_nm[s] >= 0.0)) {
            crs::set_error("crs_analyze_tolerance: sigma %d is negative or NaN", s);
            return CRS_ERR_INVALID;
        }

    // Aperture cells (focal.cpp:42-57): site, position, ideal focusing phase.
    const double p = period_um;
    const double center = (n_cells - 1) / 2.0;
    const double k = 2.0 * kPi / wavelength_um;
    const double R_ap = diameter_um / 2.0;
    const double f = focal_um;
    std::vector<int64_t> site;
    std::vector<double> px, py, phi;
    for (int iy = 0; iy < n_cells; ++iy)
        for (int ix = 0; ix < n_cells; ++ix) {
            const double x = (ix - center) * p;
            const double y = (iy - center) * p;
            const double r = std::sqrt(x * x + y * y);
            if (r > R_ap) continue;
            site.push_back(static_cast<int64_t>(iy) * n_cells + ix);
            px.push_back(x);
            py.push_back(y);
            phi.push_back(-k * (std::sqrt(r * r + f * f) - f));
        }
    const int ncell = static_cast<int>(site.size());
    if (ncell == 0) {
        crs::set_error("crs_analyze_tolerance: no cell inside the aperture");
        return CRS_ERR_INVALID;
    }

    const double dl = wavelength_um * focal_um / diameter_um;
    const double W = std::max(6.0 * dl, 4.0);  // focal.cpp:61-64
    const double step = 2.0 * W / (kNG - 1);
    const int64_t ng2 = static_cast<int64_t>(kNG) * kNG;
    const int n = std::max(1, trials);
    const int64_t total = static_cast<int64_t>(nsigmas) * n;
    // Trials per launch: enough to fill the GPU, few enough that the host copy of the
    // transmissions stays under ~256 MB.
    const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(256, (int64_t(16) << 20) / ncell));
    const size_t nfill = static_cast<size_t>(n_cells) * n_cells;

    std::vector<double> st(total);
    std::vector<double> fills(nfill);
    double peak_ideal = 0.0;
    for (int64_t b0 = 0; b0 < total; b0 += batch) {
        const int64_t nb = std::min(batch, total - b0);
        const bool first = b0 == 0;
        // Set 0: the ideal lens polar(1, phi_ideal) (focal.cpp:72), set 1 + i: trial b0 + i.
        std::vector<double> t(2 * static_cast<size_t>(ncell) * (1 + nb));
        for (int q = 0; q < ncell; ++q) {
            t[2 * q] = std::cos(phi[q]);
            t[2 * q + 1] = std::sin(phi[q]);
        }
        for (int64_t i = 0; i < nb; ++i) {
            const int64_t idx = b0 + i;
            const double sigma_nm = sigmas_nm[idx / n];
            const int trial = static_cast<int>(idx % n);
            // tolerance.cpp:20-31: CD error (nm) to fill error, one engine per trial.
            const double sigma_fill = (sigma_nm / 1000.0) / period_um;
            std::mt19937 rng(seed + static_cast<unsigned>(trial) * 7919u +
                             static_cast<unsigned>(sigma_nm * 13.0));
            std::normal_distribution<double> g(0.0, sigma_fill);
            for (size_t c = 0; c < nfill; ++c)
                fills[c] = std::clamp(fill_map[c] + g(rng), 0.05, 0.95);
            double* ts = t.data() + 2 * static_cast<size_t>(ncell) * (1 + i);
            for (int q = 0; q < ncell; ++q) {
                const int e = nearest_fill(lib_fill, nlib, fills[site[q]]);
                const std::complex<double> v = std::polar(lib_amp[e], lib_phase[e]);
                ts[2 * q] = v.real();
                ts[2 * q + 1] = v.imag();
            }
        }
        std::vector<crs_plane> planes(nb + (first ? 1 : 0));
        for (int64_t i = 0; i < nb; ++i) {
            crs_plane& P = planes[i];
            P = crs_plane{};
            P.k = k;
            P.x0 = -W;
            P.sx = step;
            P.y0 = -W;
            P.sy = step;
            P.z0 = f;
            P.nx = kNG;
            P.ny = kNG;
            P.tset = static_cast<int32_t>(1 + i);
            P.offset = i * ng2;
        }
        if (first) {  // the ideal peak at the focal centre, once (the same in every trial)
            crs_plane& P = planes[nb];
            P = crs_plane{};
            P.k = k;
            P.z0 = f;
            P.nx = 1;
            P.ny = 1;
            P.tset = 0;
            P.offset = nb * ng2;
        }
        std::vector<double> I(nb * ng2 + (first ? 1 : 0));
        crs_status s = crs_forward_host(mode, px.data(), py.data(), t.data(), ncell,
                                        static_cast<int>(1 + nb), planes.data(),
                                        static_cast<int>(planes.size()), nullptr, I.data());
        if (s != CRS_OK) return s;
        if (first) peak_ideal = I[nb * ng2];
        for (int64_t i = 0; i < nb; ++i) {
            const double peak = *std::max_element(I.begin() + i * ng2, I.begin() + (i + 1) * ng2);
            st[b0 + i] = peak_ideal > 0 ? peak / peak_ideal : 0.0;  // focal.cpp:140
        }
    }

    // tolerance.cpp:37-46: mean, population standard deviation, worst, in trial order.
    for (int s = 0; s < nsigmas; ++s) {
        double sum = 0.0, sum2 = 0.0, worst = 1e300;
        for (int i = 0; i < n; ++i) {
            const double v = st[static_cast<size_t>(s) * n + i];
            sum += v;
            sum2 += v * v;
            worst = std::min(worst, v);
        }
        const double mean = sum / n;
        const double var = std::max(0.0, sum2 / n - mean * mean);
        out4[4 * s] = sigmas_nm[s];
        out4[4 * s + 1] = mean;
        out4[4 * s + 2] = std::sqrt(var);
        out4[4 * s + 3] = worst;
    }
    if (strehl) std::copy(st.begin(), st.end(), strehl);
    return CRS_OK;
}
