// ref_driver.cpp — runs the reference's own CPU propagation and analysis code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/ref.mk from this file plus the
// reference sources where they lie (src/celeris/analysis/{focal,throughfocus,
// chromatic,mtf,field}.cpp, src/celeris/design/{metalens,phase_profile}.cpp); the
// binary lands in oracle/_ref/ and is used to pin the FP64 restatement in
// oracle/rs_oracle.c and to write golden vectors.  Nothing here is shipped.
//
// Protocol: `ref_driver <command> <in.f64> <out.f64>`.  Both files are flat
// little-endian float64 arrays; integers travel as float64.
//
//   pillars      : npil, px[npil], py[npil], (re,im)[npil], cx, cy, z, lambda, n, W
//                  -> intensity[n*n]       (celeris::propagate_pillars, focal.cpp:205)
//   lens commands share a lens block:
//                  n_cells, period, fill_map[n_cells^2], nlib, fill[nlib],
//                  phase[nlib], amplitude[nlib]
//   focus        : lens, f, lambda, D      -> strehl, phase_strehl, fwhm, dl, ee
//                                             (celeris::analyze_focus, focal.cpp:32)
//   psf          : lens, f, lambda, D, n, W -> intensity[n*n]  (compute_psf, focal.cpp:148)
//   throughfocus : lens, f, lambda, D      -> N, z[N], I[N], z_peak, dof,
//                                             cnx, cnz, caustic[cnz*cnx]
//                                             (analyze_through_focus, throughfocus.cpp:15)
//   chromatic    : lens, f, lambda, D, lmin, lmax, n -> n x (lambda, f, rel)
//                                             (analyze_chromatic, chromatic.cpp:9)
//   chromatic_disp : lens, f, lambda, D, lmin, lmax, n, nlibs, nlibs x (lambda,
//                  fill[nlib], phase[nlib], amplitude[nlib]) -> n x (lambda, f, rel)
//                                             (analyze_chromatic_dispersive, chromatic.cpp:60)
//   mtf          : lens, f, lambda, D      -> m, freq[m], mtf[m], ideal[m], cutoff
//                                             (analyze_mtf, mtf.cpp:10)
//   fieldgrid    : lens, f, lambda, D, max_angle, n_half, psf_n -> n, then n*n x
//                  (theta_x, theta_y, rel_strehl, fwhm_x, fwhm_y, cx, cy)
//                                             (analyze_field_grid, field.cpp:150)
//   psffield     : lens, f, lambda, D, angle, n, W, on_axis_peak -> angle, cx, rel_strehl,
//                  intensity[n*n]             (compute_psf_field, focal.cpp:240)
//   fov          : lens, f, lambda, D, m, angles[m] -> m x (angle, rel_strehl, shift)
//                                             (analyze_field_of_view, field.cpp:12)
//   widefov      : lens, f, lambda, D_lens, D_stop, stop_dist, m, angles[m]
//                  -> m x (angle, rel_strehl, shift)   (analyze_wide_fov, field.cpp:78)
//   design       : nlib, fill[nlib], phase[nlib], amplitude[nlib], lambda, period, f, D,
//                  amplitude_weight -> n, rms_phase_error_deg, mean_amplitude, fill_map[n*n]
//                                             (design_metalens, metalens.cpp:230-273)
//   tolerance    : lens, f, lambda, D, nsig, sigmas[nsig], trials, seed
//                  -> nsig x (sigma, mean, std, worst)   (analyze_tolerance, tolerance.cpp:12)
//   pgm          : nx, ny, gamma, values[nx*ny] -> ok; the image itself is written to
//                  <out.f64>.pgm              (write_pgm, io/image.cpp:10)
//   achro        : n_cells, period, D, f, nb, lambda[nb], center_index, center_lambda,
//                  natoms, per atom (amplitude[nb], phase[nb]), atom_index[n_cells^2]
//                  -> nb x (lambda, f, rel)   (verify_achromatic_focus, achromatic.cpp:288)
//   pbachro      : n_cells, period, D, f, handedness, nb, lambda[nb], center_index,
//                  center_lambda, natoms, per atom (re, im)[nb], atom_index[n_cells^2],
//                  rotation[n_cells^2] -> nb x (lambda, f, rel)
//                                             (verify_pb_achromatic_focus, pb_achromatic.cpp:255)
#include "celeris/analysis/chromatic.hpp"
#include "celeris/analysis/field.hpp"
#include "celeris/analysis/focal.hpp"
#include "celeris/analysis/mtf.hpp"
#include "celeris/analysis/throughfocus.hpp"
#include "celeris/analysis/tolerance.hpp"
#include "celeris/design/metalens.hpp"
#include "celeris/design/achromatic.hpp"
#include "celeris/design/pb_achromatic.hpp"
#include "celeris/io/image.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace celeris;

namespace {

struct Reader {
    std::vector<double> v;
    std::size_t i = 0;
    double next() {
        if (i >= v.size()) { std::fprintf(stderr, "ref_driver: input too short\n"); std::exit(2); }
        return v[i++];
    }
    int next_int() { return static_cast<int>(next()); }
};

std::vector<double> read_all(const char* path) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    std::vector<double> out;
    double buf[4096];
    std::size_t got;
    while ((got = std::fread(buf, sizeof(double), 4096, f)) > 0) out.insert(out.end(), buf, buf + got);
    std::fclose(f);
    return out;
}

void write_all(const char* path, const std::vector<double>& v) {
    std::FILE* f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

void read_lens(Reader& r, MetalensDesign& lens, UnitCellLibrary& lib) {
    lens.n_cells = r.next_int();
    lens.period_um = r.next();
    lens.fill_map.resize(static_cast<std::size_t>(lens.n_cells) * lens.n_cells);
    for (double& x : lens.fill_map) x = r.next();
    lens.rms_phase_error_deg = 0.0;
    lens.mean_amplitude = 0.0;
    const int nlib = r.next_int();
    lib.period_um = lens.period_um;
    lib.wavelength_um = 0.0;
    lib.thickness_um = 0.0;
    lib.fill.resize(nlib);
    lib.phase.resize(nlib);
    lib.amplitude.resize(nlib);
    for (double& x : lib.fill) x = r.next();
    for (double& x : lib.phase) x = r.next();
    for (double& x : lib.amplitude) x = r.next();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_driver <command> <in.f64> <out.f64>\n");
        return 2;
    }
    const std::string cmd = argv[1];
    Reader r{read_all(argv[2])};
    std::vector<double> out;

    if (cmd == "pillars") {
        const int npil = r.next_int();
        std::vector<double> px(npil), py(npil);
        std::vector<cdouble> t(npil);
        for (double& x : px) x = r.next();
        for (double& y : py) y = r.next();
        for (cdouble& c : t) { const double re = r.next(); const double im = r.next(); c = {re, im}; }
        const double cx = r.next(), cy = r.next(), z = r.next(), lam = r.next();
        const int n = r.next_int();
        const double W = r.next();
        PsfMap m = propagate_pillars(px, py, t, cx, cy, z, lam, n, W);
        out = m.intensity;
    } else if (cmd == "pgm") {
        const int nx = r.next_int(), ny = r.next_int();
        const double gamma = r.next();
        std::vector<double> v(static_cast<std::size_t>(std::max(nx, 0)) * std::max(ny, 0));
        for (double& x : v) x = r.next();
        out = {write_pgm(std::string(argv[3]) + ".pgm", nx, ny, v, gamma) ? 1.0 : 0.0};
    } else if (cmd == "design") {
        UnitCellLibrary lib;
        const int nlib = r.next_int();
        lib.fill.resize(nlib);
        lib.phase.resize(nlib);
        lib.amplitude.resize(nlib);
        for (double& x : lib.fill) x = r.next();
        for (double& x : lib.phase) x = r.next();
        for (double& x : lib.amplitude) x = r.next();
        lib.wavelength_um = r.next();
        lib.period_um = r.next();
        lib.thickness_um = 0.0;
        const double f = r.next(), D = r.next(), w = r.next();
        MetalensDesign d = design_metalens(lib, f, D, w);
        out = {static_cast<double>(d.n_cells), d.rms_phase_error_deg, d.mean_amplitude};
        out.insert(out.end(), d.fill_map.begin(), d.fill_map.end());
    } else if (cmd == "achro" || cmd == "pbachro") {
        const int n_cells = r.next_int();
        const double period = r.next(), D = r.next(), f = r.next();
        const int hand = cmd == "pbachro" ? r.next_int() : 1;
        const int nb = r.next_int();
        std::vector<double> lam(nb);
        for (double& x : lam) x = r.next();
        const int center_index = r.next_int();
        const double center_lambda = r.next();
        const int natoms = r.next_int();
        const std::size_t nsite = static_cast<std::size_t>(n_cells) * n_cells;
        std::vector<AchromaticFocalPoint> pts;
        if (cmd == "achro") {
            DispersiveLibrary lib;
            lib.wavelengths_um = lam;
            lib.center_index = center_index;
            lib.center_wavelength_um = center_lambda;
            lib.period_um = period;
            lib.atoms.resize(natoms);
            for (DispersiveAtom& a : lib.atoms) {
                a.amplitude.resize(nb);
                a.phase_rad.resize(nb);
                for (double& x : a.amplitude) x = r.next();
                for (double& x : a.phase_rad) x = r.next();
            }
            AchromaticDesign d;
            d.n_cells = n_cells;
            d.period_um = period;
            d.diameter_um = D;
            d.focal_length_um = f;
            d.center_wavelength_um = center_lambda;
            d.atom_index.resize(nsite);
            for (int& x : d.atom_index) x = r.next_int();
            pts = verify_achromatic_focus(lib, d);
        } else {
            DispersivePbLibrary lib;
            lib.wavelengths_um = lam;
            lib.center_index = center_index;
            lib.center_wavelength_um = center_lambda;
            lib.period_um = period;
            lib.atoms.resize(natoms);
            for (DispersivePbAtom& a : lib.atoms) {
                a.cross.resize(nb);
                for (cdouble& c : a.cross) { const double re = r.next(); const double im = r.next(); c = {re, im}; }
            }
            PbAchromaticDesign d;
            d.n_cells = n_cells;
            d.period_um = period;
            d.diameter_um = D;
            d.focal_length_um = f;
            d.handedness = hand;
            d.center_wavelength_um = center_lambda;
            d.atom_index.resize(nsite);
            for (int& x : d.atom_index) x = r.next_int();
            d.rotation_rad.resize(nsite);
            for (double& x : d.rotation_rad) x = r.next();
            pts = verify_pb_achromatic_focus(lib, d);
        }
        for (const auto& p : pts) out.insert(out.end(), {p.wavelength_um, p.focal_length_um, p.rel_peak});
    } else {
        MetalensDesign lens;
        UnitCellLibrary lib;
        read_lens(r, lens, lib);
        const double f = r.next(), lam = r.next(), D = r.next();
        if (cmd == "focus") {
            FocalAnalysis a = analyze_focus(lens, lib, f, lam, D);
            out = {a.strehl, a.phase_strehl, a.fwhm_um, a.diffraction_limit_um, a.encircled_energy};
        } else if (cmd == "psf") {
            const int n = r.next_int();
            const double W = r.next();
            PsfMap m = compute_psf(lens, lib, f, lam, D, n, W);
            out = m.intensity;
        } else if (cmd == "throughfocus") {
            ThroughFocus tf = analyze_through_focus(lens, lib, f, lam, D);
            out.push_back(static_cast<double>(tf.z_um.size()));
            for (float z : tf.z_um) out.push_back(z);
            for (float v : tf.intensity) out.push_back(v);
            out.push_back(tf.z_peak_um);
            out.push_back(tf.dof_um);
            out.push_back(tf.caustic_nx);
            out.push_back(tf.caustic_nz);
            for (float v : tf.caustic) out.push_back(v);
        } else if (cmd == "chromatic_disp") {
            // lmin, lmax, n, nlibs, then nlibs x (lambda, fill[nlib], phase[nlib],
            // amplitude[nlib]) on the design library's fill grid; build_lib_at(lambda)
            // returns the table whose lambda matches (analyze_chromatic_dispersive asks
            // for the design wavelength and then each lambda_i).
            const double lmin = r.next(), lmax = r.next();
            const int n = r.next_int(), nlibs = r.next_int();
            std::vector<UnitCellLibrary> tables(nlibs);
            for (UnitCellLibrary& t : tables) {
                t = lib;
                t.wavelength_um = r.next();
                for (double& x : t.fill) x = r.next();
                for (double& x : t.phase) x = r.next();
                for (double& x : t.amplitude) x = r.next();
            }
            auto build_lib_at = [&](double wl) -> UnitCellLibrary {
                for (const UnitCellLibrary& t : tables)
                    if (std::fabs(t.wavelength_um - wl) <= 1e-12 * wl) return t;
                std::fprintf(stderr, "ref_driver: no library table at lambda %.17g\n", wl);
                std::exit(3);
            };
            auto pts = analyze_chromatic_dispersive(lens, build_lib_at, f, lam, D, lmin, lmax, n);
            for (const auto& p : pts) {
                out.push_back(p.wavelength_um);
                out.push_back(p.focal_length_um);
                out.push_back(p.rel_peak);
            }
        } else if (cmd == "chromatic") {
            const double lmin = r.next(), lmax = r.next();
            const int n = r.next_int();
            auto pts = analyze_chromatic(lens, lib, f, lam, D, lmin, lmax, n);
            for (const auto& p : pts) {
                out.push_back(p.wavelength_um);
                out.push_back(p.focal_length_um);
                out.push_back(p.rel_peak);
            }
        } else if (cmd == "mtf") {
            MtfCurve c = analyze_mtf(lens, lib, f, lam, D);
            out.push_back(static_cast<double>(c.mtf.size()));
            for (float v : c.freq_cyc_mm) out.push_back(v);
            for (float v : c.mtf) out.push_back(v);
            for (float v : c.mtf_ideal) out.push_back(v);
            out.push_back(c.cutoff_cyc_mm);
        } else if (cmd == "fieldgrid") {
            const double max_angle = r.next();
            const int n_half = r.next_int(), psf_n = r.next_int();
            FieldGrid g = analyze_field_grid(lens, lib, f, lam, D, max_angle, n_half, psf_n);
            out.push_back(g.n);
            for (const auto& p : g.points) {
                out.insert(out.end(), {p.theta_x_deg, p.theta_y_deg, p.rel_strehl, p.fwhm_x_um,
                                       p.fwhm_y_um, p.cx_um, p.cy_um});
            }
        } else if (cmd == "psffield") {
            const double angle = r.next();
            const int n = r.next_int();
            const double W = r.next(), on_axis_peak = r.next();
            FieldPsf p = compute_psf_field(lens, lib, f, lam, D, angle, n, W, on_axis_peak);
            out = {p.angle_deg, p.cx_um, p.rel_strehl};
            out.insert(out.end(), p.psf.intensity.begin(), p.psf.intensity.end());
        } else if (cmd == "fov" || cmd == "widefov") {
            double Dstop = 0.0, stop_dist = 0.0;
            if (cmd == "widefov") {
                Dstop = r.next();
                stop_dist = r.next();
            }
            const int m = r.next_int();
            std::vector<double> angles(m);
            for (double& a : angles) a = r.next();
            std::vector<FieldPoint> pts =
                cmd == "fov" ? analyze_field_of_view(lens, lib, f, lam, D, angles)
                             : analyze_wide_fov(lens, lib, f, lam, D, Dstop, stop_dist, angles);
            for (const auto& p : pts) out.insert(out.end(), {p.angle_deg, p.rel_strehl, p.spot_shift_um});
        } else if (cmd == "tolerance") {
            const int nsig = r.next_int();
            std::vector<double> sig(nsig);
            for (double& x : sig) x = r.next();
            const int trials = r.next_int();
            const unsigned seed = static_cast<unsigned>(r.next());
            for (const ToleranceResult& t : analyze_tolerance(lens, lib, f, lam, D, sig, trials, seed))
                out.insert(out.end(), {t.sigma_nm, t.mean_strehl, t.std_strehl, t.worst_strehl});
        } else {
            std::fprintf(stderr, "ref_driver: unknown command %s\n", cmd.c_str());
            return 2;
        }
    }
    write_all(argv[3], out);
    return 0;
}
