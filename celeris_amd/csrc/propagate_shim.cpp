// propagate_shim.cpp — the reference's C++ GPU-propagation names, so a caller built
// against src/celeris/cuda/propagate.hpp links against this library unchanged.
// Each forwards to the C ABI; see include/celeris_rs.h for the behaviour kept.
#include "../../include/celeris/cuda/propagate.hpp"
#include "../../include/celeris_rs.h"

namespace celeris::cuda {

bool available() { return crs_available() != 0; }

const char* device_name() { return crs_device_name(); }

bool propagate_psf(const double* px, const double* py, const std::complex<double>* pt, int npil,
                   double cx, double cy, double z, double k, int n, double half_window,
                   double* out) {
    return crs_propagate_psf(px, py, reinterpret_cast<const double*>(pt), npil, cx, cy, z, k, n,
                             half_window, out) != 0;
}

bool propagate_zx(const double* px, const double* py, const std::complex<double>* pt, int npil,
                  double xmin, double xmax, int nx, double zmin, double zmax, int nz, double k,
                  double* out) {
    return crs_propagate_zx(px, py, reinterpret_cast<const double*>(pt), npil, xmin, xmax, nx,
                            zmin, zmax, nz, k, out) != 0;
}

}  // namespace celeris::cuda
