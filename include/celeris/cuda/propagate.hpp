#pragma once
// Drop-in replacement header for the reference's src/celeris/cuda/propagate.hpp:
// the same namespace, names and signatures (propagate.hpp:16-38), implemented by
// libceleris_rs.so on MI355X (celeris_amd/csrc/propagate_shim.cpp -> include/celeris_rs.h).
// Callers (analysis/focal.cpp, analysis/throughfocus.cpp, cli/bench.cpp) compile
// against it unchanged.
//
//   available()      nonzero HIP device count                     (propagate.hpp:20)
//   device_name()    name of device 0, "" if none                  (propagate.hpp:21)
//   propagate_psf    |E|^2 on an n x n window at distance z        (propagate.hpp:27-30)
//   propagate_zx     |E|^2 on an nz x nx slice in the plane y = 0   (propagate.hpp:35-38)
// Both return false, leaving `out` untouched, when no device is visible, the sizes
// are rejected (npil <= 0, n <= 0; nx <= 1 or nz <= 1) or a HIP call fails, so the
// caller falls back to its CPU loop exactly as with the reference.

#include <complex>

namespace celeris::cuda {

bool available();
const char* device_name();

bool propagate_psf(const double* px, const double* py, const std::complex<double>* pt, int npil,
                   double cx, double cy, double z, double k, int n, double half_window,
                   double* out);

bool propagate_zx(const double* px, const double* py, const std::complex<double>* pt, int npil,
                  double xmin, double xmax, int nx, double zmin, double zmax, int nz, double k,
                  double* out);

}  // namespace celeris::cuda
