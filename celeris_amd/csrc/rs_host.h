// rs_host.h — host-side helpers shared by the C ABI's translation units (rs_api.cpp,
// rs_analysis.cpp, rs_tolerance.cpp).  Not part of the public ABI.
#pragma once

namespace crs {

// The calling thread's message for crs_last_error(): set_error formats and replaces it,
// clear_error empties it (each entry point clears it on entry).
void set_error(const char* fmt, ...);
void clear_error();

}  // namespace crs
