// Kernel instantiations for N = 16384 samples; GN_T_16384 threads per instance.
// At 1024 threads each thread owns 16 points and the LDS exchange runs in
// split (real, then imaginary) mode; 512 threads (32 points per thread) is
// the alternative layout.
#include "kernels.hpp"
#ifndef GN_T_16384
#define GN_T_16384 1024
#endif
GN_DEFINE_GRID(16384, GN_T_16384)
