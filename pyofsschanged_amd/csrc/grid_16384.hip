// Kernel instantiations for N = 16384 samples; GN_T_16384 threads per instance.
// At 1024 threads each thread owns 16 points and the LDS exchange runs in
// split (real, then imaginary) mode; 512 threads (32 points per thread) is
// the alternative layout. GN_C_16384 is the schedule of the spectral-domain
// RK4IP kernels: 8 puts lane-pair passes in the 16384-point transform
// (radices 32L, 32L, 16), 0x100 stage passes in blocks of 2 points, 0x6000
// the passes of rk4ip_sp and the Raman transforms unrolled.
#include "kernels.hpp"
#ifndef GN_T_16384
#define GN_T_16384 1024
#endif
#ifndef GN_C_16384
#define GN_C_16384 0x6108
#endif
GN_DEFINE_GRID(16384, GN_T_16384, GN_C_16384)
