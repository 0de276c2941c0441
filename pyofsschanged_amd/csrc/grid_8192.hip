// Kernel instantiations for N = 8192 samples; GN_T_8192 threads per instance.
// GN_C_8192 is the schedule of the spectral-domain RK4IP kernels: 8 puts a
// lane-pair pass in the 8192-point transform (radices 32L, 16, 16), 0x200
// stage passes in blocks of 8 points, 0x6000 the passes of rk4ip_sp and the
// Raman transforms unrolled.
#include "kernels.hpp"
#ifndef GN_T_8192
#define GN_T_8192 512
#endif
#ifndef GN_C_8192
#define GN_C_8192 0x6208
#endif
GN_DEFINE_GRID(8192, GN_T_8192, GN_C_8192)
