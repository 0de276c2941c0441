// Kernel instantiations for N = 256 samples; GN_T_256 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_256
#define GN_T_256 64
#endif
#ifndef GN_C_256
#define GN_C_256 0
#endif
GN_DEFINE_GRID(256, GN_T_256, GN_C_256)
