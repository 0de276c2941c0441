// Kernel instantiations for N = 1024 samples; GN_T_1024 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_1024
#define GN_T_1024 128
#endif
#ifndef GN_C_1024
#define GN_C_1024 0
#endif
GN_DEFINE_GRID(1024, GN_T_1024, GN_C_1024)
