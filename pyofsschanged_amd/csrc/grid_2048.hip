// Kernel instantiations for N = 2048 samples; GN_T_2048 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_2048
#define GN_T_2048 256
#endif
#ifndef GN_C_2048
#define GN_C_2048 0
#endif
GN_DEFINE_GRID(2048, GN_T_2048, GN_C_2048)
