// Kernel instantiations for N = 4096 samples; GN_T_4096 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_4096
#define GN_T_4096 512
#endif
#ifndef GN_C_4096
#define GN_C_4096 0
#endif
GN_DEFINE_GRID(4096, GN_T_4096, GN_C_4096)
