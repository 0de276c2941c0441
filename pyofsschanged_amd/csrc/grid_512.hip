// Kernel instantiations for N = 512 samples; GN_T_512 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_512
#define GN_T_512 64
#endif
#ifndef GN_C_512
#define GN_C_512 0
#endif
GN_DEFINE_GRID(512, GN_T_512, GN_C_512)
