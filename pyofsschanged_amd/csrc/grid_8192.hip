// Kernel instantiations for N = 8192 samples; GN_T_8192 threads per instance.
#include "kernels.hpp"
#ifndef GN_T_8192
#define GN_T_8192 512
#endif
GN_DEFINE_GRID(8192, GN_T_8192)
