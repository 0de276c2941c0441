// Kernel instantiations for N = 512 samples, 64 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(512, 64)
