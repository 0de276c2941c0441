// Kernel instantiations for N = 256 samples, 64 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(256, 64)
