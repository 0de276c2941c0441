// Kernel instantiations for N = 2048 samples, 256 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(2048, 256)
