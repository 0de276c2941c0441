// Kernel instantiations for N = 8192 samples, 512 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(8192, 512)
