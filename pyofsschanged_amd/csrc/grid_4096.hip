// Kernel instantiations for N = 4096 samples, 512 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(4096, 512)
