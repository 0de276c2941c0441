// Kernel instantiations for N = 1024 samples, 128 threads per instance.
#include "kernels.hpp"
GN_DEFINE_GRID(1024, 128)
