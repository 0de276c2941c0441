// Schedule 1 of the 4096-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, lane-pair passes in the
// Raman convolution's 2048-point transform.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(4096, 512, 0x80, 1)
