// Schedule 2 of the 16384-sample grid's spectral-domain adaptive RK4IP
// kernels, one of the autotuner's candidates: 1024 threads, lane-pair passes
// in the 16384-point transform, unrolled passes, stage blocks of 4 points.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(16384, 1024, 0x6008, 2)
