// Schedule 2 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, lane-pair passes in both
// transforms, unrolled passes, stage blocks of 8 points.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(8192, 512, 0x6288, 2)
