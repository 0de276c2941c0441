// Schedule 3 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, a lane-pair pass in the
// 8192-point transform, stage passes in blocks of 2.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(8192, 512, 0x108, 3)
