// Schedule 3 of the 4096-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, lane-pair passes in both
// transforms.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(4096, 512, 0x88, 3)
