// Schedule 1 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, the plain plans (no
// lane-pair pass).
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(8192, 512, 0x00, 1)
