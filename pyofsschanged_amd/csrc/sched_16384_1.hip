// Schedule 1 of the 16384-sample grid's spectral-domain adaptive RK4IP
// kernels, one of the autotuner's candidates: 1024 threads, the plain plans
// (no lane-pair pass).
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(16384, 1024, 0x00, 1)
