// Schedule 1 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 512 threads, a lane-pair pass in the
// 8192-point transform, the passes of rk4ip_sp and raman_real in their loops
// (stage blocks of 4 points).
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(8192, 512, 0x08, 1)
