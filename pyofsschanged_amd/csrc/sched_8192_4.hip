// Schedule 4 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: 1024 threads (8 points each), a lane-pair
// pass in the 8192-point transform.
#include "kernels.hpp"
GN_DEFINE_SCHEDULE(8192, 1024, 0x08, 4)
