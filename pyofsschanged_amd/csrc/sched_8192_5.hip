// Schedule 5 of the 8192-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: two 512-thread workgroups per instance
// (half and fine steps beside the coarse step), a lane-pair pass in the
// 8192-point transform.
#include "kernels.hpp"
GN_DEFINE_PAIR_SCHEDULE(8192, 512, 0x08, 5)
