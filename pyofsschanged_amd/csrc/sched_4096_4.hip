// Schedule 4 of the 4096-sample grid's spectral-domain adaptive RK4IP kernels,
// one of the autotuner's candidates: two 512-thread workgroups per instance
// (half and fine steps beside the coarse step), plain plans.
#include "kernels.hpp"
GN_DEFINE_PAIR_SCHEDULE(4096, 512, 0x00, 4)
