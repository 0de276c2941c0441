// Schedule 3 of the 16384-sample grid's spectral-domain adaptive RK4IP
// kernels, one of the autotuner's candidates: two 1024-thread workgroups per
// instance (half and fine steps beside the coarse step), lane-pair passes in
// the 16384-point transform.
#include "kernels.hpp"
GN_DEFINE_PAIR_SCHEDULE(16384, 1024, 0x08, 3)
