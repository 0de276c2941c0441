// Ensemble kernels (mixed 4096/8192/16384 grids, one launch) at 1024 threads
// per realisation.
#include "ensemble.hpp"
GN_DEFINE_ENSEMBLE(1024)
