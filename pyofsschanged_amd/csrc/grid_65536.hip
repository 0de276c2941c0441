// Launch table of the 65536-sample grid (two-level transform, fields in
// workspace slots; big.hpp).
#include "big.hpp"

GN_DEFINE_BIG_GRID(65536)
