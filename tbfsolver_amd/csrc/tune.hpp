// tune.hpp -- lookup in the tuning table (tune.cpp).
#pragma once
#include <array>

namespace tbfg {
// the tuned shape of a kernel family on a level (kind: TBFG_TUNE_*), or false
bool tune_lookup(int kind, int nx, int ny, int nz, int elem_bytes, int out[3]);
}  // namespace tbfg
