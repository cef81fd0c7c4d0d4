// tune.cpp -- the tuning table: per level shape and precision, the tile shape
// a kernel family launches with (set by the tuner, tbfsolver_amd/tune.py, or
// from its saved table).  Entries are looked up when a kernel is launched
// (sweeps) or when a handle is created (the operator product); without an
// entry the built-in order of preference applies.
#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "tbfgpu.h"
#include "tune.hpp"

namespace tbfg {
namespace {
using Key = std::tuple<int, int, int, int, int>;  // kind, nx, ny, nz, bytes per value
std::mutex g_mu;
std::map<Key, std::array<int, 3>>& table()
{
    static std::map<Key, std::array<int, 3>> t;
    return t;
}
}  // namespace

bool tune_lookup(int kind, int nx, int ny, int nz, int elem_bytes, int out[3])
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (table().empty()) return false;
    auto it = table().find(Key{kind, nx, ny, nz, elem_bytes});
    if (it == table().end()) return false;
    for (int i = 0; i < 3; ++i) out[i] = it->second[i];
    return true;
}
}  // namespace tbfg

extern "C" int tbfg_tune_set(int32_t kind, const int32_t dims[3], int32_t elem_bytes, const int32_t shape[3])
{
    if (kind < TBFG_TUNE_SWEEP || kind > TBFG_TUNE_OP || !dims || !shape || (elem_bytes != 4 && elem_bytes != 8)) {
        tbfg::set_error("tbfg_tune_set: invalid argument");
        return TBFG_EINVAL;
    }
    std::lock_guard<std::mutex> lk(tbfg::g_mu);
    const tbfg::Key key{kind, dims[0], dims[1], dims[2], elem_bytes};
    if (shape[0] == 0 && shape[1] == 0 && shape[2] == 0)
        tbfg::table().erase(key);
    else
        tbfg::table()[key] = {shape[0], shape[1], shape[2]};
    return TBFG_OK;
}

extern "C" int tbfg_tune_clear(void)
{
    std::lock_guard<std::mutex> lk(tbfg::g_mu);
    tbfg::table().clear();
    return TBFG_OK;
}

extern "C" int tbfg_tune_get(int32_t kind, const int32_t dims[3], int32_t elem_bytes, int32_t shape[3])
{
    if (!dims || !shape) {
        tbfg::set_error("tbfg_tune_get: invalid argument");
        return TBFG_EINVAL;
    }
    int s[3] = {0, 0, 0};
    tbfg::tune_lookup(kind, dims[0], dims[1], dims[2], elem_bytes, s);
    for (int i = 0; i < 3; ++i) shape[i] = s[i];
    return TBFG_OK;
}
