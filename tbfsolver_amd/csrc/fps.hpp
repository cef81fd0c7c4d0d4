// fps.hpp -- the FAST_MODE solver handle: FFT in x and z, tridiagonal in y
// (fastPoissonSolver.f90 of the reference).
#pragma once
#include "common.hpp"

struct tbfg_fps;

namespace tbfg {
int fps_create(const tbfg_grid_desc* g, const tbfg_fps_opts* o, tbfg_fps** out);
int fps_solve_dev(tbfg_fps* h, const double* s, double* p, void* stream);
void fps_destroy(tbfg_fps* h);
const tbfg_grid_desc* fps_grid(const tbfg_fps* h);
// the constant-coefficient operator the FFT solver inverts (q = L p), FP64
int fps_apply_operator_dev(tbfg_fps* h, const double* p, double* q, void* stream);
int64_t fps_footprint(const tbfg_fps* h);
}  // namespace tbfg
