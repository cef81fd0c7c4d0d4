// mg_kernels.hpp -- device code of the multigrid-preconditioned CG path.
//
// Index conventions inside kernels are 0-based: cell (i,j,k), i fastest,
// c = i + nx*(j + ny*k).  x and z wrap periodically, y clamps to the first/last
// interior cell (homogeneous Neumann ghost = interior, boundaryField_S.f90:713-725).
// The red/black colour is the reference's parity of the 1-based index sum
// (rbgs.f90:121): colour(i,j,k) = (i+j+k+3) mod 2 with 0 = red (updated first).
#pragma once
#include "common.hpp"

namespace tbfg {

constexpr int RED = 0;
constexpr int BLACK = 1;
constexpr int kMaxTail = 16;  // levels handled by the single-workgroup tail kernel

// A bf16 storage value: the upper half of an FP32 word.  Only the finest
// level's multigrid right-hand side and coefficients are stored in it (the
// arithmetic stays FP32, the smoother's iterate stays FP32); it converts to
// float exactly.
struct bf16s {
    uint16_t bits;
    __host__ __device__ operator float() const { return __builtin_bit_cast(float, (uint32_t)bits << 16); }
};
// float -> bf16, rounded to nearest even (NaN stays a quiet NaN)
__host__ __device__ inline bf16s to_bf16(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return bf16s{(uint16_t)((u >> 16) | 0x40u)};
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16s{(uint16_t)(u >> 16)};
}
// a value converted to a storage type (bf16 through FP32)
template <class TD, class TS>
__host__ __device__ inline TD store_as(TS v)
{
    return (TD)v;
}
template <>
__host__ __device__ inline bf16s store_as<bf16s, float>(float v)
{
    return to_bf16(v);
}
template <>
__host__ __device__ inline bf16s store_as<bf16s, double>(double v)
{
    return to_bf16((float)v);
}

// Scalars of one PCG solve, resident on the device so that no host round trip
// sits between the kernels of an iteration.
struct PcgState {
    double delta;  // r.z of the current direction (pcg.f90:120,131)
    double alpha;  // delta / (d.Ad)             (pcg.f90:222)
    double gamma;  // direction update factor     (pcg.f90:132)
    double rr;     // ||r||^2
    double qq;     // ||q||^2
    double res;    // ||r|| / ||q||
    double zmean;  // arithmetic mean removed from the preconditioned residual
    double tol;
    int32_t iter;      // direction updates made
    int32_t max_iter;
    int32_t done;      // 1: stop (converged, max_iter, or numeric failure)
    int32_t status;    // TBFG_OK / TBFG_ENOTCONVERGED / TBFG_ENUMERIC
    int32_t flexible;  // 0 Fletcher-Reeves, 1 Polak-Ribiere
    int32_t pad;
    // coarsest-level solves of this PCG solve (diagnostics)
    int32_t coarse_calls;
    int32_t coarse_iters;     // inner iterations summed over calls
    int32_t coarse_iters_max;
    int32_t pad2;
    double coarse_res_max;    // worst final relative residual
    // V-cycles per preconditioner application under a residual target (tolMG)
    int32_t mg_stop;       // 1: no further cycle in this application
    int32_t mg_cyc_total;  // cycles of this solve
    int32_t mg_cyc_max;    // most cycles in one application
    int32_t pad3;
};

// The coarsest-level solve run by multi-workgroup kernels (levels too large for
// the tail): the scalars of its flexible CG.  done: stop (converged, a zero
// right-hand side or the outer solve already stopped).
struct InnerState {
    double shift;  // (sum V q) / N, removed from the right-hand side
    double rr0;    // sum V r^2 at the start
    double delta;  // sum V r z
    double alpha;
    double gamma;
    double rlast;  // sum V r^2 after the last update
    int32_t done;
    int32_t it;
};

// Partial sums of a grid reduction: NV doubles per block of the producing
// kernel, added in block order by a one-block finalize kernel.
struct Reducer {
    double* partials;    // >= blocks * NV
};

template <class T>
struct TailLevel {
    LevelDev<T> geo;
    T* p;
    T* q;
    const T* beta;
    // the level's place in the LDS-resident tail (elements from the dynamic
    // LDS base): p, q, beta, the table block
    int lds_p, lds_q, lds_beta, lds_tab;
    // ceil(2^32 / d) for d = nx, ny, nx/2, ny/2: the tail's index arithmetic
    // divides cell indices by these with one multiply-high, exact while
    // cells * max(nx, ny) < 2^32 (checked when the plan is built)
    uint32_t dm_nx, dm_ny, dm_hx, dm_hy;
};

constexpr int kTailSegs = 2 * kMaxTail + 1;

template <class T>
struct TailPlan {
    int nlev;          // number of tail levels (index 0 = finest tail level)
    int user_last;     // tail index of the reference's coarsest level
    int bottom;        // tail index of the dense bottom level
    int n_pre, n_post;
    int cmaxit;        // inner PCG iterations on the coarsest user level
    int rev_post;      // post-smoothing black then red (the tail is the inner V-cycle's bottom part)
    T ctol;            // its relative residual target
    T *cx, *cr, *cd, *cad;  // inner PCG vectors (coarsest user level)
    const T* cvol;     // cell volumes of that level
    const double* G;   // bottom dense operator, row-major: x = G q
    PcgState* st;      // receives the coarse-solve diagnostics (may be null)
    TailLevel<T> lv[kMaxTail];
    // the arrays the LDS-resident tail stages from HBM (beta and the tables of
    // every level, q of the first): source, destination (elements from the
    // LDS base), and the first element of each in the concatenated list
    int nseg;
    const T* seg_src[kTailSegs];
    int seg_dst[kTailSegs];
    int seg_first[kTailSegs + 1];
};

}  // namespace tbfg
