// mg_launch.hpp -- host-side launchers of the multigrid / PCG kernels.
#pragma once
#include "mg_kernels.hpp"

namespace tbfg {


// the operator product's launch: plane-marching tiles of tx x (nt / tx) cells
// and kc planes, or (tx == 0: no tile divides the level) column segments.
// TBFG_OP_TILE ("TXxNT", or "columns") and TBFG_OP_KC override the choice.
struct OpPlan {
    int tx = 0, nt = 0, tiles_x = 0, ntiles = 0, kc = 0;
    unsigned nblocks = 0;
};
OpPlan op_plan(const LevelDev<double>& L);
// doubles of reduction scratch the CG kernels of a level need (with its operator launch)
size_t reduction_slots(const LevelDev<double>& L, const OpPlan& plan);

int launch_apply_op(const LevelDev<double>& L, const double* beta, const double* p, double* q, void* stream);
int launch_op_dot(const LevelDev<double>& L, const OpPlan& plan, const double* beta, const double* d, double* qv,
                  PcgState* st, const Reducer& red, void* stream);
// the direction d = (z - m) + gamma d_old into dnew, qv = A d, d.qv and psi +=
// alpha_prev d_old in one pass (plane-marching tiles: plan.tx != 0)
struct KernelEvents;
template <class TZ>
int launch_op_dir(const LevelDev<double>& L, const OpPlan& plan, const double* beta, const TZ* z, const double* dold,
                  double* dnew, double* psi, int init, double* qv, PcgState* st, const Reducer& red, void* stream,
                  const KernelEvents* kev = nullptr);
// psi += alpha d of the last iteration that ran (d in d0 or d1 by its parity)
int launch_psi_last(size_t n, double* psi, const double* d0, const double* d1, const PcgState* st, void* stream);
// the state of a new solve: zero but tol, max_iter and flexible
int launch_state_init(PcgState* st, double tol, int32_t max_iter, int32_t flexible, void* stream);
// r = q - A psi, ||r||^2, ||q||^2; r32 (optional): the multigrid's copy of r,
// TR = float or bf16s.  With plan.tx the operator's plane-march tiles.
template <class TR>
int launch_residual_init(const LevelDev<double>& L, const OpPlan& plan, const double* beta, const double* q,
                         const double* psi, double* r, TR* r32, PcgState* st, const Reducer& red, void* stream);
template <class TR>
int launch_pcg_update(size_t n, double* psi, double* r, TR* r32, const double* d, const double* qv,
                      PcgState* st, const Reducer& red, void* stream);
template <class TZ>
int launch_pcg_dots(size_t n, const double* r, const TZ* z, const double* qv, int init, PcgState* st,
                    const Reducer& red, void* stream);
template <class TZ>
int launch_pcg_dir(size_t n, const TZ* z, double* d, int init, const PcgState* st, void* stream);

template <class T, class TQ>
int launch_rbgs(const LevelDev<T>& L, const T* beta, const TQ* q, T* p, int colour, bool first, const int* done,
                void* stream);
// one full red/black sweep, out of place (pin -> pout); levels with a tile
// shape dividing (nx, ny).  With dots: also the partial sums of k_pcg_dots for
// z = pout (finish with launch_fin_dots over *nblocks_out partials).
struct SweepDots {
    const double* r = nullptr;
    const double* qv = nullptr;
    int init = 0;
    Reducer red{nullptr};
    size_t capacity = 0;  // doubles of red.partials
};
// events that a timed launch records around the kernel itself (the
// dispatch's own begin and end, hipExtLaunchKernel)
struct KernelEvents {
    hipEvent_t start, stop;
};
template <class T>
bool sweep_supported(const LevelDev<T>& L);
// one full red/black sweep of a small level, out of place, from bricks staged
// in LDS (levels whose dimensions the brick divides)
template <class T>
bool brick_supported(const LevelDev<T>& L);
template <class T, class TQ>
int launch_brick_sweep(const LevelDev<T>& L, const T* beta, const TQ* q, const T* pin, T* pout, bool first,
                       const int* done, void* stream);
// (TB: the storage type of beta, T or, on the finest level, bf16s)
template <class T, class TQ, class TB = T>
int launch_sweep(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* pin, T* pout, bool first,
                 const int* done, void* stream, const SweepDots* dots = nullptr, unsigned* nblocks_out = nullptr,
                 const KernelEvents* kev = nullptr);
// solveMG's stop test after V-cycle cyc: sets PcgState::mg_stop
// (T0: the storage type of the finest level's beta and q)
template <class T, class T0>
int launch_mg_res(const LevelDev<T>& L, const T0* beta, const T0* q, const T* p, double tol, int cyc, int max_cycles,
                  PcgState* st, const Reducer& red, void* stream);

// the coarsest-level solve by multi-workgroup kernels, in phases driven by the
// host: Start (project q, r = q, x = 0), FirstDir (d = z), OpUpdate (ad = A d,
// alpha, x and r updated, q = r, stop test), NextDir (gamma, d = z + gamma d),
// Finish (p = x, diagnostics)
enum class InnerOp { Start, FirstDir, OpUpdate, NextDir, Finish };
template <class T>
struct InnerArgs {
    const T* beta;
    const T* vol;
    T *q, *r, *x, *d, *ad;  // q: the level's right-hand side, also the inner V-cycle's
    const T* z;              // the inner V-cycle's output
    T* p;                    // receives x
    double tol2;             // (relative residual target)^2
};
template <class T>
int launch_inner(InnerOp op, const LevelDev<T>& L, const InnerArgs<T>& a, InnerState* is, PcgState* st,
                 const Reducer& red, void* stream);
int launch_fin_dots(unsigned nblocks, size_t ncell, int init, PcgState* st, const Reducer& red, void* stream);
template <class T, class TQ, class TB = T>
int launch_resid_restrict(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* p, T* qc, const int* done,
                          void* stream);
template <class TF, class T>
int launch_restrict(const LevelDev<T>& L, const TF* fine, T* coarse, void* stream);
template <class T>
int launch_prolong_add(const LevelDev<T>& L, const T* pc, T* p, const int* done, void* stream);
template <class TS, class TD>
int launch_convert(size_t n, const TS* src, TD* dst, void* stream);
template <class T>
int launch_tail(const TailPlan<T>* plan_dev, const int* done, size_t lds_bytes, void* stream);
// dynamic LDS the LDS-resident tail kernel may use on the current device (0: none)
template <class T>
size_t tail_lds_capacity();
template <class T>
int launch_dense_setup(const LevelDev<double>& L, const T* beta, const double* vol, double* work, double* G,
                       void* stream);

}  // namespace tbfg
