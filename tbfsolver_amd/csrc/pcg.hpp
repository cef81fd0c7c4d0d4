// pcg.hpp -- the MG_MODE solver handle: PCG with a geometric multigrid
// preconditioner (pcg.f90, multiGrid.f90, rbgs.f90, poissMat.f90 of the reference).
#pragma once
#include <memory>

#include "fps.hpp"
#include "mg_launch.hpp"

struct tbfg_pcg {
    tbfg_grid_desc g{};
    tbfg_pcg_params prm{};
    bool mixed = false;
    bool b16 = false;  // mixed, with the finest level's multigrid q and beta stored in bf16
    int nuser = 0;   // reference levels (nLevels)
    int ntot = 0;    // reference levels + sub-levels of the coarsest-solve hierarchy
    int t0 = 0;      // first level run by the tail kernel
    int kin = -1;    // coarsest user level when its solve runs as multi-workgroup kernels (too large for the tail)
    size_t n0 = 0;   // finest cells

    std::vector<tbfg::HostLevel> hl;
    std::vector<std::unique_ptr<tbfg::LevelTables<double>>> tabd;  // FP64 tables (all levels)
    std::vector<std::unique_ptr<tbfg::LevelTables<float>>> tabf;   // FP32 tables (mixed)

    // per-level MG arrays (element type float in mixed mode, double otherwise);
    // level 0: p = z, q = r (double), beta = caller's (FP64) or a float (b16: bf16) copy (mixed)
    std::vector<tbfg::DevBuf> p, q, beta;
    std::vector<tbfg::DevBuf> palt;     // second p buffer of the levels smoothed out of place
    std::vector<char> fused;            // level smoothed out of place, one launch per sweep: 1 plane march, 2 bricks
    bool fuse_dots = true;              // PCG dot products formed by the finest level's last sweep
    size_t tail_lds = 0;                // dynamic LDS of the LDS-resident tail kernel (0: HBM form)
    tbfg::OpPlan opp;                   // launch of the finest level's operator product
    bool fuse_op = false;               // direction + operator + deferred psi update in one pass (d2: the pair)
    bool queued_fft = false;            // the last queued solve used the FFT preconditioner (pcg_result)
    tbfg::DevBuf r, d, d2, qv;          // CG vectors (FP64)
    tbfg::DevBuf r32;                   // FP32 (b16: bf16) copy of r, the multigrid's right-hand side (mixed)
    tbfg::DevBuf st, partials, counter; // PcgState + reduction scratch
    tbfg::DevBuf plan;                  // TailPlan<T>
    tbfg::DevBuf G, dwork, vol;         // dense bottom operator and its set-up scratch
    tbfg::DevBuf cx, cr, cd, cad, cvol; // inner CG of the coarsest-level solve
    tbfg::DevBuf instate;               // its InnerState (multi-workgroup form)
    int32_t* hin = nullptr;             // pinned copy of InnerState::done, ::it (the host's stop test)
    int inner_guess = 8;                // iterations launched before the first stop test (last solve's count)
    int32_t* hmg = nullptr;             // pinned copy of PcgState::mg_stop (solveMG's loop, mg_tol > 0)
    tbfg::DevBuf hpsi, hbeta, hq;       // device copies for the host-array entry point
    tbfg_fps* fps = nullptr;            // FFT preconditioner (TBFG_PRECOND_FFT)
    tbfg::DevBuf zfft;                  // its output z (FP64)
    tbfg::PcgState* hst = nullptr;      // pinned mirror of the device state (final, after a solve)
    tbfg::PcgState* hring = nullptr;    // pinned copies of the state after consecutive iterations
    hipEvent_t hev[2] = {nullptr, nullptr};
    double* hstage = nullptr;           // pinned staging (3 * n0 doubles)
    int64_t footprint = 0;

    // kernel timing (tbfg_pcg_profile_begin/end)
    int prof_kernel = -1;
    std::vector<hipEvent_t> prof_ev;  // pairs
    int prof_used = 0;
    int64_t prof_launches = 0;
    double prof_ms = 0.0;

    ~tbfg_pcg();
};

namespace tbfg {
int pcg_create(const tbfg_grid_desc* g, const tbfg_pcg_params* prm, tbfg_pcg** out);
int pcg_solve_dev(tbfg_pcg* h, double* psi, const double* beta, const double* q, int32_t* iters, double* res,
                  void* stream);
int pcg_solve_queued_dev(tbfg_pcg* h, double* psi, const double* beta, const double* q, int32_t iterations,
                         void* stream);
int pcg_result(tbfg_pcg* h, int32_t* iters, double* res);
int pcg_precondition_dev(tbfg_pcg* h, const double* beta, const double* r, double* z, void* stream);
int pcg_profile_begin(tbfg_pcg* h, int kernel);
int pcg_profile_end(tbfg_pcg* h, int64_t* launches, double* total_ms, double* bytes_per_launch);
}  // namespace tbfg
