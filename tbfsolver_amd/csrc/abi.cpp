// abi.cpp -- the exported C ABI (include/tbfgpu.h).  Every entry point converts
// exceptions and device failures into status codes; nothing aborts.
#include <cstring>
#include <new>

#include "fields.hpp"
#include "fps.hpp"
#include "mg_launch.hpp"
#include "pcg.hpp"

namespace tbfg {
const char* last_error();
int make_velocity(const tbfg_grid_desc* g, uint64_t seed, double* ux, double* uy, double* uz);
int make_bubble_density(const tbfg_grid_desc* g, int nb, double D, uint64_t seed, int nref, double rho_l,
                        double rho_g, double* rho, double* centres);
}  // namespace tbfg

using namespace tbfg;

#define ABI_GUARD(...)                                         \
    try {                                                      \
        clear_error();                                         \
        __VA_ARGS__                                            \
    } catch (const std::bad_alloc&) {                          \
        set_error("host allocation failed");                   \
        return TBFG_ENOMEM;                                    \
    } catch (...) {                                            \
        set_error("unexpected host exception");                \
        return TBFG_EINVAL;                                    \
    }

namespace {

// Fortran-bounded host arrays: lb/ub per axis, i fastest.
struct FArr {
    int64_t lb[3], ub[3], ext[3];
    size_t at(int64_t i, int64_t j, int64_t k) const
    {
        return (size_t)(i - lb[0]) + (size_t)ext[0] * ((size_t)(j - lb[1]) + (size_t)ext[1] * (size_t)(k - lb[2]));
    }
};

int make_farr(const int64_t* lb, const int64_t* ub, const int32_t* n, const char* name, FArr& f)
{
    if (!lb || !ub) {
        set_error("%s: NULL bounds", name);
        return TBFG_EINVAL;
    }
    for (int a = 0; a < 3; ++a) {
        if (lb[a] > 1 || ub[a] < n[a] || ub[a] - lb[a] + 1 > (int64_t)n[a] + 2 * 64) {
            set_error("%s: bounds %lld..%lld on axis %d do not hold cells 1..%d", name, (long long)lb[a],
                      (long long)ub[a], a, n[a]);
            return TBFG_EINVAL;
        }
        if (1 - lb[a] > n[a] || ub[a] - n[a] > n[a]) {
            set_error("%s: halo on axis %d is wider than the grid", name, a);
            return TBFG_EINVAL;
        }
        f.lb[a] = lb[a];
        f.ub[a] = ub[a];
        f.ext[a] = ub[a] - lb[a] + 1;
    }
    return TBFG_OK;
}

// (host copies of whole fields: one plane per OpenMP iteration, so the
// host-array entry points are not bound by one core's copy rate)
void gather(const double* src, const FArr& f, const int32_t* n, double* dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= n[2]; ++k)
        for (int64_t j = 1; j <= n[1]; ++j) {
            const double* row = src + f.at(1, j, k);
            double* out = dst + (size_t)n[0] * ((size_t)(j - 1) + (size_t)n[1] * (size_t)(k - 1));
            std::memcpy(out, row, sizeof(double) * (size_t)n[0]);
        }
}

// interior write, then updateBoundaries for a cell field of the channel:
// Neumann ghost layer in y, then periodic halos in x and z over the full
// extent of the other axes (field.f90:950-974, boundaryField_S.f90:713-725).
void scatter_fill(const double* src, const FArr& f, const int32_t* n, double* dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= n[2]; ++k)
        for (int64_t j = 1; j <= n[1]; ++j)
            std::memcpy(dst + f.at(1, j, k), src + (size_t)n[0] * ((size_t)(j - 1) + (size_t)n[1] * (size_t)(k - 1)),
                        sizeof(double) * (size_t)n[0]);
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= n[2]; ++k)
        for (int64_t i = 1; i <= n[0]; ++i) {
            if (f.lb[1] <= 0) dst[f.at(i, 0, k)] = dst[f.at(i, 1, k)];
            if (f.ub[1] >= n[1] + 1) dst[f.at(i, n[1] + 1, k)] = dst[f.at(i, n[1], k)];
        }
#pragma omp parallel for schedule(static)
    for (int64_t k = 1; k <= n[2]; ++k)
        for (int64_t j = f.lb[1]; j <= f.ub[1]; ++j) {
            for (int64_t i = f.lb[0]; i <= 0; ++i) dst[f.at(i, j, k)] = dst[f.at(i + n[0], j, k)];
            for (int64_t i = n[0] + 1; i <= f.ub[0]; ++i) dst[f.at(i, j, k)] = dst[f.at(i - n[0], j, k)];
        }
    for (int64_t j = f.lb[1]; j <= f.ub[1]; ++j)
        for (int64_t i = f.lb[0]; i <= f.ub[0]; ++i) {
            for (int64_t k = f.lb[2]; k <= 0; ++k) dst[f.at(i, j, k)] = dst[f.at(i, j, k + n[2])];
            for (int64_t k = n[2] + 1; k <= f.ub[2]; ++k) dst[f.at(i, j, k)] = dst[f.at(i, j, k - n[2])];
        }
}

int ensure(DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return TBFG_OK;
    return b.alloc(bytes);
}

}  // namespace

extern "C" {

const char* tbfg_last_error(void) { return tbfg::last_error(); }

const char* tbfg_version(void) { return "tbfgpu 0.1 (gfx950)"; }

int tbfg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ------------------------------------------------------------------ FAST

int tbfg_fps_create(const tbfg_grid_desc* grid, const tbfg_fps_opts* opts, tbfg_fps** out)
{
    ABI_GUARD(return fps_create(grid, opts, out);)
}

int tbfg_fps_solve_dev(tbfg_fps* h, const double* s_dev, double* p_dev, void* stream)
{
    ABI_GUARD(return fps_solve_dev(h, s_dev, p_dev, stream);)
}

int tbfg_fps_solve(tbfg_fps* h, const double* s, const int64_t s_lb[3], const int64_t s_ub[3], double* p,
                   const int64_t p_lb[3], const int64_t p_ub[3])
{
    ABI_GUARD({
        if (!h || !s || !p) {
            set_error("tbfg_fps_solve: NULL argument");
            return TBFG_EINVAL;
        }
        const int32_t* n = fps_grid(h)->n;
        FArr fs, fp;
        TBFG_TRY(make_farr(s_lb, s_ub, n, "s", fs));
        TBFG_TRY(make_farr(p_lb, p_ub, n, "p", fp));
        const size_t cells = (size_t)n[0] * n[1] * n[2];
        std::vector<double> stage(cells);
        DevBuf ds, dp;
        TBFG_TRY(ds.alloc(cells * 8));
        TBFG_TRY(dp.alloc(cells * 8));
        gather(s, fs, n, stage.data());
        TBFG_HIP(hipMemcpy(ds.p, stage.data(), cells * 8, hipMemcpyHostToDevice));
        TBFG_TRY(fps_solve_dev(h, ds.as<double>(), dp.as<double>(), nullptr));
        TBFG_HIP(hipMemcpy(stage.data(), dp.p, cells * 8, hipMemcpyDeviceToHost));
        scatter_fill(stage.data(), fp, n, p);
        return TBFG_OK;
    })
}

int tbfg_fps_apply_operator_dev(tbfg_fps* h, const double* p_dev, double* q_dev, void* stream)
{
    ABI_GUARD({
        if (!h || !p_dev || !q_dev) {
            set_error("tbfg_fps_apply_operator_dev: NULL argument");
            return TBFG_EINVAL;
        }
        return fps_apply_operator_dev(h, p_dev, q_dev, stream);
    })
}

int64_t tbfg_fps_footprint(const tbfg_fps* h) { return fps_footprint(h); }

int tbfg_fps_destroy(tbfg_fps* h)
{
    ABI_GUARD({
        fps_destroy(h);
        return TBFG_OK;
    })
}

// -------------------------------------------------------------------- MG

void tbfg_pcg_default_params(tbfg_pcg_params* p)
{
    if (!p) return;
    // channel_18/specs/pcg_solver:21-31 with one V-cycle per application
    p->n_levels = 4;
    p->n_pre = 2;
    p->n_post = 2;
    p->max_iter = 100;
    p->tol = 1e-8;
    p->mg_cycles = 1;
    p->precision = TBFG_PREC_FP64;
    p->coarse_max_iter = -1;
    p->flexible = -1;
    p->coarse_tol = 0.0;
    p->preconditioner = TBFG_PRECOND_MG;
    p->mg_max_iter = 100;
    p->mg_tol = 0.0;
}

int tbfg_pcg_create(const tbfg_grid_desc* grid, const tbfg_pcg_params* prm, tbfg_pcg** out)
{
    ABI_GUARD(return pcg_create(grid, prm, out);)
}

int tbfg_pcg_solve_dev(tbfg_pcg* h, double* psi_dev, const double* beta_dev, const double* q_dev, int32_t* iters,
                       double* rel_res, void* stream)
{
    ABI_GUARD(return pcg_solve_dev(h, psi_dev, beta_dev, q_dev, iters, rel_res, stream);)
}

int tbfg_pcg_solve_queued_dev(tbfg_pcg* h, double* psi_dev, const double* beta_dev, const double* q_dev,
                              int32_t iterations, void* stream)
{
    ABI_GUARD(return pcg_solve_queued_dev(h, psi_dev, beta_dev, q_dev, iterations, stream);)
}

int tbfg_pcg_result(tbfg_pcg* h, int32_t* iters, double* res)
{
    ABI_GUARD(return pcg_result(h, iters, res);)
}

int tbfg_pcg_precondition_dev(tbfg_pcg* h, const double* beta_dev, const double* r_dev, double* z_dev, void* stream)
{
    ABI_GUARD({
        TBFG_TRY(pcg_precondition_dev(h, beta_dev, r_dev, z_dev, stream));
        TBFG_HIP(hipStreamSynchronize((hipStream_t)stream));
        return TBFG_OK;
    })
}

int tbfg_pcg_solve(tbfg_pcg* h, double* psi, const int64_t psi_lb[3], const int64_t psi_ub[3], const double* beta,
                   const int64_t beta_lb[3], const int64_t beta_ub[3], const double* q, const int64_t q_lb[3],
                   const int64_t q_ub[3], int32_t* iters, double* rel_res)
{
    ABI_GUARD({
        if (!h || !psi || !beta || !q) {
            set_error("tbfg_pcg_solve: NULL argument");
            return TBFG_EINVAL;
        }
        const int32_t* n = h->g.n;
        FArr fpsi, fb, fq;
        TBFG_TRY(make_farr(psi_lb, psi_ub, n, "psi", fpsi));
        TBFG_TRY(make_farr(beta_lb, beta_ub, n, "beta", fb));
        TBFG_TRY(make_farr(q_lb, q_ub, n, "q", fq));
        const size_t cells = h->n0;
        if (!h->hstage) TBFG_HIP(hipHostMalloc((void**)&h->hstage, 3 * cells * 8, hipHostMallocDefault));
        TBFG_TRY(ensure(h->hpsi, cells * 8));
        TBFG_TRY(ensure(h->hbeta, cells * 8));
        TBFG_TRY(ensure(h->hq, cells * 8));
        double* s0 = h->hstage;
        double* s1 = s0 + cells;
        double* s2 = s1 + cells;
        gather(psi, fpsi, n, s0);
        gather(beta, fb, n, s1);
        gather(q, fq, n, s2);
        TBFG_HIP(hipMemcpy(h->hpsi.p, s0, cells * 8, hipMemcpyHostToDevice));
        TBFG_HIP(hipMemcpy(h->hbeta.p, s1, cells * 8, hipMemcpyHostToDevice));
        TBFG_HIP(hipMemcpy(h->hq.p, s2, cells * 8, hipMemcpyHostToDevice));
        int rc = pcg_solve_dev(h, h->hpsi.as<double>(), h->hbeta.as<double>(), h->hq.as<double>(), iters, rel_res,
                               nullptr);
        if (rc != TBFG_OK && rc != TBFG_ENOTCONVERGED) return rc;
        TBFG_HIP(hipMemcpy(s0, h->hpsi.p, cells * 8, hipMemcpyDeviceToHost));
        scatter_fill(s0, fpsi, n, psi);
        return rc;
    })
}

int64_t tbfg_pcg_footprint(const tbfg_pcg* h) { return h ? h->footprint : 0; }

int tbfg_pcg_profile_begin(tbfg_pcg* h, int32_t kernel) { ABI_GUARD(return pcg_profile_begin(h, kernel);) }

int tbfg_pcg_coarse_stats(const tbfg_pcg* h, int32_t* calls, int32_t* iters, int32_t* iters_max, double* res_max)
{
    if (!h || !h->hst) {
        set_error("tbfg_pcg_coarse_stats: no solve yet");
        return TBFG_EINVAL;
    }
    if (calls) *calls = h->hst->coarse_calls;
    if (iters) *iters = h->hst->coarse_iters;
    if (iters_max) *iters_max = h->hst->coarse_iters_max;
    if (res_max) *res_max = h->hst->coarse_res_max;
    return TBFG_OK;
}

int tbfg_pcg_mg_stats(const tbfg_pcg* h, int32_t* cycles_total, int32_t* cycles_max)
{
    if (!h || !h->hst) {
        set_error("tbfg_pcg_mg_stats: no solve yet");
        return TBFG_EINVAL;
    }
    if (cycles_total) *cycles_total = h->hst->mg_cyc_total;
    if (cycles_max) *cycles_max = h->hst->mg_cyc_max;
    return TBFG_OK;
}

int tbfg_pcg_profile_end(tbfg_pcg* h, int64_t* launches, double* total_ms, double* bytes_per_launch)
{
    ABI_GUARD(return pcg_profile_end(h, launches, total_ms, bytes_per_launch);)
}

int tbfg_pcg_destroy(tbfg_pcg* h)
{
    ABI_GUARD({
        delete h;
        return TBFG_OK;
    })
}

// ---------------------------------------------------------------- fields

int tbfg_apply_operator_dev(const tbfg_grid_desc* grid, const double* beta_dev, const double* p_dev, double* q_dev,
                            void* stream)
{
    ABI_GUARD({
        TBFG_TRY(validate_grid(grid));
        if (!beta_dev || !p_dev || !q_dev) {
            set_error("tbfg_apply_operator_dev: NULL argument");
            return TBFG_EINVAL;
        }
        HostLevel L = make_level(grid, 0);
        LevelTables<double> tab;
        TBFG_TRY(tab.build(L, nullptr));
        TBFG_TRY(launch_apply_op(tab.view, beta_dev, p_dev, q_dev, stream));
        // the tables are freed on return: wait for the kernel that reads them
        TBFG_HIP(hipStreamSynchronize((hipStream_t)stream));
        return TBFG_OK;
    })
}

int tbfg_divergence_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy, const double* uz, double* div,
                        void* stream)
{
    ABI_GUARD(return divergence_dev(grid, ux, uy, uz, div, stream);)
}

int tbfg_beta_dev(const tbfg_grid_desc* grid, double alpha_dt, const double* rho, double* beta, void* stream)
{
    ABI_GUARD(return beta_dev(grid, alpha_dt, rho, beta, stream);)
}

int tbfg_project_mg_dev(const tbfg_grid_desc* grid, double alpha_dt, const double* rho, const double* psi,
                        double* ux, double* uy, double* uz, void* stream)
{
    ABI_GUARD(return project_mg_dev(grid, alpha_dt, rho, psi, ux, uy, uz, stream);)
}

int tbfg_source_fast_dev(const tbfg_grid_desc* grid, double alpha_dt, double rho0, const double* rho,
                         const double* ux, const double* uy, const double* uz, const double* psi1,
                         const double* psi2, const double* const st1[3], const double* const st2[3], double* s,
                         void* stream)
{
    ABI_GUARD(return source_fast_dev(grid, alpha_dt, rho0, rho, ux, uy, uz, psi1, psi2, st1, st2, s, stream);)
}

int tbfg_project_fast_dev(const tbfg_grid_desc* grid, double alpha_dt, double rho0, const double* rho,
                          const double* psi, const double* psi1, const double* psi2, const double* const st1[3],
                          const double* const st2[3], double* ux, double* uy, double* uz, void* stream)
{
    ABI_GUARD(return project_fast_dev(grid, alpha_dt, rho0, rho, psi, psi1, psi2, st1, st2, ux, uy, uz, stream);)
}

int tbfg_update_pressure_dev(const tbfg_grid_desc* grid, const double* psi, double* p, double* psi1, double* psi2,
                             const double* const st[3], double* const st1[3], double* const st2[3], void* stream)
{
    ABI_GUARD(return update_pressure_dev(grid, psi, p, psi1, psi2, st, st1, st2, stream);)
}

int tbfg_max_divergence_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy, const double* uz,
                            double* out, void* stream)
{
    ABI_GUARD(return max_divergence_dev(grid, ux, uy, uz, out, stream);)
}

int tbfg_max_divergence_noalloc_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy,
                                    const double* uz, double* out, void* stream)
{
    ABI_GUARD(return max_divergence_noalloc_dev(grid, ux, uy, uz, out, stream);)
}

int64_t tbfg_material_work_doubles(const tbfg_grid_desc* grid)
{
    return material_work_doubles(grid);
}

int tbfg_material_props_dev(const tbfg_grid_desc* grid, double rho_l, double rho_g, double mu_l, double mu_g,
                            const double* c, double* cs, double* rho, double* mu, double* work, void* stream)
{
    ABI_GUARD(return material_props_dev(grid, rho_l, rho_g, mu_l, mu_g, c, cs, rho, mu, work, stream);)
}

// ---------------------------------------------------------------- inputs

int tbfg_make_velocity(const tbfg_grid_desc* grid, uint64_t seed, double* ux, double* uy, double* uz)
{
    ABI_GUARD(return make_velocity(grid, seed, ux, uy, uz);)
}

int tbfg_make_bubble_density(const tbfg_grid_desc* grid, int32_t nb, double diameter, uint64_t seed, int32_t nref,
                             double rho_l, double rho_g, double* rho, double* centres)
{
    ABI_GUARD(return make_bubble_density(grid, nb, diameter, seed, nref, rho_l, rho_g, rho, centres);)
}

}  // extern "C"
