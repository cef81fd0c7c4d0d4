// fields.hpp -- RHS, coefficient, projection and continuity-error kernels.
#pragma once
#include <memory>

#include "common.hpp"

namespace tbfg {
int divergence_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz, double* div,
                   void* stream);
int max_divergence_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz, double* out,
                       void* stream);
int max_divergence_noalloc_dev(const tbfg_grid_desc* g, const double* ux, const double* uy, const double* uz,
                               double* out, void* stream);
int beta_dev(const tbfg_grid_desc* g, double alpha_dt, const double* rho, double* beta, void* stream);
int source_fast_dev(const tbfg_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* ux,
                    const double* uy, const double* uz, const double* psi1, const double* psi2,
                    const double* const st1[3], const double* const st2[3], double* s, void* stream);
int project_fast_dev(const tbfg_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* psi,
                     const double* psi1, const double* psi2, const double* const st1[3],
                     const double* const st2[3], double* ux, double* uy, double* uz, void* stream);
int update_pressure_dev(const tbfg_grid_desc* g, const double* psi, double* p, double* psi1, double* psi2,
                        const double* const st[3], double* const st1[3], double* const st2[3], void* stream);
int project_mg_dev(const tbfg_grid_desc* g, double alpha_dt, const double* rho, const double* psi, double* ux,
                   double* uy, double* uz, void* stream);
int64_t material_work_doubles(const tbfg_grid_desc* g);
int material_props_dev(const tbfg_grid_desc* g, double rho_l, double rho_g, double mu_l, double mu_g, const double* c,
                       double* cs, double* rho, double* mu, double* work, void* stream);
}  // namespace tbfg
