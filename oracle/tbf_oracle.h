/*
 * tbf_oracle.h -- CPU FP64 restatement of TBFsolver's pressure-projection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the GPU library:
 * only the test suite, the smoke check and the CPU-baseline leg of the bench
 * may load it.  The product never links, loads or calls it.
 *
 * Conventions (shared with the GPU library's C ABI, restated independently):
 *  - cell-centred fields hold interior cells only, Fortran order, i (x) fastest:
 *    idx(i,j,k) = (i-1) + nx*((j-1) + ny*(k-1)) with 1-based i,j,k;
 *  - x and z are periodic, y is bounded by walls with homogeneous Neumann for the
 *    pressure (ghost = first interior cell), as in the channel set-up of the
 *    reference (src/fields/field/boundaryField_S.f90:713-725, field.f90:708-946);
 *  - staggered velocity: ux(i,j,k) is the x-face between cells i and i+1 for
 *    i=1..nx (face 0 == face nx, periodic), uz likewise in z; uy has ny+1 faces
 *    j=0..ny per column (stored j-index 0..ny), faces 0 and ny are the walls.
 */
#ifndef TBF_ORACLE_H
#define TBF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t n[3];       /* cells along x, y, z                      */
    double  L[3];       /* domain lengths                            */
    int32_t stretch[3]; /* 0 = uniform, 1 = tanh (grid.f90:293-330)  */
    double  gamma;      /* tanh stretching factor (reference: 1.4)   */
} or_grid_desc;

typedef struct {
    int32_t n_levels;      /* MG levels including the finest (multiGrid.f90:81)   */
    int32_t n_pre;         /* pre-smoothing RBGS sweeps  (nPreSweep)              */
    int32_t n_post;        /* post-smoothing RBGS sweeps (nPostSweep)             */
    int32_t max_iter;      /* maxIterPCG                                          */
    double  tol;           /* tolPCG on ||r||_2/||q||_2                           */
    int32_t mg_cycles;     /* fixed V-cycles per preconditioner application       */
    int32_t flexible;      /* 0: Fletcher-Reeves (pcg.f90:130-133), 1: Polak-Ribiere */
    int32_t coarse_mode;   /* 0: exact coarsest solve, 1: K inner cycles on a sub-hierarchy */
    int32_t coarse_cycles; /* K for coarse_mode 1                                 */
    int32_t mg_max_iter;   /* maxIterMG (used when mg_tol > 0)                    */
    double  mg_tol;        /* tolMG: > 0 runs solveMG's loop (V-cycles while the
                              finest residual exceeds it), 0 mg_cycles fixed      */
} or_pcg_params;

/* grid coordinates with `ghost` extra layers: xc has n+2*ghost entries for
 * i = 1-ghost..n+ghost, xf has n+2*ghost+1 entries for i = -ghost..n+ghost. */
void or_axis_coords(int32_t n, double L, int32_t stretch, double gamma,
                    int32_t ghost, double* xc, double* xf);

/* FFT building blocks (unnormalised, sign -1 forward, +1 backward). */
void or_fft_c2c(int32_t n, const double* in_ri, double* out_ri, int32_t sign);

/* 7-point variable-coefficient operator q = A p (pcg.f90:144-183). */
void or_apply_op(const or_grid_desc* g, const double* beta, const double* p, double* q);

/* FAST_MODE direct solve: lap(p) = s, p volume-weighted mean free (fastPoissonSolver.f90:337-423). */
int or_fft_solve(const or_grid_desc* g, const double* s, double* p);
/* the constant-coefficient operator the FFT solver inverts: modified wavenumbers in x,z, 3-point y. */
void or_fft_operator(const or_grid_desc* g, const double* p, double* q);

/* one tridiagonal column solve with lambda = lx+lz (both real), complex RHS interleaved. */
void or_tridiag(const or_grid_desc* g, double lam, int32_t pinned, double* rhs_ri);

/* MG building blocks on level `lev` (lev 0 = finest) of an n_levels hierarchy. */
void or_level_dims(const or_grid_desc* g, int32_t lev, int32_t* n3);
void or_rbgs(const or_grid_desc* g, int32_t lev, const double* beta, const double* q,
             double* p, int32_t sweeps, int32_t remove_mean);
void or_residual(const or_grid_desc* g, int32_t lev, const double* beta, const double* q,
                 const double* p, double* res);
void or_restrict(const or_grid_desc* g, int32_t lev_fine, const double* fine, double* coarse);
void or_prolong_add(const or_grid_desc* g, int32_t lev_fine, const double* coarse, double* fine);
int  or_coarse_exact(const or_grid_desc* g, int32_t lev, const double* beta, const double* q,
                     double* p);

/* MG_MODE: PCG + multigrid (pcg.f90:103-140).  psi in/out (warm start). */
int or_pcg_solve(const or_grid_desc* g, const or_pcg_params* prm, const double* beta,
                 const double* q, double* psi, int32_t* iters, double* rel_res);
/* one preconditioner application z = M r. */
/* V-cycles of the preconditioner applications of the last or_pcg_solve /
 * or_mg_apply call in this process: total, largest per application, and
 * those of the last application (the one after the final update, which the
 * reference makes before its stop test, pcg.f90:122-136). */
void or_last_mg_cycles(int32_t* total, int32_t* max_per_apply, int32_t* last);
int or_mg_apply(const or_grid_desc* g, const or_pcg_params* prm, const double* beta,
                const double* r, double* z);

/* RHS, coefficient, projection and divergence (poissonEqn.f90:155-342, momentumEqn.f90:795-1090,
 * auxiliaryRoutines.f90:42-91). */
void or_divergence(const or_grid_desc* g, const double* ux, const double* uy, const double* uz,
                   double* div);
void or_beta(const or_grid_desc* g, double alpha_dt, const double* rho, double* beta);
void or_project_mg(const or_grid_desc* g, double alpha_dt, const double* rho, const double* psi,
                   double* ux, double* uy, double* uz);
void or_project_fast(const or_grid_desc* g, double alpha_dt, double rho0, const double* rho,
                     const double* psi, const double* psi1, const double* psi2,
                     const double* const st1[3], const double* const st2[3],
                     double* ux, double* uy, double* uz);
void or_source_fast(const or_grid_desc* g, double alpha_dt, double rho0, const double* rho,
                    const double* ux, const double* uy, const double* uz,
                    const double* psi1, const double* psi2,
                    const double* const st1[3], const double* const st2[3], double* s);
void or_update_pressure(const or_grid_desc* g, const double* psi, double* p, double* psi1, double* psi2,
                        const double* const st[3], double* const st1[3], double* const st2[3]);
double or_max_abs(int64_t n, const double* a);
void or_material_props(const or_grid_desc* g, double rho_l, double rho_g, double mu_l, double mu_g,
                       const double* c, double* cs, double* rho, double* mu);
double or_volume_mean(const or_grid_desc* g, const double* p);

#ifdef __cplusplus
}
#endif
#endif
