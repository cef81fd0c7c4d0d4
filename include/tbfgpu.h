/*
 * tbfgpu.h -- C ABI of the MI355X pressure-projection engine for TBFsolver.
 *
 * A drop-in for the reference's two pressure-solver plug-ins, selected in the
 * reference at compile time by POIS_OPT (src/Makefile:3,17-24):
 *   FAST_MODE: type(fastPoissonSolver) / fastPoissonSolverCTOR / solveFPS
 *   MG_MODE:   type(pcg) / pcgCTOR / solvePCG (multigrid-preconditioned CG)
 * plus the device-resident pieces around them (RHS, coefficient, projection,
 * continuity error) and the input mirrors used by the benchmarks.
 *
 * Conventions
 *  - Fields are FP64.  Host-array entry points take Fortran arrays with their
 *    lower/upper bounds (lb/ub per axis, as lbound/ubound return them); only
 *    the interior 1..n is read, and outputs get their halo filled the way the
 *    reference's updateBoundaries does for the channel (periodic x and z,
 *    homogeneous Neumann in y).
 *  - Device entry points (*_dev) take device pointers to interior-only arrays,
 *    Fortran order (i fastest), and a HIP stream handle passed as void*
 *    (NULL = default stream).  They are asynchronous unless stated otherwise.
 *  - Every entry point returns a status code; tbfg_last_error() gives the
 *    message of the calling thread's last failure.  Nothing aborts.
 *  - A handle may be used by one host thread at a time; distinct handles are
 *    independent.
 */
#ifndef TBFGPU_H
#define TBFGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference aborts through mpiABORT, errorHandler.f90:30-50) */
#define TBFG_OK 0
#define TBFG_EINVAL 1        /* bad arguments, grid or hierarchy              */
#define TBFG_ENOMEM 2        /* device or host allocation failed             */
#define TBFG_ENOTCONVERGED 3 /* max iterations reached; result still written */
#define TBFG_ENUMERIC 4      /* non-finite values in input or iteration      */
#define TBFG_EDEVICE 5       /* HIP runtime error                            */
#define TBFG_EIO 6           /* file missing, short or not writable          */

#define TBFG_PREC_FP64 0  /* all arithmetic and storage in FP64                  */
#define TBFG_PREC_MIXED 1 /* FP32 inner solver inside an FP64 outer iteration    */
/* PCG only: TBFG_PREC_MIXED with the finest level's multigrid right-hand side
 * and coefficients stored in bf16 (the smoother's iterate, the arithmetic and
 * every coarser level stay FP32; the outer iteration and its stop test stay
 * FP64).  Where the finest level is not smoothed by the plane-march kernel
 * (too small, or a shape it does not tile) it acts as TBFG_PREC_MIXED. */
#define TBFG_PREC_MIXED_BF16 2

/* Grid: x and z periodic, y walls.  stretch[a] = 1 selects the hyperbolic
 * tangent distribution of grid.f90:293-330 with factor gamma (the reference
 * fixes gamma = 1.4 at grid.f90:27). */
typedef struct tbfg_grid_desc {
    int32_t n[3];
    double L[3];
    int32_t stretch[3];
    double gamma;
} tbfg_grid_desc;

/* ---------------------------------------------------------------- library */

/* Message of the calling thread's last failed call ("" if none). */
const char* tbfg_last_error(void);
/* Library version string. */
const char* tbfg_version(void);
/* Number of visible devices (0 on a host without a GPU; no device is initialised
 * by the host-only entry points below). */
int tbfg_device_count(void);

/* ---------------------------------------------------------- FFT solver (FAST) */

typedef struct tbfg_fps tbfg_fps;

typedef struct tbfg_fps_opts {
    int32_t precision;     /* TBFG_PREC_FP64 or TBFG_PREC_MIXED (FFT solver)          */
    int32_t refine_sweeps; /* MIXED only: FP64 iterative-refinement sweeps (default 2) */
} tbfg_fps_opts;

/* Replaces fastPoissonSolverCTOR (src/fastPoissonSolver/fastPoissonSolver.f90:123-156):
 * eigenvalues, y coefficients and device work buffers.  x and z must be uniform,
 * nx and nz in {2^a 3^b}, even, at most 4096 (beyond 1024 points only lengths whose every FFT pass
 * has at most 1024 butterflies in FP64, 2048 in mixed precision: n / r <= that for each radix r of
 * n = 4^a 2^b 3^c, so every length up to 2048, and 3072 and 4096), ny <= 1024. */
int tbfg_fps_create(const tbfg_grid_desc* grid, const tbfg_fps_opts* opts, tbfg_fps** out);

/* Replaces solveFPS (fastPoissonSolver.f90:337-384): solves lap(p) = s with the
 * FFT-x/z + tridiagonal-y operator, removes the volume-weighted mean of p
 * (:388-423) and fills p's halo (updateBoundaries).  s and p are host arrays
 * with Fortran bounds s_lb..s_ub, p_lb..p_ub (interior 1..n inside). Blocking. */
int tbfg_fps_solve(tbfg_fps* h, const double* s, const int64_t s_lb[3], const int64_t s_ub[3],
                   double* p, const int64_t p_lb[3], const int64_t p_ub[3]);

/* Device-resident variant: s_dev, p_dev interior-only (nx*ny*nz doubles). */
int tbfg_fps_solve_dev(tbfg_fps* h, const double* s_dev, double* p_dev, void* stream);

/* The constant-coefficient operator the FFT solver inverts (modified wavenumbers
 * in x and z, the 3-point y operator of fastPoissonSolver.f90:255-276): q = L p. */
int tbfg_fps_apply_operator_dev(tbfg_fps* h, const double* p_dev, double* q_dev, void* stream);

/* Device bytes held by the handle. */
int64_t tbfg_fps_footprint(const tbfg_fps* h);

int tbfg_fps_destroy(tbfg_fps* h);

/* -------------------------------------------------------- PCG + multigrid (MG) */

typedef struct tbfg_pcg tbfg_pcg;

typedef struct tbfg_pcg_params {
    int32_t n_levels;      /* nLevels     (channel_18/specs/pcg_solver:21)           */
    int32_t n_pre;         /* nPreSweep                                              */
    int32_t n_post;        /* nPostSweep                                             */
    int32_t max_iter;      /* maxIterPCG                                             */
    double tol;            /* tolPCG on ||r||_2 / ||q||_2                             */
    int32_t mg_cycles;     /* V-cycles per preconditioner application (default 1)    */
    int32_t precision;     /* TBFG_PREC_FP64 (reference algorithm), TBFG_PREC_MIXED or
                              TBFG_PREC_MIXED_BF16                                   */
    int32_t coarse_max_iter; /* coarsest reference level: > 0 inner CG iterations to
                                coarse_tol, 0 the V-cycle continues below it (one
                                cycle), -1 by precision (FP64 40, mixed 0)           */
    int32_t flexible;        /* 0 Fletcher-Reeves, 1 Polak-Ribiere, -1 by precision  */
    double coarse_tol;       /* relative residual of the coarsest-level solve
                                (0: 1e-10 in FP64, 1e-5 in mixed precision)          */
    int32_t preconditioner;  /* TBFG_PRECOND_MG (the reference's multigrid) or
                                TBFG_PRECOND_FFT: z = L^-1 r with L the constant-
                                coefficient operator of the FFT solver (FP32 FFTs in
                                mixed precision); x, z uniform, nx, nz 2^a 3^b      */
    int32_t mg_max_iter;     /* maxIterMG: most V-cycles per application (mg_tol > 0) */
    double mg_tol;           /* tolMG (multiGrid.f90:165-182, 236-279): > 0 each
                                application runs V-cycles until the finest level's
                                ||q - A p|| / ||q|| after its last smoothing is at
                                most mg_tol (at least one cycle, at most
                                mg_max_iter; channel_18/specs/pcg_solver: 0.4, 100);
                                0: mg_cycles cycles.  Must be < 1.                  */
} tbfg_pcg_params;

#define TBFG_PRECOND_MG 0
#define TBFG_PRECOND_FFT 1

/* Fills *p with the reference's defaults for the channel case. */
void tbfg_pcg_default_params(tbfg_pcg_params* p);

/* Replaces pcgCTOR + multiGridCTOR (src/pcg/pcg.f90:62-99, src/multiGrid/multiGrid.f90:72-112):
 * builds the level hierarchy and allocates the device work space.  Every axis
 * must be divisible by 2^(n_levels-1) (multiGrid.f90:116-145). */
int tbfg_pcg_create(const tbfg_grid_desc* grid, const tbfg_pcg_params* prm, tbfg_pcg** out);

/* Replaces solvePCG (pcg.f90:103-140): psi (in: warm start, out: solution with halo
 * filled), beta = alpha*dt/rho and q = div(u*) are host Fortran arrays.  *iters
 * receives the number of direction updates (the reference's iter_ is iters+1 on
 * convergence, pcg.f90:401), *rel_res the final ||r||/||q||.  Returns
 * TBFG_ENOTCONVERGED (with psi written) when max_iter is reached.  Blocking. */
int tbfg_pcg_solve(tbfg_pcg* h, double* psi, const int64_t psi_lb[3], const int64_t psi_ub[3],
                   const double* beta, const int64_t beta_lb[3], const int64_t beta_ub[3],
                   const double* q, const int64_t q_lb[3], const int64_t q_ub[3], int32_t* iters,
                   double* rel_res);

/* Device-resident variant (interior-only arrays).  Synchronises the stream to
 * test convergence. */
int tbfg_pcg_solve_dev(tbfg_pcg* h, double* psi_dev, const double* beta_dev, const double* q_dev,
                       int32_t* iters, double* rel_res, void* stream);

/* No counterpart in the reference: solvePCG (pcg.f90:103-140) enqueued without any
 * host synchronisation, for callers that capture a whole batch of solves in one HIP
 * graph.  Exactly `iterations` PCG iterations are queued (1..max_iter); the device's
 * stop test still ends the solve, and iterations queued after it return at once, so
 * the result equals tbfg_pcg_solve_dev's when `iterations` is at least the count that
 * solve needs.  Asynchronous; the stream carries only kernels.  Not for a handle whose
 * solve must synchronise (tolMG loop, coarsest level larger than the tail kernel takes,
 * no pre-smoothing, kernel profiling): TBFG_EINVAL. */
int tbfg_pcg_solve_queued_dev(tbfg_pcg* h, double* psi_dev, const double* beta_dev, const double* q_dev,
                              int32_t iterations, void* stream);
/* The outcome of the last queued solve, read after its work has finished: iterations,
 * relative residual, and TBFG_OK, TBFG_ENOTCONVERGED (max_iter reached, or not converged
 * within the iterations queued) or TBFG_ENUMERIC.  Blocking.  The read is ordered with no
 * stream (a library-internal non-blocking one, legal while other threads capture graphs):
 * the caller synchronises the solve's stream first. */
int tbfg_pcg_result(tbfg_pcg* h, int32_t* iters, double* res);

/* Applies the handle's multigrid preconditioner once: z = M r (device arrays). */
int tbfg_pcg_precondition_dev(tbfg_pcg* h, const double* beta_dev, const double* r_dev, double* z_dev,
                              void* stream);

/* Device bytes held by the handle (for admission by HBM footprint). */
int64_t tbfg_pcg_footprint(const tbfg_pcg* h);

/* Kernel timing with HIP events recorded on the solve's stream around every
 * launch of one kernel family (no counterpart in the reference, whose only
 * timing is the wall clock of poissonEqn.f90:463-489):
 *   0 red/black sweep on the finest level         1 fused residual+restriction, finest
 *     (one full sweep, or its two half-sweeps)
 *   2 operator apply + d.Ad (PCG)                 3 psi/r update + ||r|| (PCG)
 * tbfg_pcg_profile_end returns the launches seen since _begin, their summed
 * duration and the algorithmic bytes one launch must move (each input value
 * read once, each output written once).  Launches queued ahead of a solve's
 * stop test that found it already stopped (shorter than a quarter of that
 * solve's median launch) are not counted.  Families 0 and 2 (plane-marching
 * forms) are timed by events the dispatch itself records, so the interval
 * holds that kernel alone. */
#define TBFG_KERNEL_SWEEP_FINE 0
#define TBFG_KERNEL_RESTRICT_FINE 1
#define TBFG_KERNEL_SPMV 2
#define TBFG_KERNEL_UPDATE 3
int tbfg_pcg_profile_begin(tbfg_pcg* h, int32_t kernel);

/* Diagnostics of the handle's last solve: coarsest-level solves made, their
 * inner iterations (sum and maximum) and the worst final relative residual. */
int tbfg_pcg_coarse_stats(const tbfg_pcg* h, int32_t* calls, int32_t* iters, int32_t* iters_max, double* res_max);
/* V-cycles of the last solve's preconditioner applications: the sum and the
 * largest per application (the reference's multiGrid iter_, multiGrid.f90:236-279).
 * A solve of n updates applies the preconditioner n times (the first, then
 * after every update but the last, whose stop test comes first); the
 * reference applies it once more, after the final update, before its test
 * (pcg.f90:122-136).  After tbfg_pcg_precondition_dev with mg_tol > 0: that
 * application's cycles. */
int tbfg_pcg_mg_stats(const tbfg_pcg* h, int32_t* cycles_total, int32_t* cycles_max);
int tbfg_pcg_profile_end(tbfg_pcg* h, int64_t* launches, double* total_ms, double* bytes_per_launch);

int tbfg_pcg_destroy(tbfg_pcg* h);

/* ------------------------------------------------------- stencils and fields */

/* computeQv (pcg.f90:144-183): q = A p with A = div(beta grad .), channel ghosts. */
int tbfg_apply_operator_dev(const tbfg_grid_desc* grid, const double* beta_dev, const double* p_dev,
                            double* q_dev, void* stream);

/* MG computeSource (poissonEqn.f90:155-190) / auxiliaryRoutines.f90:66-79:
 * div = du/dx + dv/dy + dw/dz.  ux,uz: nx*ny*nz faces (face i between cells i
 * and i+1, periodic); uy: nx*(ny+1)*nz faces, j = 0..ny (walls at 0 and ny). */
int tbfg_divergence_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy, const double* uz,
                        double* div, void* stream);

/* computeBeta (poissonEqn.f90:311-342): beta = alpha_dt / rho. */
int tbfg_beta_dev(const tbfg_grid_desc* grid, double alpha_dt, const double* rho, double* beta,
                  void* stream);

/* MG makeVelocityDivFree -> addPressureGrad (momentumEqn.f90:795-869,980-993). */
int tbfg_project_mg_dev(const tbfg_grid_desc* grid, double alpha_dt, const double* rho, const double* psi,
                        double* ux, double* uy, double* uz, void* stream);

/* FAST_MODE computeSource + computeOldPressDiv (poissonEqn.f90:193-305), nl = 2:
 * s = rho0/(alpha dt) div u* + div[(1 - rho0 (1/rho)_f) G_f] with the face value
 * G_f = d_f(2 psi1 - psi2) - 2 st1_f + st2_f (poissonEqn.f90:255-301), psi1 =
 * psi^n, psi2 = psi^(n-1) (psi%ptrf_, psi%ptrf_%ptrf_), and st1 = st^n, st2 =
 * st^(n-1) the old surface-tension levels (st%ptrf_, st%ptrf_%ptrf_), and (1/rho)_f
 * the face mean of 1/rho.  st1 and st2 point to three device arrays each (x, y, z
 * components, laid out like ux, uy, uz; the y wall faces are read as the reference
 * reads them, and updatePressure keeps them zero).  Any level, or component, may
 * be NULL, meaning zero (the first steps; no surface tension).
 * rho0 = min(rho_l, rho_g) (poissonEqn.f90:127).  x, z uniform. */
int tbfg_source_fast_dev(const tbfg_grid_desc* grid, double alpha_dt, double rho0, const double* rho,
                         const double* ux, const double* uy, const double* uz, const double* psi1,
                         const double* psi2, const double* const st1[3], const double* const st2[3], double* s,
                         void* stream);

/* FAST_MODE makeVelocityDivFree + computeOldPressGrad (momentumEqn.f90:999-1127):
 * u_f -= alpha dt [(1/rho0) d psi + ((1/rho)_f - 1/rho0) G_f], G_f as above
 * (momentumEqn.f90:1104-1121), on the periodic x/z faces and interior y faces. */
int tbfg_project_fast_dev(const tbfg_grid_desc* grid, double alpha_dt, double rho0, const double* rho,
                          const double* psi, const double* psi1, const double* psi2, const double* const st1[3],
                          const double* const st2[3], double* ux, double* uy, double* uz, void* stream);

/* updatePressure (poissonEqn.f90:374-415) with clean_st_prev_bc (:419-459), one
 * kernel: p = psi (p may be NULL); storeOldField: psi2 <- psi1 <- psi;
 * storeOldFieldV: st2 <- st1 <- st (st NULL = zero surface tension); then the old
 * st levels are zeroed on the Neumann walls, i.e. the y faces j = 0 and j = ny
 * (x and z are periodic).  st1 and st2 both NULL skip the surface-tension levels
 * (st must then be NULL too).  The reference's restart copies p0 and st0 are
 * psi2 and st2 after the call. */
int tbfg_update_pressure_dev(const tbfg_grid_desc* grid, const double* psi, double* p, double* psi1, double* psi2,
                             const double* const st[3], double* const st1[3], double* const st2[3], void* stream);

/* computeContinuityError (auxiliaryRoutines.f90:42-91): returns max|div u| (not
 * multiplied by dt) in *out.  Blocking. */
int tbfg_max_divergence_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy, const double* uz,
                            double* out, void* stream);

/* No counterpart in the reference: tbfg_max_divergence_dev for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  The
 * same result, bit for bit, without a hipMalloc or hipFree: the max word and the grid
 * metrics live in a static device array of 196,864 bytes per GPU, held until the
 * process ends, so every axis may have at most 4096 cells (more: TBFG_EINVAL).  Every
 * call on a GPU shares that array; calls are serialised per GPU and each returns with
 * its stream drained, so calls from several threads do not overlap (overlapping use of
 * the array would race).  Blocking; *out is a host double. */
int tbfg_max_divergence_noalloc_dev(const tbfg_grid_desc* grid, const double* ux, const double* uy,
                                    const double* uz, double* out, void* stream);

/* updateMaterialProps + smoothVF (VOF.f90:197-262): two passes of cellToVertex /
 * vertexToCell (interpolation.f90:188-276) smooth the volume fraction c into cs
 * (ghosts: periodic x, z; the wall row in y), then rho = rho_g cs + (1-cs) rho_l
 * and mu = mu_g cs + (1-cs) mu_l (mu may be NULL).  Device arrays of nx*ny*nz
 * cells; work: tbfg_material_work_doubles(grid) doubles of device scratch
 * (the vertex field).  c and cs must not alias. */
int64_t tbfg_material_work_doubles(const tbfg_grid_desc* grid);
int tbfg_material_props_dev(const tbfg_grid_desc* grid, double rho_l, double rho_g, double mu_l, double mu_g,
                            const double* c, double* cs, double* rho, double* mu, double* work, void* stream);

/* ------------------------------------------------------------------- tuning table */

/* No counterpart in the reference (its kernels are OpenMP loops).  The tile
 * shape a kernel family launches with on a level of dims[3] cells in a
 * precision of elem_bytes (4: the FP32 multigrid, 8: FP64), chosen by a tuner
 * (tbfsolver_amd/tune.py) instead of the built-in preference order:
 *   TBFG_TUNE_SWEEP       full sweep from a non-zero start   shape = {tx, ty, threads}
 *   TBFG_TUNE_SWEEP_EDGE  first sweep / sweep with the dots  shape = {tx, ty, threads}
 *   TBFG_TUNE_OP          FP64 operator product (level 0)    shape = {tx, threads, 0}, {-1, 0, 0} = columns
 * A shape the level does not divide, or without a kernel, is ignored.  Sweeps
 * read the table at every launch, the operator when a handle is created.
 * shape {0, 0, 0} removes the entry; tbfg_tune_get returns {0, 0, 0} for none. */
#define TBFG_TUNE_SWEEP 0
#define TBFG_TUNE_SWEEP_EDGE 1
#define TBFG_TUNE_OP 2
int tbfg_tune_set(int32_t kind, const int32_t dims[3], int32_t elem_bytes, const int32_t shape[3]);
int tbfg_tune_get(int32_t kind, const int32_t dims[3], int32_t elem_bytes, int32_t shape[3]);
int tbfg_tune_clear(void);

/* ------------------------------------------------------------- field files (host) */

/* Field types of the reference (setIndexesBounds, field.f90:1746-1789): cell
 * centres, and faces staggered in x, y, z (file range 0..n on that axis). */
#define TBFG_FIELD_CL 0
#define TBFG_FIELD_SX 1
#define TBFG_FIELD_SY 2
#define TBFG_FIELD_SZ 3

/* Number of values a field file of this type holds for an n[0] x n[1] x n[2] grid. */
int tbfg_field_count(int type, const int32_t n[3], int64_t* count);

/* writeField (field.f90:453-476): the file holds the int32 count, then the FP64
 * values of the type's index range, column-major (x fastest).  f is a host array
 * in the engine layout: nx*ny*nz values (cl; sx and sz hold faces 1..n, whose
 * periodic image is written as face 0), nx*(ny+1)*nz for sy (faces 0..ny). */
int tbfg_field_write(const char* path, int type, const int32_t n[3], const double* f);

/* readField / readInternalField (field.f90:331-346,419-443): reads such a file
 * into the engine layout; a count other than the grid's is TBFG_EINVAL (the
 * reference aborts: "Dimension of field ... not equal to mesh size"). */
int tbfg_field_read(const char* path, int type, const int32_t n[3], double* f);

/* The name of time folder n (writeTimeFolder, time.f90:317: format I6, adjusted
 * left): "0", "12", ...  out receives at most len bytes including the NUL. */
int tbfg_time_folder_name(int32_t n, char* out, int32_t len);

/* --------------------------------------------------------- input mirrors (host) */

/* Seeded staggered velocity, u,v,w ~ N(0,1) on unique faces, v = 0 on the walls
 * (counter-based generator: splitmix64 -> uniform -> Box-Muller). Host arrays. */
int tbfg_make_velocity(const tbfg_grid_desc* grid, uint64_t seed, double* ux, double* uy, double* uz);

/* Non-overlapping spheres (periodic x/z distance), centres uniform in
 * [0,Lx) x [R, Ly-R] x [0,Lz) by rejection; VOF per init_Bubble_vf
 * (initialConditions.f90:192-398) with nref^3 sub-sampling of interface cells,
 * two cell->vertex->cell smoothing passes (VOF.f90:235-262) and
 * rho = rho_g*cs + rho_l*(1-cs) (VOF.f90:222-223).  centres (3*nb doubles) may be
 * NULL.  Host arrays. */
int tbfg_make_bubble_density(const tbfg_grid_desc* grid, int32_t nb, double diameter, uint64_t seed,
                             int32_t nref, double rho_l, double rho_g, double* rho, double* centres);

#ifdef __cplusplus
}
#endif
#endif
