/*
 * tbf_oracle.c -- CPU FP64 restatement of TBFsolver's pressure-projection path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by the test suite, the smoke
 * check and the CPU-baseline leg of the bench, never by the product.
 *
 * Each function names the reference routine it restates (file:line relative to
 * the reference source tree).  The restatement is written from the reference's
 * behaviour; the product's GPU code is an independent implementation.
 *
 * Deviations that do not change the converged answer are listed where they occur:
 *   - halos are not stored; ghost values are read through periodic wrap (x, z) and
 *     Neumann clamp (y), which is what updateBoundaries leaves in the halo for the
 *     channel set-up (field.f90:950-974 with boundaryField_S.f90:713-725);
 *   - the coarsest multigrid solve (MUMPS in poissMat.f90:289-312) is an exact
 *     solve of the same Lagrange-augmented system by conjugate gradients to
 *     round-off (the system's unique solution is the same);
 *   - FFTs are an own mixed-radix DFT (FFTW in fastPoissonSolver.f90:353-371);
 *     the singular (kx,kz)=(0,0) column pins x_ny = 0 instead of the epsilon
 *     patch at fastPoissonSolver.f90:318-322; both differ by a constant that the
 *     volume-weighted mean removal (:388-423) takes out.
 *
 * Parity status: every exported function is pinned by a test in tests/ (closed
 * forms, numpy DFTs, dense solves on tiny grids, invariants); see the test
 * file headers.
 */
#include "tbf_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OG 3 /* ghost layers kept in coordinate arrays */

typedef struct {
    int n;
    double* buf;
    double *c, *f, *dc, *df, *ip, *im; /* offset pointers, see ax_build */
} ax_t;

typedef struct {
    int n[3];
    ax_t ax[3];
} lev_t;

/* ------------------------------------------------------------------ grid */

/* grid.f90:235-330 -- uniform and tanh coordinates of centres and faces. */
static double coord_c(int i, int n, double L, int stretch, double gamma)
{
    if (!stretch) {
        double delta = L / n;
        return 0.5 * delta + (i - 1) * delta;
    }
    return 0.5 * L * (1.0 + tanh(-gamma + 2.0 * gamma * (i - 0.5) / n) / tanh(gamma));
}

static double coord_f(int i, int n, double L, int stretch, double gamma)
{
    if (!stretch) {
        double delta = L / n;
        return i * delta;
    }
    return 0.5 * L * (1.0 + tanh(-gamma + 2.0 * gamma * i / n) / tanh(gamma));
}

void or_axis_coords(int32_t n, double L, int32_t stretch, double gamma, int32_t ghost,
                    double* xc, double* xf)
{
    for (int i = 1 - ghost; i <= n + ghost; ++i) xc[i - (1 - ghost)] = coord_c(i, n, L, stretch, gamma);
    for (int i = -ghost; i <= n + ghost; ++i) xf[i + ghost] = coord_f(i, n, L, stretch, gamma);
}

/* grid.f90:334-399 metrics; pcg.f90:433-466 / rbgs.f90:489-521 inverse products. */
static void ax_build(ax_t* a, int n, double L, int stretch, double gamma)
{
    int len = n + 2 * OG + 4;
    a->n = n;
    a->buf = (double*)calloc((size_t)6 * len, sizeof(double));
    int off = OG + 1;
    a->c = a->buf + 0 * len + off;
    a->f = a->buf + 1 * len + off;
    a->dc = a->buf + 2 * len + off;
    a->df = a->buf + 3 * len + off;
    a->ip = a->buf + 4 * len + off;
    a->im = a->buf + 5 * len + off;
    for (int i = 1 - OG; i <= n + OG; ++i) a->c[i] = coord_c(i, n, L, stretch, gamma);
    for (int i = -OG; i <= n + OG; ++i) a->f[i] = coord_f(i, n, L, stretch, gamma);
    for (int i = 2 - OG; i <= n + OG; ++i) a->dc[i] = a->c[i] - a->c[i - 1];
    for (int i = 1 - OG; i <= n + OG; ++i) a->df[i] = a->f[i] - a->f[i - 1];
    for (int i = 1; i <= n; ++i) {
        a->ip[i] = 1.0 / (a->dc[i + 1] * a->df[i]);
        a->im[i] = 1.0 / (a->dc[i] * a->df[i]);
    }
}

static void lev_build(lev_t* l, const or_grid_desc* g, int lev)
{
    for (int a = 0; a < 3; ++a) {
        l->n[a] = g->n[a] >> lev;
        ax_build(&l->ax[a], l->n[a], g->L[a], g->stretch[a], g->gamma);
    }
}

static void lev_free(lev_t* l)
{
    for (int a = 0; a < 3; ++a) free(l->ax[a].buf);
}

void or_level_dims(const or_grid_desc* g, int32_t lev, int32_t* n3)
{
    for (int a = 0; a < 3; ++a) n3[a] = g->n[a] >> lev;
}

static inline int wrapi(int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); }
static inline int clampi(int j, int n) { return j < 1 ? 1 : (j > n ? n : j); }

#define IDX(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))

/* ----------------------------------------------------- operator and smoother */

/* pcg.f90:163-176 (computeQv) with rbgs.f90:300-315 coefficients. */
static inline double op_cell(const lev_t* L, const double* beta, const double* p, int i, int j, int k)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const int E = wrapi(i + 1, nx), W = wrapi(i - 1, nx);
    const int N = clampi(j + 1, ny), S = clampi(j - 1, ny);
    const int F = wrapi(k + 1, nz), B = wrapi(k - 1, nz);
    const double bc = beta[IDX(i, j, k)];
    const double aR = 0.5 * (beta[IDX(E, j, k)] + bc) * L->ax[0].ip[i];
    const double aL = 0.5 * (bc + beta[IDX(W, j, k)]) * L->ax[0].im[i];
    const double aT = 0.5 * (beta[IDX(i, N, k)] + bc) * L->ax[1].ip[j];
    const double aBo = 0.5 * (bc + beta[IDX(i, S, k)]) * L->ax[1].im[j];
    const double aF = 0.5 * (beta[IDX(i, j, F)] + bc) * L->ax[2].ip[k];
    const double aBa = 0.5 * (bc + beta[IDX(i, j, B)]) * L->ax[2].im[k];
    const double pc = p[IDX(i, j, k)];
    return aR * (p[IDX(E, j, k)] - pc) - aL * (pc - p[IDX(W, j, k)]) + aT * (p[IDX(i, N, k)] - pc) -
           aBo * (pc - p[IDX(i, S, k)]) + aF * (p[IDX(i, j, F)] - pc) - aBa * (pc - p[IDX(i, j, B)]);
}

static void op_apply(const lev_t* L, const double* beta, const double* p, double* q)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) q[IDX(i, j, k)] = op_cell(L, beta, p, i, j, k);
}

void or_apply_op(const or_grid_desc* g, const double* beta, const double* p, double* q)
{
    lev_t L;
    lev_build(&L, g, 0);
    op_apply(&L, beta, p, q);
    lev_free(&L);
}

/* rbgs.f90:276-333 (computeResiduals): res = q - A p. */
static void residual(const lev_t* L, const double* beta, const double* q, const double* p, double* res)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) res[IDX(i, j, k)] = q[IDX(i, j, k)] - op_cell(L, beta, p, i, j, k);
}

/* deterministic sums: one partial per k-plane, added in plane order. */
static double sum_planes(int nx, int ny, int nz, const double* a, const double* b, int mode)
{
    double* part = (double*)malloc(sizeof(double) * (size_t)nz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double s = 0.0;
        const size_t o = (size_t)k * nx * ny;
        for (size_t t = 0; t < (size_t)nx * ny; ++t) {
            double v = mode == 0 ? a[o + t] : (mode == 1 ? a[o + t] * b[o + t] : a[o + t] * a[o + t]);
            s += v;
        }
        part[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < nz; ++k) s += part[k];
    free(part);
    return s;
}

/* rbgs.f90:403-446 (correctFullNeumann): subtract the arithmetic mean. */
static void remove_mean(const lev_t* L, double* p)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const size_t n = (size_t)nx * ny * nz;
    double m = sum_planes(nx, ny, nz, p, NULL, 0) / (double)n;
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < n; ++t) p[t] -= m;
}

/* rbgs.f90:186-272 (iterateGS): red list then black list, colour = parity of i+j+k. */
static void rbgs_sweeps(const lev_t* L, const double* beta, const double* q, double* p, int sweeps,
                        int rm_mean)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    for (int it = 0; it < sweeps; ++it) {
        for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(static)
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j) {
                    const int i0 = (((1 + j + k) & 1) == colour) ? 1 : 2; /* (i+j+k) mod 2 == colour */
                    for (int i = i0; i <= nx; i += 2) {
                        const int E = wrapi(i + 1, nx), W = wrapi(i - 1, nx);
                        const int N = clampi(j + 1, ny), S = clampi(j - 1, ny);
                        const int F = wrapi(k + 1, nz), B = wrapi(k - 1, nz);
                        const double bc = beta[IDX(i, j, k)];
                        const double aR = 0.5 * (beta[IDX(E, j, k)] + bc) * L->ax[0].ip[i];
                        const double aL = 0.5 * (bc + beta[IDX(W, j, k)]) * L->ax[0].im[i];
                        const double aT = 0.5 * (beta[IDX(i, N, k)] + bc) * L->ax[1].ip[j];
                        const double aBo = 0.5 * (bc + beta[IDX(i, S, k)]) * L->ax[1].im[j];
                        const double aF = 0.5 * (beta[IDX(i, j, F)] + bc) * L->ax[2].ip[k];
                        const double aBa = 0.5 * (bc + beta[IDX(i, j, B)]) * L->ax[2].im[k];
                        p[IDX(i, j, k)] = (-q[IDX(i, j, k)] + (aR * p[IDX(E, j, k)] + aL * p[IDX(W, j, k)]) +
                                           (aT * p[IDX(i, N, k)] + aBo * p[IDX(i, S, k)]) +
                                           (aF * p[IDX(i, j, F)] + aBa * p[IDX(i, j, B)])) /
                                          (aR + aL + aT + aBo + aF + aBa);
                    }
                }
        }
        if (rm_mean) remove_mean(L, p);
    }
}

/* interpolation.f90:80-124 (restrictionOp3D) over fine/coarse centres, field.f90:1476-1505. */
static void restrict_lv(const lev_t* Lf, const lev_t* Lc, const double* fine, double* coarse)
{
    const int nxf = Lf->n[0], nyf = Lf->n[1];
    const int nxc = Lc->n[0], nyc = Lc->n[1], nzc = Lc->n[2];
    const double *xf = Lf->ax[0].c, *yf = Lf->ax[1].c, *zf = Lf->ax[2].c;
    const double *xc = Lc->ax[0].c, *yc = Lc->ax[1].c, *zc = Lc->ax[2].c;
#define FI(i, j, k) ((size_t)((i) - 1) + (size_t)nxf * ((size_t)((j) - 1) + (size_t)nyf * (size_t)((k) - 1)))
#define CI(i, j, k) ((size_t)((i) - 1) + (size_t)nxc * ((size_t)((j) - 1) + (size_t)nyc * (size_t)((k) - 1)))
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzc; ++k)
        for (int j = 1; j <= nyc; ++j)
            for (int i = 1; i <= nxc; ++i) {
                const double w1 = (xf[2 * i] - xc[i]) * (yf[2 * j] - yc[j]) * (zf[2 * k] - zc[k]);
                const double w2 = (xc[i] - xf[2 * i - 1]) * (yf[2 * j] - yc[j]) * (zf[2 * k] - zc[k]);
                const double w3 = (xf[2 * i] - xc[i]) * (yc[j] - yf[2 * j - 1]) * (zf[2 * k] - zc[k]);
                const double w4 = (xc[i] - xf[2 * i - 1]) * (yc[j] - yf[2 * j - 1]) * (zf[2 * k] - zc[k]);
                const double w5 = (xf[2 * i] - xc[i]) * (yf[2 * j] - yc[j]) * (zc[k] - zf[2 * k - 1]);
                const double w6 = (xc[i] - xf[2 * i - 1]) * (yf[2 * j] - yc[j]) * (zc[k] - zf[2 * k - 1]);
                const double w7 = (xf[2 * i] - xc[i]) * (yc[j] - yf[2 * j - 1]) * (zc[k] - zf[2 * k - 1]);
                const double w8 = (xc[i] - xf[2 * i - 1]) * (yc[j] - yf[2 * j - 1]) * (zc[k] - zf[2 * k - 1]);
                const double V = (xf[2 * i] - xf[2 * i - 1]) * (yf[2 * j] - yf[2 * j - 1]) * (zf[2 * k] - zf[2 * k - 1]);
                coarse[CI(i, j, k)] =
                    (w1 * fine[FI(2 * i - 1, 2 * j - 1, 2 * k - 1)] + w2 * fine[FI(2 * i, 2 * j - 1, 2 * k - 1)] +
                     w3 * fine[FI(2 * i - 1, 2 * j, 2 * k - 1)] + w4 * fine[FI(2 * i, 2 * j, 2 * k - 1)] +
                     w5 * fine[FI(2 * i - 1, 2 * j - 1, 2 * k)] + w6 * fine[FI(2 * i, 2 * j - 1, 2 * k)] +
                     w7 * fine[FI(2 * i - 1, 2 * j, 2 * k)] + w8 * fine[FI(2 * i, 2 * j, 2 * k)]) /
                    V;
            }
}

/* interpolation.f90:129-181 (prolongationOp3D) with coarse ghosts, then unarySum
 * (multiGrid.f90:222-223): fine += P coarse. */
static void prolong_add_lv(const lev_t* Lf, const lev_t* Lc, const double* coarse, double* fine)
{
    const int nxf = Lf->n[0], nyf = Lf->n[1], nzf = Lf->n[2];
    const int nxc = Lc->n[0], nyc = Lc->n[1], nzc = Lc->n[2];
    const double *xf = Lf->ax[0].c, *yf = Lf->ax[1].c, *zf = Lf->ax[2].c;
    const double *xc = Lc->ax[0].c, *yc = Lc->ax[1].c, *zc = Lc->ax[2].c;
#define CG(i, j, k) coarse[CI(wrapi(i, nxc), clampi(j, nyc), wrapi(k, nzc))]
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzf; ++k)
        for (int j = 1; j <= nyf; ++j)
            for (int i = 1; i <= nxf; ++i) {
                const int ih = i / 2, jh = j / 2, kh = k / 2;
                const double w1 = (xc[ih + 1] - xf[i]) * (yc[jh + 1] - yf[j]) * (zc[kh + 1] - zf[k]);
                const double w2 = (xf[i] - xc[ih]) * (yc[jh + 1] - yf[j]) * (zc[kh + 1] - zf[k]);
                const double w3 = (xc[ih + 1] - xf[i]) * (yf[j] - yc[jh]) * (zc[kh + 1] - zf[k]);
                const double w4 = (xf[i] - xc[ih]) * (yf[j] - yc[jh]) * (zc[kh + 1] - zf[k]);
                const double w5 = (xc[ih + 1] - xf[i]) * (yc[jh + 1] - yf[j]) * (zf[k] - zc[kh]);
                const double w6 = (xf[i] - xc[ih]) * (yc[jh + 1] - yf[j]) * (zf[k] - zc[kh]);
                const double w7 = (xc[ih + 1] - xf[i]) * (yf[j] - yc[jh]) * (zf[k] - zc[kh]);
                const double w8 = (xf[i] - xc[ih]) * (yf[j] - yc[jh]) * (zf[k] - zc[kh]);
                const double V = (xc[ih + 1] - xc[ih]) * (yc[jh + 1] - yc[jh]) * (zc[kh + 1] - zc[kh]);
                const double v = (w1 * CG(ih, jh, kh) + w2 * CG(ih + 1, jh, kh) + w3 * CG(ih, jh + 1, kh) +
                                  w4 * CG(ih + 1, jh + 1, kh) + w5 * CG(ih, jh, kh + 1) +
                                  w6 * CG(ih + 1, jh, kh + 1) + w7 * CG(ih, jh + 1, kh + 1) +
                                  w8 * CG(ih + 1, jh + 1, kh + 1)) /
                                 V;
                fine[FI(i, j, k)] = fine[FI(i, j, k)] + v;
            }
#undef CG
}
#undef FI
#undef CI

/* poissMat.f90:714-1290 + 289-312: exact solve of the V-scaled symmetric system
 * -V*A x + lambda = -V*q, sum(x) = 0.  M = -V*A is symmetric positive semi-definite
 * with null space span(1); lambda = mean(-V*q) projects the right-hand side onto
 * range(M) and CG from x=0 stays orthogonal to 1, so the CG limit is the unique
 * Lagrange solution. */
static int coarse_exact_lv(const lev_t* L, const double* beta, const double* q, double* x)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const size_t n = (size_t)nx * ny * nz;
    double* V = (double*)malloc(sizeof(double) * n);
    double* b = (double*)malloc(sizeof(double) * n);
    double* r = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    double* Md = (double*)malloc(sizeof(double) * n);
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                V[IDX(i, j, k)] = L->ax[0].df[i] * L->ax[1].df[j] * L->ax[2].df[k];
    double lam = 0.0;
    for (size_t t = 0; t < n; ++t) {
        b[t] = -V[t] * q[t];
        lam += b[t];
    }
    lam /= (double)n;
    for (size_t t = 0; t < n; ++t) {
        b[t] -= lam;
        x[t] = 0.0;
        r[t] = b[t];
        d[t] = b[t];
    }
    double rr = 0.0, bb = 0.0;
    for (size_t t = 0; t < n; ++t) rr += r[t] * r[t];
    bb = rr;
    int it = 0, maxit = (int)(20 * n + 100);
    if (bb > 0.0) {
        for (it = 0; it < maxit; ++it) {
            op_apply(L, beta, d, Md);
            double dMd = 0.0;
            for (size_t t = 0; t < n; ++t) {
                Md[t] *= -V[t];
                dMd += d[t] * Md[t];
            }
            if (dMd <= 0.0) break;
            double a = rr / dMd, rr2 = 0.0, rm = 0.0;
            for (size_t t = 0; t < n; ++t) {
                x[t] += a * d[t];
                r[t] -= a * Md[t];
                rm += r[t];
            }
            /* r stays in range(M) = {sum = 0}; rounding would otherwise leave a
             * component along the null vector that CG cannot remove, x would
             * drift along it and lose digits when its mean is taken out */
            rm /= (double)n;
            for (size_t t = 0; t < n; ++t) {
                r[t] -= rm;
                rr2 += r[t] * r[t];
            }
            if (rr2 <= 1e-30 * bb) break;
            double gmm = rr2 / rr;
            rr = rr2;
            for (size_t t = 0; t < n; ++t) d[t] = r[t] + gmm * d[t];
        }
    }
    double m = 0.0;
    for (size_t t = 0; t < n; ++t) m += x[t];
    m /= (double)n;
    for (size_t t = 0; t < n; ++t) x[t] -= m;
    free(V);
    free(b);
    free(r);
    free(d);
    free(Md);
    return it;
}

/* ---------------------------------------------------------- exported blocks */

void or_rbgs(const or_grid_desc* g, int32_t lev, const double* beta, const double* q, double* p,
             int32_t sweeps, int32_t remove_mean_flag)
{
    lev_t L;
    lev_build(&L, g, lev);
    rbgs_sweeps(&L, beta, q, p, sweeps, remove_mean_flag);
    lev_free(&L);
}

void or_residual(const or_grid_desc* g, int32_t lev, const double* beta, const double* q, const double* p,
                 double* res)
{
    lev_t L;
    lev_build(&L, g, lev);
    residual(&L, beta, q, p, res);
    lev_free(&L);
}

void or_restrict(const or_grid_desc* g, int32_t lev_fine, const double* fine, double* coarse)
{
    lev_t Lf, Lc;
    lev_build(&Lf, g, lev_fine);
    lev_build(&Lc, g, lev_fine + 1);
    restrict_lv(&Lf, &Lc, fine, coarse);
    lev_free(&Lf);
    lev_free(&Lc);
}

void or_prolong_add(const or_grid_desc* g, int32_t lev_fine, const double* coarse, double* fine)
{
    lev_t Lf, Lc;
    lev_build(&Lf, g, lev_fine);
    lev_build(&Lc, g, lev_fine + 1);
    prolong_add_lv(&Lf, &Lc, coarse, fine);
    lev_free(&Lf);
    lev_free(&Lc);
}

int or_coarse_exact(const or_grid_desc* g, int32_t lev, const double* beta, const double* q, double* p)
{
    lev_t L;
    lev_build(&L, g, lev);
    int it = coarse_exact_lv(&L, beta, q, p);
    lev_free(&L);
    return it;
}

/* --------------------------------------------------------------- multigrid */

typedef struct {
    int nlev;       /* user levels (multiGrid.f90:81)                  */
    int ntot;       /* user levels + sub-levels used by coarse_mode 1  */
    lev_t* L;
    double **beta, **p, **q, **res;
    const or_pcg_params* prm;
} mg_t;

static size_t lev_cells(const lev_t* L) { return (size_t)L->n[0] * L->n[1] * L->n[2]; }

static int mg_build(mg_t* m, const or_grid_desc* g, const or_pcg_params* prm, const double* beta)
{
    m->prm = prm;
    m->nlev = prm->n_levels;
    for (int l = 1; l < m->nlev; ++l)
        for (int a = 0; a < 3; ++a)
            if ((g->n[a] % (1 << l)) != 0) return -1; /* multiGrid.f90:116-145 */
    int ntot = m->nlev;
    if (prm->coarse_mode == 1) {
        /* sub-hierarchy: keep halving while every dimension is even, down to
         * at most 64 cells, solved exactly at the bottom */
        for (;;) {
            int ok = 1;
            long cells = 1;
            for (int a = 0; a < 3; ++a) {
                int nn = g->n[a] >> (ntot - 1);
                cells *= nn;
                if (nn % 2 != 0 || nn < 2) ok = 0;
            }
            if (!ok || cells <= 64) break;
            ntot++;
        }
    }
    m->ntot = ntot;
    m->L = (lev_t*)calloc(ntot, sizeof(lev_t));
    m->beta = (double**)calloc(ntot, sizeof(double*));
    m->p = (double**)calloc(ntot, sizeof(double*));
    m->q = (double**)calloc(ntot, sizeof(double*));
    m->res = (double**)calloc(ntot, sizeof(double*));
    for (int l = 0; l < ntot; ++l) {
        lev_build(&m->L[l], g, l);
        size_t n = lev_cells(&m->L[l]);
        m->beta[l] = (double*)malloc(sizeof(double) * n);
        m->p[l] = (double*)calloc(n, sizeof(double));
        m->q[l] = (double*)calloc(n, sizeof(double));
        m->res[l] = (double*)calloc(n, sizeof(double));
    }
    memcpy(m->beta[0], beta, sizeof(double) * lev_cells(&m->L[0]));
    /* beta is restricted every V-cycle in multiGrid.f90:210; it does not change
     * within a solve, so restricting once per solve gives the same values. */
    for (int l = 1; l < ntot; ++l) restrict_lv(&m->L[l - 1], &m->L[l], m->beta[l - 1], m->beta[l]);
    return 0;
}

static void mg_free(mg_t* m)
{
    for (int l = 0; l < m->ntot; ++l) {
        lev_free(&m->L[l]);
        free(m->beta[l]);
        free(m->p[l]);
        free(m->q[l]);
        free(m->res[l]);
    }
    free(m->L);
    free(m->beta);
    free(m->p);
    free(m->q);
    free(m->res);
}

static void vcycle(mg_t* m, int l, int last, int reset);

/* level `last` of the user hierarchy: exact solve, or K cycles on the sub-hierarchy. */
static void coarsest(mg_t* m, int l)
{
    if (m->prm->coarse_mode == 0 || m->ntot == m->nlev) {
        coarse_exact_lv(&m->L[l], m->beta[l], m->q[l], m->p[l]);
        return;
    }
    memset(m->p[l], 0, sizeof(double) * lev_cells(&m->L[l]));
    for (int c = 0; c < m->prm->coarse_cycles; ++c) vcycle(m, l, m->ntot - 1, 0);
}

/* multiGrid.f90:186-232 (mgVcycle), levels counted from the finest (0). */
static void vcycle(mg_t* m, int l, int last, int reset)
{
    const or_pcg_params* prm = m->prm;
    if (l == last) {
        if (last == m->nlev - 1)
            coarsest(m, l);
        else
            coarse_exact_lv(&m->L[l], m->beta[l], m->q[l], m->p[l]);
        return;
    }
    if (reset) memset(m->p[l], 0, sizeof(double) * lev_cells(&m->L[l]));
    rbgs_sweeps(&m->L[l], m->beta[l], m->q[l], m->p[l], prm->n_pre, 1);
    residual(&m->L[l], m->beta[l], m->q[l], m->p[l], m->res[l]);
    restrict_lv(&m->L[l], &m->L[l + 1], m->res[l], m->q[l + 1]);
    vcycle(m, l + 1, last, 1);
    prolong_add_lv(&m->L[l], &m->L[l + 1], m->p[l + 1], m->p[l]);
    rbgs_sweeps(&m->L[l], m->beta[l], m->q[l], m->p[l], prm->n_post, 1);
}

static int32_t g_mg_cycles_total = 0, g_mg_cycles_max = 0, g_mg_cycles_last = 0;

void or_last_mg_cycles(int32_t* total, int32_t* max_per_apply, int32_t* last)
{
    *total = g_mg_cycles_total;
    *max_per_apply = g_mg_cycles_max;
    *last = g_mg_cycles_last;
}

/* multiGrid.f90:165-182 (solveMG); the finest level starts from zero
 * (pcg.f90:118,128).  mg_tol = 0: a fixed number of V-cycles.  mg_tol > 0:
 * continueIterating (multiGrid.f90:236-279) before every cycle -- stop after
 * mg_max_iter cycles, else continue while the finest level's residual
 * ||q - A p|| / (||q|| + tiny) exceeds tolMG.  That residual is the one
 * solveRBGS leaves after the finest level's last smoothing (rbgs.f90:168-181,
 * computeResiduals rbgs.f90:276-333); before the first cycle p = 0.  The
 * right-hand side is PCG's r, assigned on the interior only (pcg.f90:117,349),
 * so its halo adds nothing to the norm. */
static void mg_apply_built(mg_t* m, const double* r, double* z)
{
    const lev_t* L = &m->L[0];
    size_t n = lev_cells(L);
    memcpy(m->q[0], r, sizeof(double) * n);
    memset(m->p[0], 0, sizeof(double) * n);
    int cycles = 0;
    if (m->prm->mg_tol > 0.0) {
        const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
        const double tiny = 2.2250738585072014e-308;
        const double qn = sqrt(sum_planes(nx, ny, nz, m->q[0], NULL, 2));
        double res = qn / (qn + tiny);
        while (cycles != m->prm->mg_max_iter && res > m->prm->mg_tol) {
            vcycle(m, 0, m->nlev - 1, 0);
            ++cycles;
            residual(L, m->beta[0], m->q[0], m->p[0], m->res[0]);
            res = sqrt(sum_planes(nx, ny, nz, m->res[0], NULL, 2)) / (qn + tiny);
        }
    } else {
        for (int c = 0; c < m->prm->mg_cycles; ++c) vcycle(m, 0, m->nlev - 1, 0);
        cycles = m->prm->mg_cycles;
    }
    g_mg_cycles_total += cycles;
    g_mg_cycles_last = cycles;
    if (cycles > g_mg_cycles_max) g_mg_cycles_max = cycles;
    memcpy(z, m->p[0], sizeof(double) * n);
}

int or_mg_apply(const or_grid_desc* g, const or_pcg_params* prm, const double* beta, const double* r, double* z)
{
    mg_t m;
    if (mg_build(&m, g, prm, beta)) return -1;
    g_mg_cycles_total = g_mg_cycles_max = 0;
    mg_apply_built(&m, r, z);
    mg_free(&m);
    return 0;
}

/* pcg.f90:103-140 (solvePCG) with the stop test of pcg.f90:359-429.  Returns 0 on
 * convergence, 1 when max_iter updates were made without reaching tol.  `iters`
 * counts direction updates (the reference's iter_ counter is iters+1 on
 * convergence, pcg.f90:401). */
int or_pcg_solve(const or_grid_desc* g, const or_pcg_params* prm, const double* beta, const double* q,
                 double* psi, int32_t* iters, double* rel_res)
{
    mg_t m;
    if (mg_build(&m, g, prm, beta)) return -1;
    g_mg_cycles_total = g_mg_cycles_max = 0;
    const lev_t* L = &m.L[0];
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const size_t n = lev_cells(L);
    double* r = (double*)malloc(sizeof(double) * n);
    double* d = (double*)malloc(sizeof(double) * n);
    double* d0 = (double*)malloc(sizeof(double) * n);
    double* qv = (double*)malloc(sizeof(double) * n);
    residual(L, beta, q, psi, r);
    mg_apply_built(&m, r, d);
    double delta = sum_planes(nx, ny, nz, r, d, 1);
    const double qn = sqrt(sum_planes(nx, ny, nz, q, NULL, 2));
    int it = 0, status = 1;
    double res = 0.0;
    for (;;) {
        res = sqrt(sum_planes(nx, ny, nz, r, NULL, 2)) / sqrt(qn * qn + 2.2250738585072014e-308);
        if (res <= prm->tol) {
            status = 0;
            break;
        }
        if (it == prm->max_iter) break;
        it++;
        op_apply(L, beta, d, qv);
        const double alpha = delta / sum_planes(nx, ny, nz, d, qv, 1);
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < n; ++t) {
            psi[t] = psi[t] + alpha * d[t];
            r[t] = r[t] - alpha * qv[t];
            d0[t] = d[t];
        }
        mg_apply_built(&m, r, d);
        const double delta0 = delta;
        delta = sum_planes(nx, ny, nz, r, d, 1);
        double gmm;
        if (prm->flexible)
            gmm = -alpha * sum_planes(nx, ny, nz, d, qv, 1) / delta0;
        else
            gmm = delta / delta0;
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < n; ++t) d[t] = d[t] + gmm * d0[t];
    }
    *iters = it;
    *rel_res = res;
    free(r);
    free(d);
    free(d0);
    free(qv);
    mg_free(&m);
    return status;
}

/* ----------------------------------------------------------------- FFT path */

/* mixed-radix recursive DFT, y[k] = sum_m x[m*xs] exp(sign*2*pi*i*m*k/n). */
static void dft_rec(int n, const double complex* x, int xs, double complex* y, int sign, double complex* tmp)
{
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int r = 0;
    for (int f = 2; f <= n; ++f)
        if (n % f == 0) {
            r = f;
            break;
        }
    int mlen = n / r;
    if (r == n && n > 5) { /* prime length: direct DFT */
        for (int k = 0; k < n; ++k) {
            double complex s = 0;
            for (int t = 0; t < n; ++t) {
                double ang = sign * 2.0 * M_PI * (double)((long)t * k % n) / n;
                s += x[t * xs] * cexp(I * ang);
            }
            y[k] = s;
        }
        return;
    }
    for (int q = 0; q < r; ++q) dft_rec(mlen, x + q * xs, xs * r, y + q * mlen, sign, tmp + n);
    for (int k = 0; k < mlen; ++k)
        for (int s = 0; s < r; ++s) {
            double complex acc = 0;
            int kk = k + mlen * s;
            for (int q = 0; q < r; ++q) {
                double ang = sign * 2.0 * M_PI * (double)((long)q * kk % n) / n;
                acc += y[q * mlen + k] * cexp(I * ang);
            }
            tmp[kk] = acc;
        }
    memcpy(y, tmp, sizeof(double complex) * n);
}

static void dft(int n, const double complex* x, double complex* y, int sign)
{
    double complex* tmp = (double complex*)malloc(sizeof(double complex) * 4 * (size_t)n + 64);
    dft_rec(n, x, 1, y, sign, tmp);
    free(tmp);
}

void or_fft_c2c(int32_t n, const double* in_ri, double* out_ri, int32_t sign)
{
    dft(n, (const double complex*)in_ri, (double complex*)out_ri, sign);
}

/* fastPoissonSolver.f90:255-276 (initOffDiagMatrix) */
static void offdiag(const lev_t* L, double* am, double* ap)
{
    const int ny = L->n[1];
    for (int j = 1; j <= ny; ++j) {
        am[j - 1] = 1.0 / (L->ax[1].dc[j] * L->ax[1].df[j]);
        ap[j - 1] = 1.0 / (L->ax[1].dc[j + 1] * L->ax[1].df[j]);
    }
    am[0] = 0.0;
    ap[ny - 1] = 0.0;
}

/* Thomas solve of am_j x_{j-1} + (lam-am_j-ap_j) x_j + ap_j x_{j+1} = d_j
 * (fastPoissonSolver.f90:303-330 without the row scaling, which does not change
 * the solution).  Pinned: x_n = 0 and rows 1..n-1 solved (singular (0,0) mode). */
static void thomas(int n, const double* am, const double* ap, double lam, int pinned, double complex* d,
                   double* cp)
{
    int m = pinned ? n - 1 : n;
    if (pinned) d[n - 1] = 0.0;
    if (m <= 0) return;
    double b = lam - am[0] - ap[0];
    cp[0] = ap[0] / b;
    d[0] = d[0] / b;
    for (int j = 1; j < m; ++j) {
        double den = (lam - am[j] - ap[j]) - am[j] * cp[j - 1];
        cp[j] = ap[j] / den;
        d[j] = (d[j] - am[j] * d[j - 1]) / den;
    }
    for (int j = m - 2; j >= 0; --j) d[j] = d[j] - cp[j] * d[j + 1];
}

void or_tridiag(const or_grid_desc* g, double lam, int32_t pinned, double* rhs_ri)
{
    lev_t L;
    lev_build(&L, g, 0);
    int ny = L.n[1];
    double* am = (double*)malloc(sizeof(double) * ny);
    double* ap = (double*)malloc(sizeof(double) * ny);
    double* cp = (double*)malloc(sizeof(double) * ny);
    offdiag(&L, am, ap);
    thomas(ny, am, ap, lam, pinned, (double complex*)rhs_ri, cp);
    free(am);
    free(ap);
    free(cp);
    lev_free(&L);
}

/* fastPoissonSolver.f90:215-251 (computeEigenvalues) */
static void eigen(const lev_t* L, int a, double* lam)
{
    const int n = L->n[a];
    const double dx = L->ax[a].c[2] - L->ax[a].c[1];
    for (int i = 1; i <= n; ++i) {
        double v = (i - 1) * 2.0 * M_PI;
        lam[i - 1] = 2.0 * (cos(v / n) - 1.0) / (dx * dx);
    }
}

/* fastPoissonSolver.f90:337-384 (solveFPS) + :388-423 (zeroAveragePressure). */
int or_fft_solve(const or_grid_desc* g, const double* s, double* p)
{
    if (g->stretch[0] || g->stretch[2]) return -1; /* grid.f90:219-230 */
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int nh = nx / 2 + 1;
    double* lx = (double*)malloc(sizeof(double) * nx);
    double* lz = (double*)malloc(sizeof(double) * nz);
    double* am = (double*)malloc(sizeof(double) * ny);
    double* ap = (double*)malloc(sizeof(double) * ny);
    eigen(&L, 0, lx);
    eigen(&L, 2, lz);
    offdiag(&L, am, ap);
    /* spectral array S[kx][kz][j], j fastest */
    double complex* S = (double complex*)malloc(sizeof(double complex) * (size_t)nh * nz * ny);
#define SP(kx, kz, j) S[((size_t)(kx) * nz + (size_t)(kz)) * ny + (size_t)(j)]
    /* x r2c then z c2c forward */
#pragma omp parallel
    {
        double complex* a = (double complex*)malloc(sizeof(double complex) * (nx > nz ? nx : nz));
        double complex* b = (double complex*)malloc(sizeof(double complex) * (nx > nz ? nx : nz));
        double complex* col = (double complex*)malloc(sizeof(double complex) * (size_t)nh * nz);
#pragma omp for schedule(static)
        for (int j = 1; j <= ny; ++j) {
            for (int k = 1; k <= nz; ++k) {
                for (int i = 1; i <= nx; ++i) a[i - 1] = s[IDX(i, j, k)];
                dft(nx, a, b, -1);
                for (int kx = 0; kx < nh; ++kx) col[(size_t)kx * nz + (k - 1)] = b[kx];
            }
            for (int kx = 0; kx < nh; ++kx) {
                dft(nz, col + (size_t)kx * nz, b, -1);
                for (int kz = 0; kz < nz; ++kz) SP(kx, kz, j - 1) = b[kz];
            }
        }
        free(a);
        free(b);
        free(col);
    }
    /* tridiagonal in y per (kx,kz), fastPoissonSolver.f90:280-333 */
#pragma omp parallel
    {
        double* cp = (double*)malloc(sizeof(double) * ny);
#pragma omp for schedule(static)
        for (int kz = 0; kz < nz; ++kz)
            for (int kx = 0; kx < nh; ++kx) {
                int pinned = (kx == 0 && kz == 0);
                thomas(ny, am, ap, lx[kx] + lz[kz], pinned, &SP(kx, kz, 0), cp);
            }
        free(cp);
    }
    /* inverse z, inverse x (c2r), scale by 1/(nx*nz) */
#pragma omp parallel
    {
        double complex* a = (double complex*)malloc(sizeof(double complex) * (nx > nz ? nx : nz));
        double complex* b = (double complex*)malloc(sizeof(double complex) * (nx > nz ? nx : nz));
        double complex* col = (double complex*)malloc(sizeof(double complex) * (size_t)nh * nz);
#pragma omp for schedule(static)
        for (int j = 1; j <= ny; ++j) {
            for (int kx = 0; kx < nh; ++kx) {
                for (int kz = 0; kz < nz; ++kz) a[kz] = SP(kx, kz, j - 1);
                dft(nz, a, col + (size_t)kx * nz, +1);
            }
            for (int k = 1; k <= nz; ++k) {
                for (int kx = 0; kx < nh; ++kx) a[kx] = col[(size_t)kx * nz + (k - 1)];
                for (int kx = nh; kx < nx; ++kx) a[kx] = conj(a[nx - kx]);
                if (nx % 2 == 0) a[nx / 2] = creal(a[nx / 2]);
                a[0] = creal(a[0]);
                dft(nx, a, b, +1);
                for (int i = 1; i <= nx; ++i) p[IDX(i, j, k)] = creal(b[i - 1]) / (double)(nx * nz);
            }
        }
        free(a);
        free(b);
        free(col);
    }
#undef SP
    /* zeroAveragePressure */
    double m = or_volume_mean(g, p);
    const size_t n = (size_t)nx * ny * nz;
    for (size_t t = 0; t < n; ++t) p[t] -= m;
    free(S);
    free(lx);
    free(lz);
    free(am);
    free(ap);
    lev_free(&L);
    return 0;
}

void or_fft_operator(const or_grid_desc* g, const double* p, double* q)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const double dx = L.ax[0].c[2] - L.ax[0].c[1], dz = L.ax[2].c[2] - L.ax[2].c[1];
    double* am = (double*)malloc(sizeof(double) * ny);
    double* ap = (double*)malloc(sizeof(double) * ny);
    offdiag(&L, am, ap);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double pc = p[IDX(i, j, k)];
                double vx = (p[IDX(wrapi(i + 1, nx), j, k)] - 2.0 * pc + p[IDX(wrapi(i - 1, nx), j, k)]) / (dx * dx);
                double vz = (p[IDX(i, j, wrapi(k + 1, nz))] - 2.0 * pc + p[IDX(i, j, wrapi(k - 1, nz))]) / (dz * dz);
                double vy = am[j - 1] * (p[IDX(i, clampi(j - 1, ny), k)] - pc) +
                            ap[j - 1] * (p[IDX(i, clampi(j + 1, ny), k)] - pc);
                q[IDX(i, j, k)] = vx + vy + vz;
            }
    free(am);
    free(ap);
    lev_free(&L);
}

double or_volume_mean(const or_grid_desc* g, const double* p)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    double s = 0.0, vg = 0.0;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double V = L.ax[0].df[i] * L.ax[1].df[j] * L.ax[2].df[k];
                s += p[IDX(i, j, k)] * V;
                vg += V;
            }
    lev_free(&L);
    return s / vg;
}

/* ----------------------------------------------- RHS, projection, divergence */

#define UYI(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)(j) + (size_t)(ny + 1) * (size_t)((k) - 1)))

/* poissonEqn.f90:176-188 (MG computeSource), auxiliaryRoutines.f90:66-79 */
void or_divergence(const or_grid_desc* g, const double* ux, const double* uy, const double* uz, double* div)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double dx = (ux[IDX(i, j, k)] - ux[IDX(wrapi(i - 1, nx), j, k)]) / L.ax[0].df[i];
                double dy = (uy[UYI(i, j, k)] - uy[UYI(i, j - 1, k)]) / L.ax[1].df[j];
                double dz = (uz[IDX(i, j, k)] - uz[IDX(i, j, wrapi(k - 1, nz))]) / L.ax[2].df[k];
                div[IDX(i, j, k)] = dx + dy + dz;
            }
    lev_free(&L);
}

/* poissonEqn.f90:311-342 (computeBeta) */
void or_beta(const or_grid_desc* g, double alpha_dt, const double* rho, double* beta)
{
    const size_t n = (size_t)g->n[0] * g->n[1] * g->n[2];
    for (size_t t = 0; t < n; ++t) beta[t] = alpha_dt / rho[t];
}

/* momentumEqn.f90:795-869 (addPressureGrad) over the setEqnBounds ranges
 * (momentumEqn.f90:1134-1190): all periodic x/z faces, interior y faces. */
void or_project_mg(const or_grid_desc* g, double alpha_dt, const double* rho, const double* psi, double* ux,
                   double* uy, double* uz)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int E = wrapi(i + 1, nx), F = wrapi(k + 1, nz);
                double inv = 0.5 * (1.0 / rho[IDX(E, j, k)] + 1.0 / rho[IDX(i, j, k)]);
                ux[IDX(i, j, k)] = ux[IDX(i, j, k)] - alpha_dt * inv * (psi[IDX(E, j, k)] - psi[IDX(i, j, k)]) / L.ax[0].dc[i + 1];
                inv = 0.5 * (1.0 / rho[IDX(i, j, F)] + 1.0 / rho[IDX(i, j, k)]);
                uz[IDX(i, j, k)] = uz[IDX(i, j, k)] - alpha_dt * inv * (psi[IDX(i, j, F)] - psi[IDX(i, j, k)]) / L.ax[2].dc[k + 1];
                if (j < ny) {
                    inv = 0.5 * (1.0 / rho[IDX(i, j + 1, k)] + 1.0 / rho[IDX(i, j, k)]);
                    uy[UYI(i, j, k)] = uy[UYI(i, j, k)] - alpha_dt * inv * (psi[IDX(i, j + 1, k)] - psi[IDX(i, j, k)]) / L.ax[1].dc[j + 1];
                }
            }
    lev_free(&L);
}

/* Old surface-tension levels of the FAST splitting: st1 = st^n (st%ptrf_), st2 =
 * st^(n-1) (st%ptrf_%ptrf_), each three staggered components laid out like the
 * velocity (x, z: face i / k after cell i / k; y: faces 0..ny).  A NULL level is
 * zero. */
#define STC(st, c, idx) ((st)[c] ? (st)[c][idx] : 0.0)

/* momentumEqn.f90:999-1127 (FAST makeVelocityDivFree + computeOldPressGrad, nl = 2):
 * u_f -= a dt [(1/rho0) d_f psi + ((1/rho)_f - 1/rho0) G_f] with
 * G_f = d_f(2 psi1 - psi2) - 2 st1_f + st2_f (momentumEqn.f90:1104-1121), on all
 * periodic x/z faces and the interior y faces. */
void or_project_fast(const or_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* psi,
                     const double* psi1, const double* psi2, const double* const st1[3],
                     const double* const st2[3], double* ux, double* uy, double* uz)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const double dxc = L.ax[0].dc[1], dzc = L.ax[2].dc[1];
    const double inv0 = 1.0 / rho0;
    const double* s1[3] = {st1 ? st1[0] : 0, st1 ? st1[1] : 0, st1 ? st1[2] : 0};
    const double* s2[3] = {st2 ? st2[0] : 0, st2 ? st2[1] : 0, st2 ? st2[2] : 0};
#define PH(a, i, j, k) ((a) ? (a)[IDX(i, j, k)] : 0.0)
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int E = wrapi(i + 1, nx), F = wrapi(k + 1, nz);
                const size_t c = IDX(i, j, k);
                double inv = 0.5 * (1.0 / rho[IDX(E, j, k)] + 1.0 / rho[c]);
                double dpr = (psi[IDX(E, j, k)] - psi[c]) / dxc;
                double dpr0 = (2.0 * (PH(psi1, E, j, k) - PH(psi1, i, j, k)) - (PH(psi2, E, j, k) - PH(psi2, i, j, k))) / dxc;
                dpr0 = dpr0 - 2.0 * STC(s1, 0, c) + STC(s2, 0, c);
                ux[c] = ux[c] - alpha_dt * (inv0 * dpr + (inv - inv0) * dpr0);
                inv = 0.5 * (1.0 / rho[IDX(i, j, F)] + 1.0 / rho[c]);
                dpr = (psi[IDX(i, j, F)] - psi[c]) / dzc;
                dpr0 = (2.0 * (PH(psi1, i, j, F) - PH(psi1, i, j, k)) - (PH(psi2, i, j, F) - PH(psi2, i, j, k))) / dzc;
                dpr0 = dpr0 - 2.0 * STC(s1, 2, c) + STC(s2, 2, c);
                uz[c] = uz[c] - alpha_dt * (inv0 * dpr + (inv - inv0) * dpr0);
                if (j < ny) {
                    const size_t yf = UYI(i, j, k);
                    const double dyc = L.ax[1].dc[j + 1];
                    inv = 0.5 * (1.0 / rho[IDX(i, j + 1, k)] + 1.0 / rho[c]);
                    dpr = (psi[IDX(i, j + 1, k)] - psi[c]) / dyc;
                    dpr0 = (2.0 * (PH(psi1, i, j + 1, k) - PH(psi1, i, j, k)) - (PH(psi2, i, j + 1, k) - PH(psi2, i, j, k))) / dyc;
                    dpr0 = dpr0 - 2.0 * STC(s1, 1, yf) + STC(s2, 1, yf);
                    uy[yf] = uy[yf] - alpha_dt * (inv0 * dpr + (inv - inv0) * dpr0);
                }
            }
    lev_free(&L);
}

/* poissonEqn.f90:193-305 (FAST computeSource + computeOldPressDiv, nl = 2):
 * s = rho0/(a dt) div u* + div[(1 - rho0 (1/rho)_f) G_f], G_f as in or_project_fast
 * (computeOldPressDiv, poissonEqn.f90:255-301: the face after the cell reads the
 * st values of its own face index, the face before it those of index - 1). */
void or_source_fast(const or_grid_desc* g, double alpha_dt, double rho0, const double* rho, const double* ux,
                    const double* uy, const double* uz, const double* psi1, const double* psi2,
                    const double* const st1[3], const double* const st2[3], double* s)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const double dxf = L.ax[0].df[1], dzf = L.ax[2].df[1];
    const double dxc = L.ax[0].dc[1], dzc = L.ax[2].dc[1];
    const double r = 1.0 / alpha_dt;
    const double* s1[3] = {st1 ? st1[0] : 0, st1 ? st1[1] : 0, st1 ? st1[2] : 0};
    const double* s2[3] = {st2 ? st2[0] : 0, st2 ? st2[1] : 0, st2 ? st2[2] : 0};
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int E = wrapi(i + 1, nx), W = wrapi(i - 1, nx);
                const int N = clampi(j + 1, ny), S = clampi(j - 1, ny);
                const int F = wrapi(k + 1, nz), B = wrapi(k - 1, nz);
                const size_t c = IDX(i, j, k), cw = IDX(W, j, k), cb = IDX(i, j, B);
                const size_t yn = UYI(i, j, k), ys = UYI(i, j - 1, k);
                double dx = (ux[c] - ux[cw]) / dxf;
                double dy = (uy[yn] - uy[ys]) / L.ax[1].df[j];
                double dz = (uz[c] - uz[cb]) / dzf;
                double sv = (dx + dy + dz) * r * rho0;
                const double ic = 1.0 / rho[c];
                double apx = 1.0 - rho0 * 0.5 * (1.0 / rho[IDX(E, j, k)] + ic);
                double amx = 1.0 - rho0 * 0.5 * (1.0 / rho[cw] + ic);
                double apy = 1.0 - rho0 * 0.5 * (1.0 / rho[IDX(i, N, k)] + ic);
                double amy = 1.0 - rho0 * 0.5 * (1.0 / rho[IDX(i, S, k)] + ic);
                double apz = 1.0 - rho0 * 0.5 * (1.0 / rho[IDX(i, j, F)] + ic);
                double amz = 1.0 - rho0 * 0.5 * (1.0 / rho[cb] + ic);
                double dpxp = (2.0 * (PH(psi1, E, j, k) - PH(psi1, i, j, k)) - (PH(psi2, E, j, k) - PH(psi2, i, j, k))) / dxc;
                dpxp = dpxp - 2.0 * STC(s1, 0, c) + STC(s2, 0, c);
                double dpxm = (2.0 * (PH(psi1, i, j, k) - PH(psi1, W, j, k)) - (PH(psi2, i, j, k) - PH(psi2, W, j, k))) / dxc;
                dpxm = dpxm - 2.0 * STC(s1, 0, cw) + STC(s2, 0, cw);
                double dpyp = (2.0 * (PH(psi1, i, N, k) - PH(psi1, i, j, k)) - (PH(psi2, i, N, k) - PH(psi2, i, j, k))) / L.ax[1].dc[j + 1];
                dpyp = dpyp - 2.0 * STC(s1, 1, yn) + STC(s2, 1, yn);
                double dpym = (2.0 * (PH(psi1, i, j, k) - PH(psi1, i, S, k)) - (PH(psi2, i, j, k) - PH(psi2, i, S, k))) / L.ax[1].dc[j];
                dpym = dpym - 2.0 * STC(s1, 1, ys) + STC(s2, 1, ys);
                double dpzp = (2.0 * (PH(psi1, i, j, F) - PH(psi1, i, j, k)) - (PH(psi2, i, j, F) - PH(psi2, i, j, k))) / dzc;
                dpzp = dpzp - 2.0 * STC(s1, 2, c) + STC(s2, 2, c);
                double dpzm = (2.0 * (PH(psi1, i, j, k) - PH(psi1, i, j, B)) - (PH(psi2, i, j, k) - PH(psi2, i, j, B))) / dzc;
                dpzm = dpzm - 2.0 * STC(s1, 2, cb) + STC(s2, 2, cb);
                dx = (apx * dpxp - amx * dpxm) / dxf;
                dy = (apy * dpyp - amy * dpym) / L.ax[1].df[j];
                dz = (apz * dpzp - amz * dpzm) / dzf;
                s[c] = sv + dx + dy + dz;
            }
#undef PH
    lev_free(&L);
}

/* updatePressure (poissonEqn.f90:374-415) with clean_st_prev_bc (:419-459) for the
 * channel: p = psi; storeOldField (field.f90:1585-1600): psi2 <- psi1 <- psi;
 * storeOldFieldV: st2 <- st1 <- st; then the old st levels are zeroed on the
 * faces at and beyond the Neumann walls, which in the stored range 0..ny of the
 * y component are the wall faces j = 0 and j = ny (the x and z sides are
 * periodic).  p may be NULL; st (the current level) NULL means zero; st1 and st2
 * NULL skip the surface-tension levels. */
void or_update_pressure(const or_grid_desc* g, const double* psi, double* p, double* psi1, double* psi2,
                        const double* const st[3], double* const st1[3], double* const st2[3])
{
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const size_t nc = (size_t)nx * ny * nz, ny1 = (size_t)nx * (ny + 1) * nz;
    for (size_t t = 0; t < nc; ++t) {
        if (p) p[t] = psi[t];
        psi2[t] = psi1[t];
        psi1[t] = psi[t];
    }
    if (!st1 || !st2) return;
    for (int c = 0; c < 3; ++c) {
        const size_t n = c == 1 ? ny1 : nc;
        const double* cur = st ? st[c] : 0;
        for (size_t t = 0; t < n; ++t) {
            st2[c][t] = st1[c][t];
            st1[c][t] = cur ? cur[t] : 0.0;
        }
    }
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) {
            st1[1][UYI(i, 0, k)] = st2[1][UYI(i, 0, k)] = 0.0;
            st1[1][UYI(i, ny, k)] = st2[1][UYI(i, ny, k)] = 0.0;
        }
}
#undef STC

double or_max_abs(int64_t n, const double* a)
{
    double m = 0.0;
    for (int64_t t = 0; t < n; ++t) {
        double v = fabs(a[t]);
        if (v > m) m = v;
    }
    return m;
}

/* ------------------------------------------------------ material properties */

/* cellToVertex (interpolation.f90:188-226): trilinear interpolation of a cell
 * field to the vertices 0..n (faces) from the 8 surrounding centres, ghosts by
 * periodic wrap in x, z and zero gradient in y (the VOF field's walls). */
static void cell_to_vertex(const lev_t* L, const double* q, double* qv)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const ax_t *X = &L->ax[0], *Y = &L->ax[1], *Z = &L->ax[2];
#define QG(i, j, k) q[IDX(wrapi(i, nx), clampi(j, ny), wrapi(k, nz))]
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double w1 = (X->c[i + 1] - X->f[i]) * (Y->f[j] - Y->c[j]) * (Z->c[k + 1] - Z->f[k]);
                const double w2 = (X->f[i] - X->c[i]) * (Y->f[j] - Y->c[j]) * (Z->c[k + 1] - Z->f[k]);
                const double w3 = (X->c[i + 1] - X->f[i]) * (Y->f[j] - Y->c[j]) * (Z->f[k] - Z->c[k]);
                const double w4 = (X->f[i] - X->c[i]) * (Y->f[j] - Y->c[j]) * (Z->f[k] - Z->c[k]);
                const double w5 = (X->c[i + 1] - X->f[i]) * (Y->c[j + 1] - Y->f[j]) * (Z->c[k + 1] - Z->f[k]);
                const double w6 = (X->f[i] - X->c[i]) * (Y->c[j + 1] - Y->f[j]) * (Z->c[k + 1] - Z->f[k]);
                const double w7 = (X->c[i + 1] - X->f[i]) * (Y->c[j + 1] - Y->f[j]) * (Z->f[k] - Z->c[k]);
                const double w8 = (X->f[i] - X->c[i]) * (Y->c[j + 1] - Y->f[j]) * (Z->f[k] - Z->c[k]);
                const double V = (X->c[i + 1] - X->c[i]) * (Y->c[j + 1] - Y->c[j]) * (Z->c[k + 1] - Z->c[k]);
                qv[i + (size_t)(nx + 1) * (j + (size_t)(ny + 1) * k)] =
                    (w1 * QG(i, j + 1, k) + w2 * QG(i + 1, j + 1, k) + w3 * QG(i, j + 1, k + 1) +
                     w4 * QG(i + 1, j + 1, k + 1) + w5 * QG(i, j, k) + w6 * QG(i + 1, j, k) + w7 * QG(i, j, k + 1) +
                     w8 * QG(i + 1, j, k + 1)) / V;
            }
#undef QG
}

/* vertexToCell (interpolation.f90:231-276): the cell value from its 8 vertices. */
static void vertex_to_cell(const lev_t* L, const double* qv, double* q)
{
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    const ax_t *X = &L->ax[0], *Y = &L->ax[1], *Z = &L->ax[2];
#define QV(i, j, k) qv[(i) + (size_t)(nx + 1) * ((j) + (size_t)(ny + 1) * (k))]
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double w1 = (X->f[i] - X->c[i]) * (Y->c[j] - Y->f[j - 1]) * (Z->f[k] - Z->c[k]);
                const double w2 = (X->c[i] - X->f[i - 1]) * (Y->c[j] - Y->f[j - 1]) * (Z->f[k] - Z->c[k]);
                const double w3 = (X->f[i] - X->c[i]) * (Y->c[j] - Y->f[j - 1]) * (Z->c[k] - Z->f[k - 1]);
                const double w4 = (X->c[i] - X->f[i - 1]) * (Y->c[j] - Y->f[j - 1]) * (Z->c[k] - Z->f[k - 1]);
                const double w5 = (X->f[i] - X->c[i]) * (Y->f[j] - Y->c[j]) * (Z->f[k] - Z->c[k]);
                const double w6 = (X->c[i] - X->f[i - 1]) * (Y->f[j] - Y->c[j]) * (Z->f[k] - Z->c[k]);
                const double w7 = (X->f[i] - X->c[i]) * (Y->f[j] - Y->c[j]) * (Z->c[k] - Z->f[k - 1]);
                const double w8 = (X->c[i] - X->f[i - 1]) * (Y->f[j] - Y->c[j]) * (Z->c[k] - Z->f[k - 1]);
                const double V = (X->f[i] - X->f[i - 1]) * (Y->f[j] - Y->f[j - 1]) * (Z->f[k] - Z->f[k - 1]);
                q[IDX(i, j, k)] = (w1 * QV(i - 1, j, k - 1) + w2 * QV(i, j, k - 1) + w3 * QV(i - 1, j, k) +
                                   w4 * QV(i, j, k) + w5 * QV(i - 1, j - 1, k - 1) + w6 * QV(i, j - 1, k - 1) +
                                   w7 * QV(i - 1, j - 1, k) + w8 * QV(i, j - 1, k)) / V;
            }
#undef QV
}

/* updateMaterialProps + smoothVF (VOF.f90:197-262): two cell -> vertex -> cell
 * passes of the volume fraction c give cs; rho = rho_g cs + (1 - cs) rho_l and
 * mu likewise (mu may be NULL). */
void or_material_props(const or_grid_desc* g, double rho_l, double rho_g, double mu_l, double mu_g,
                       const double* c, double* cs, double* rho, double* mu)
{
    lev_t L;
    lev_build(&L, g, 0);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const size_t n = (size_t)nx * ny * nz;
    double* qv = (double*)malloc(sizeof(double) * (size_t)(nx + 1) * (ny + 1) * (nz + 1));
    cell_to_vertex(&L, c, qv);
    vertex_to_cell(&L, qv, cs);
    cell_to_vertex(&L, cs, qv);
    vertex_to_cell(&L, qv, cs);
    for (size_t t = 0; t < n; ++t) {
        rho[t] = rho_g * cs[t] + (1.0 - cs[t]) * rho_l;
        if (mu) mu[t] = mu_g * cs[t] + (1.0 - cs[t]) * mu_l;
    }
    free(qv);
    lev_free(&L);
}
