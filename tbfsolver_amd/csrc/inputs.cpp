// inputs.cpp -- host-side input mirrors for tests and benchmarks: a seeded
// staggered velocity field and the bubble density field of the reference's
// two-phase set-up (initialConditions.f90:192-398, VOF.f90:197-262,
// interpolation.f90:188-276).
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace tbfg {

// counter-based generator: splitmix64 finaliser of (seed*phi + index + 1)
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline double uniform01(uint64_t seed, uint64_t idx)
{
    return (double)(mix64(seed * 0x9E3779B97F4A7C15ULL + idx + 1ULL) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double normal01(uint64_t seed, uint64_t m)
{
    const double u1 = uniform01(seed, 2 * m), u2 = uniform01(seed, 2 * m + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

int make_velocity(const tbfg_grid_desc* g, uint64_t seed, double* ux, double* uy, double* uz)
{
    TBFG_TRY(validate_grid(g));
    if (!ux || !uy || !uz) {
        set_error("tbfg_make_velocity: NULL argument");
        return TBFG_EINVAL;
    }
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long long t = 0; t < (long long)n; ++t) {
        ux[t] = normal01(seed, (uint64_t)t);
        uz[t] = normal01(seed, n + (uint64_t)t);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t t = i + (size_t)nx * (j + (size_t)(ny + 1) * k);
                uy[t] = (j == 0 || j == ny) ? 0.0 : normal01(seed, 2 * n + t);
            }
    return TBFG_OK;
}

static inline int wrapi(int i, int n) { return ((i % n) + n) % n; }

int make_bubble_density(const tbfg_grid_desc* g, int nb, double D, uint64_t seed, int nref, double rho_l,
                        double rho_g, double* rho, double* centres)
{
    TBFG_TRY(validate_grid(g));
    if (!rho || nb < 0 || !(D > 0.0) || nref < 1) {
        set_error("tbfg_make_bubble_density: bad arguments (nb=%d D=%g nref=%d)", nb, D, nref);
        return TBFG_EINVAL;
    }
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const double R = 0.5 * D, Lx = g->L[0], Ly = g->L[1], Lz = g->L[2];
    if (Ly <= 2 * R) {
        set_error("bubble diameter %g does not fit between the walls (Ly=%g)", D, Ly);
        return TBFG_EINVAL;
    }
    // centres: uniform in [0,Lx) x [R, Ly-R] x [0,Lz), rejected when closer than
    // 2R to an accepted centre (minimum-image distance in x and z)
    std::vector<double> cx, cy, cz;
    const uint64_t s2 = seed ^ 0x5DEECE66DULL;
    uint64_t ctr = 0;
    const uint64_t max_draws = 2000000;
    while ((int)cx.size() < nb) {
        if (ctr >= 3 * max_draws) {
            set_error("could not place %d non-overlapping bubbles of diameter %g", nb, D);
            return TBFG_EINVAL;
        }
        const double x = uniform01(s2, ctr++) * Lx;
        const double y = R + uniform01(s2, ctr++) * (Ly - 2 * R);
        const double z = uniform01(s2, ctr++) * Lz;
        bool ok = true;
        for (size_t b = 0; b < cx.size() && ok; ++b) {
            double dx = std::fabs(x - cx[b]), dz = std::fabs(z - cz[b]);
            dx = std::fmin(dx, Lx - dx);
            dz = std::fmin(dz, Lz - dz);
            const double dy = y - cy[b];
            if (dx * dx + dy * dy + dz * dz < 4 * R * R) ok = false;
        }
        if (ok) {
            cx.push_back(x);
            cy.push_back(y);
            cz.push_back(z);
        }
    }
    if (centres)
        for (int b = 0; b < nb; ++b) {
            centres[3 * b] = cx[b];
            centres[3 * b + 1] = cy[b];
            centres[3 * b + 2] = cz[b];
        }
    HostLevel L = make_level(g, 0);
    const Axis &X = L.ax[0], &Y = L.ax[1], &Z = L.ax[2];
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> c(n, 0.0);
    // init_Bubble_vf: centre test, then nref^3 sub-sampling where 1..7 of the 8
    // cell vertices are inside (initialConditions.f90:240-371).  Each bubble is
    // evaluated with the periodic image of its centre nearest to the cell.
    for (int b = 0; b < nb; ++b) {
        const double x0 = cx[b], y0 = cy[b], z0 = cz[b];
        const double dxu = Lx / nx, dzu = Lz / nz;
        int ilo = (int)std::floor((x0 - R) / dxu) - 1, ihi = (int)std::floor((x0 + R) / dxu) + 1;
        int klo = (int)std::floor((z0 - R) / dzu) - 1, khi = (int)std::floor((z0 + R) / dzu) + 1;
        if (ihi - ilo + 1 >= nx) {
            ilo = 0;
            ihi = nx - 1;
        }
        if (khi - klo + 1 >= nz) {
            klo = 0;
            khi = nz - 1;
        }
#pragma omp parallel for schedule(dynamic)
        for (int kk = klo; kk <= khi; ++kk) {
            const int k = wrapi(kk, nz) + 1;
            // periodic image of the cell nearest to the bubble centre
            const double zshift = std::nearbyint((z0 - Z.c(k)) / Lz) * Lz;
            for (int j = 1; j <= ny; ++j) {
                if (Y.f(j) < y0 - R - Y.df(j) || Y.f(j - 1) > y0 + R + Y.df(j)) continue;
                for (int ii = ilo; ii <= ihi; ++ii) {
                    const int i = wrapi(ii, nx) + 1;
                    const double xshift = std::nearbyint((x0 - X.c(i)) / Lx) * Lx;
                    const double xc = X.c(i) + xshift, yc = Y.c(j), zc = Z.c(k) + zshift;
                    double val = 0.0;
                    {
                        const double ddx = xc - x0, ddy = yc - y0, ddz = zc - z0;
                        if (ddx * ddx + ddy * ddy + ddz * ddz <= R * R) val = 1.0;
                    }
                    const double xf0 = X.f(i - 1) + xshift, xf1 = X.f(i) + xshift;
                    const double yf0 = Y.f(j - 1), yf1 = Y.f(j);
                    const double zf0 = Z.f(k - 1) + zshift, zf1 = Z.f(k) + zshift;
                    int count = 0;
                    for (int v = 0; v < 8; ++v) {
                        const double vx = (v & 1) ? xf1 : xf0, vy = (v & 2) ? yf1 : yf0, vz = (v & 4) ? zf1 : zf0;
                        const double ddx = vx - x0, ddy = vy - y0, ddz = vz - z0;
                        if (ddx * ddx + ddy * ddy + ddz * ddz <= R * R) count++;
                    }
                    if (count > 0 && count < 8) {
                        const double dx = X.df(i), dy = Y.df(j), dz = Z.df(k);
                        const double hx = dx / nref, hy = dy / nref, hz = dz / nref;
                        double acc = 0.0;
                        for (int kr = 0; kr < nref; ++kr) {
                            const double z = zf0 + 0.5 * hz + kr * hz - z0;
                            for (int jr = 0; jr < nref; ++jr) {
                                const double y = yf0 + 0.5 * hy + jr * hy - y0;
                                const double yz = y * y + z * z;
                                if (yz > R * R) continue;
                                for (int ir = 0; ir < nref; ++ir) {
                                    const double x = xf0 + 0.5 * hx + ir * hx - x0;
                                    if (x * x + yz <= R * R) acc += hx * hy * hz;
                                }
                            }
                        }
                        val = acc / (dx * dy * dz);
                    }
                    if (val > 0.0) {
                        double& cell = c[(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
                        cell = std::fmin(1.0, cell + val);
                    }
                }
            }
        }
    }
    // smoothVF (VOF.f90:235-262): two cell->vertex->cell passes, zero-gradient
    // ghosts in y and periodic ghosts in x/z between the passes.
    auto cget = [&](const std::vector<double>& f, int i, int j, int k) {
        i = wrapi(i - 1, nx) + 1;
        k = wrapi(k - 1, nz) + 1;
        j = j < 1 ? 1 : (j > ny ? ny : j);
        return f[(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))];
    };
    std::vector<double> vtx((size_t)(nx + 1) * (ny + 1) * (nz + 1));
    std::vector<double> cs = c;
    for (int pass = 0; pass < 2; ++pass) {
        // cellToVertex (interpolation.f90:188-226), vertices 0..n
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= nz; ++k)
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i <= nx; ++i) {
                    const double xv = X.f(i), yv = Y.f(j), zv = Z.f(k);
                    const double ax = (X.c(i + 1) - xv), bx = (xv - X.c(i));
                    const double ay = (Y.c(j + 1) - yv), by = (yv - Y.c(j));
                    const double az = (Z.c(k + 1) - zv), bz = (zv - Z.c(k));
                    const double V = (X.c(i + 1) - X.c(i)) * (Y.c(j + 1) - Y.c(j)) * (Z.c(k + 1) - Z.c(k));
                    const double v = ax * by * az * cget(cs, i, j + 1, k) + bx * by * az * cget(cs, i + 1, j + 1, k) +
                                     ax * by * bz * cget(cs, i, j + 1, k + 1) +
                                     bx * by * bz * cget(cs, i + 1, j + 1, k + 1) + ax * ay * az * cget(cs, i, j, k) +
                                     bx * ay * az * cget(cs, i + 1, j, k) + ax * ay * bz * cget(cs, i, j, k + 1) +
                                     bx * ay * bz * cget(cs, i + 1, j, k + 1);
                    vtx[i + (size_t)(nx + 1) * (j + (size_t)(ny + 1) * k)] = v / V;
                }
        // vertexToCell (interpolation.f90:231-276)
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    auto qv = [&](int a, int b2, int c2) { return vtx[a + (size_t)(nx + 1) * (b2 + (size_t)(ny + 1) * c2)]; };
                    const double x = X.c(i), y = Y.c(j), z = Z.c(k);
                    const double ax = X.f(i) - x, bx = x - X.f(i - 1);
                    const double ay = y - Y.f(j - 1), by = Y.f(j) - y;
                    const double az = Z.f(k) - z, bz = z - Z.f(k - 1);
                    const double V = (X.f(i) - X.f(i - 1)) * (Y.f(j) - Y.f(j - 1)) * (Z.f(k) - Z.f(k - 1));
                    const double v = ax * ay * az * qv(i - 1, j, k - 1) + bx * ay * az * qv(i, j, k - 1) +
                                     ax * ay * bz * qv(i - 1, j, k) + bx * ay * bz * qv(i, j, k) +
                                     ax * by * az * qv(i - 1, j - 1, k - 1) + bx * by * az * qv(i, j - 1, k - 1) +
                                     ax * by * bz * qv(i - 1, j - 1, k) + bx * by * bz * qv(i, j - 1, k);
                    cs[(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = v / V;
                }
    }
    // updateMaterialProps (VOF.f90:222-223)
    for (size_t t = 0; t < n; ++t) rho[t] = rho_g * cs[t] + (1.0 - cs[t]) * rho_l;
    return TBFG_OK;
}

}  // namespace tbfg
