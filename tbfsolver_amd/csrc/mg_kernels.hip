// mg_kernels.hip -- CDNA4 kernels of the multigrid-preconditioned CG path:
// 7-point variable-coefficient operator, red/black Gauss-Seidel half-sweeps,
// fused residual+restriction, prolongation+correction, PCG vector updates with
// fused deterministic reductions, and the single-workgroup tail kernel that
// runs every small level (and the coarsest "exact" solve) in one launch.
#include "mg_kernels.hpp"
#include "mg_launch.hpp"
#include "reduce.hpp"
#include "smooth.hpp"
#include "tune.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace tbfg {

// ------------------------------------------------------------ device helpers

__device__ __forceinline__ int wrap_up(int i, int n) { return i + 1 < n ? i + 1 : 0; }
__device__ __forceinline__ int wrap_dn(int i, int n) { return i > 0 ? i - 1 : n - 1; }
__device__ __forceinline__ int clamp_up(int j, int n) { return j + 1 < n ? j + 1 : j; }
__device__ __forceinline__ int clamp_dn(int j) { return j > 0 ? j - 1 : j; }

// (I: the index type, size_t for arrays in HBM, int for the LDS-resident tail)
template <class I>
__device__ __forceinline__ void neighbours(int i, int j, int k, int nx, int ny, int nz, I& c, I& e, I& w, I& n, I& s,
                                           I& f, I& b)
{
    const I row = (I)nx * ((I)j + (I)ny * k);
    c = row + i;
    e = row + wrap_up(i, nx);
    w = row + wrap_dn(i, nx);
    n = (I)nx * ((I)clamp_up(j, ny) + (I)ny * k) + i;
    s = (I)nx * ((I)clamp_dn(j) + (I)ny * k) + i;
    f = (I)i + (I)nx * ((I)j + (I)ny * wrap_up(k, nz));
    b = (I)i + (I)nx * ((I)j + (I)ny * wrap_dn(k, nz));
}

// A level whose tables live in LDS (the LDS-resident tail): the same fields as
// LevelDev with pointers typed as LDS addresses, so their loads are LDS reads
// with 32-bit addresses rather than generic (flat) loads.
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;
// A pointer typed as a global address: its loads count only against the
// vector-memory counter, so waiting for an LDS read does not also wait for them
// (a generic load counts against both).
template <class T>
using glb_ptr = __attribute__((address_space(1))) T*;
template <class T>
struct AxisLds {
    lds_ptr<const T> ip, im, rlo, rhi, plo, phi;
};
template <class T>
struct LevelLds {
    int nx, ny, nz;
    AxisLds<T> a[3];
};
// index type of a level's arrays
template <class LV>
struct IdxOf {
    using type = size_t;
};
template <class T>
struct IdxOf<LevelLds<T>> {
    using type = int;
};

// Block coordinates of a 3-D block grid launched as a 1-D grid: blocks go
// round-robin to the 8 XCDs, each with its own L2, so consecutive blocks are
// renumbered to fall on one XCD -- neighbouring tiles (whose halos overlap)
// then share that XCD's L2 instead of each fetching the halo from HBM.
struct Blk3 {
    int x, y, z;
};

__device__ __forceinline__ Blk3 xcd_block3(int gx, int gy)
{
    const unsigned nb = gridDim.x;
    unsigned id = blockIdx.x;
    if ((nb & 7u) == 0) id = (id & 7u) * (nb >> 3) + (id >> 3);
    Blk3 b;
    b.x = (int)(id % (unsigned)gx);
    const unsigned r = id / (unsigned)gx;
    b.y = (int)(r % (unsigned)gy);
    b.z = (int)(r / (unsigned)gy);
    return b;
}

// face coefficients a_f = 0.5*(beta_c + beta_nb) / (dc*df) (rbgs.f90:220-226)
template <class T>
struct Coef {
    T aR, aL, aT, aB, aF, aK;
};

template <class T, class LV, class PB, class I>
__device__ __forceinline__ Coef<T> coefs(const LV& L, PB __restrict__ beta, int i, int j, int k, I c, I e, I w, I n,
                                         I s, I f, I b)
{
    const T bc = (T)beta[c];
    Coef<T> a;
    a.aR = (T)0.5 * ((T)beta[e] + bc) * L.a[0].ip[i];
    a.aL = (T)0.5 * (bc + (T)beta[w]) * L.a[0].im[i];
    a.aT = (T)0.5 * ((T)beta[n] + bc) * L.a[1].ip[j];
    a.aB = (T)0.5 * (bc + (T)beta[s]) * L.a[1].im[j];
    a.aF = (T)0.5 * ((T)beta[f] + bc) * L.a[2].ip[k];
    a.aK = (T)0.5 * (bc + (T)beta[b]) * L.a[2].im[k];
    return a;
}

// (A p)_c, pcg.f90:171-176
template <class T, class PP, class I>
__device__ __forceinline__ T apply_cell(const Coef<T>& a, PP __restrict__ p, I c, I e, I w, I n, I s, I f, I b)
{
    const T pc = (T)p[c];
    return a.aR * ((T)p[e] - pc) - a.aL * (pc - (T)p[w]) + a.aT * ((T)p[n] - pc) - a.aB * (pc - (T)p[s]) +
           a.aF * ((T)p[f] - pc) - a.aK * (pc - (T)p[b]);
}

// ---------------------------------------------------------- CG-path kernels

// q = A p (computeQv, pcg.f90:144-183), FP64, no reduction.
__global__ void k_apply_op(LevelDev<double> L, const double* __restrict__ beta, const double* __restrict__ p,
                           double* __restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    size_t c, e, w, n, s, f, b;
    neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, n, s, f, b);
    Coef<double> a = coefs<double>(L, beta, i, j, k, c, e, w, n, s, f, b);
    q[c] = apply_cell<double>(a, p, c, e, w, n, s, f, b);
}

// qv = A d and alpha = delta / (d . qv) (pcg.f90:144-224).  A thread owns a
// column segment of kOpPlanes cells: the column's d and beta are loaded once
// for the segment and its two neighbours in z (instead of three times), the
// face factors of its x and y once, so a cell costs the four in-plane
// neighbours of d and beta plus its share of the column.
template <int kOpPlanes>
__global__ void __launch_bounds__(512) k_op_dot(LevelDev<double> L, const double* __restrict__ beta,
                                                const double* __restrict__ d, double* __restrict__ qv, PcgState* st,
                                                Reducer red)
{
    if (st->done) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const Blk3 bk = xcd_block3((nx + blockDim.x - 1) / blockDim.x, (ny + blockDim.y - 1) / blockDim.y);
    const int i = bk.x * blockDim.x + threadIdx.x;
    const int j = bk.y * blockDim.y + threadIdx.y;
    const int k0 = bk.z * kOpPlanes;
    double v[1] = {0.0};
    if (i < nx && j < ny) {
        const double ipx = L.a[0].ip[i], imx = L.a[0].im[i];
        const double ipy = L.a[1].ip[j], imy = L.a[1].im[j];
        const size_t plane = (size_t)nx * ny;
        const size_t col = (size_t)nx * j + i;
        // in-plane neighbours (x periodic, y walls: the ghost is the cell itself)
        const size_t oe = (size_t)nx * j + (i + 1 < nx ? i + 1 : 0);
        const size_t ow = (size_t)nx * j + (i > 0 ? i - 1 : nx - 1);
        const size_t on = (size_t)nx * (j + 1 < ny ? j + 1 : j) + i;
        const size_t os = (size_t)nx * (j > 0 ? j - 1 : j) + i;
        double dcol[kOpPlanes + 2], bcol[kOpPlanes + 2];
#pragma unroll
        for (int u = 0; u < kOpPlanes + 2; ++u) {
            int k = k0 - 1 + u;
            k = k < 0 ? k + nz : (k >= nz ? k - nz : k);
            dcol[u] = d[plane * k + col];
            bcol[u] = beta[plane * k + col];
        }
#pragma unroll
        for (int u = 1; u <= kOpPlanes; ++u) {
            const int k = k0 + u - 1;
            if (k < nz) {
                const size_t pk = plane * k;
                const double bc = bcol[u];
                const double aR = 0.5 * (beta[pk + oe] + bc) * ipx;
                const double aL = 0.5 * (bc + beta[pk + ow]) * imx;
                const double aT = 0.5 * (beta[pk + on] + bc) * ipy;
                const double aB = 0.5 * (bc + beta[pk + os]) * imy;
                const double aF = 0.5 * (bcol[u + 1] + bc) * L.a[2].ip[k];
                const double aK = 0.5 * (bc + bcol[u - 1]) * L.a[2].im[k];
                const double dc = dcol[u];
                const double out = aR * (d[pk + oe] - dc) - aL * (dc - d[pk + ow]) + aT * (d[pk + on] - dc) -
                                   aB * (dc - d[pk + os]) + aF * (dcol[u + 1] - dc) - aK * (dc - dcol[u - 1]);
                qv[pk + col] = out;
                v[0] += dc * out;
            }
        }
    }
    block_partial<1>(v, red);
}

// The same product marching through planes: a workgroup of NT threads owns a
// TX x (NT/TX) tile and a chunk of planes; each plane of d and beta (tile + one-cell ring)
// is staged once in LDS, loaded PD plane steps ahead in PD register sets, and
// the stencil of plane k reads planes k-1, k, k+1 from LDS (one barrier per
// plane).  The in-plane neighbours of a cell then cost no HBM traffic.
constexpr int kOpMaxChunk = 64;

template <int TX, int NT>
struct OpTile {
    static constexpr int TY = NT / TX;
    static constexpr int OX = TX + 2, OS = OX * (TY + 2);
    static constexpr int ON = (OS + NT - 1) / NT;
};

// With DIR (an iteration's direction, operator product and the previous
// iteration's pressure update in one pass): the staged plane is the new
// direction d = (z - m) + gamma d_old (pcg.f90:266-290; d_old from the other
// buffer of a pair, so no workgroup overwrites what another one reads), made
// from z and d_old as the plane is stashed, ring included; the owner of a cell
// writes d and psi += alpha_prev d_old (pcg.f90:228-237 of the iteration
// before, deferred to here).  init: the first direction d = z - m, no psi.
template <class TZ>
struct OpDir {
    const TZ* z;
    const double* dold;
    double* dnew;
    double* psi;
    int init;
};

template <int TX, int NT, int PD, class TZ, bool DIR>
__global__ void __launch_bounds__(NT) k_op_march(LevelDev<double> L, const double* __restrict__ beta,
                                                  const double* __restrict__ d, double* __restrict__ qv,
                                                  PcgState* st, Reducer red, int kc, int tiles_x, int ntiles,
                                                  OpDir<TZ> od)
{
    using Sh = OpTile<TX, NT>;
    constexpr int TY = Sh::TY, OX = Sh::OX, OS = Sh::OS, ON = Sh::ON;
    __shared__ double sD[4 * OS];
    __shared__ double sB[4 * OS];
    __shared__ double sTz[2 * kOpMaxChunk];
    if (st->done) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    unsigned id = blockIdx.x;
    if ((gridDim.x & 7u) == 0) id = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int tile = (int)(id % (unsigned)ntiles);
    const int chunk = (int)(id / (unsigned)ntiles);
    const int k0 = min(chunk * kc, nz);
    const int k1 = min(nz, k0 + kc);
    const int tx0 = (tile % tiles_x) * TX;
    const int ty0 = (tile / tiles_x) * TY;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int gx = tx0 + tx, gy = ty0 + ty;  // TX divides nx, TY divides ny
    for (int t = threadIdx.x; t < k1 - k0; t += NT) {
        sTz[t] = L.a[2].ip[k0 + t];
        sTz[kOpMaxChunk + t] = L.a[2].im[k0 + t];
    }
    int off[ON];
#pragma unroll
    for (int m = 0; m < ON; ++m) {
        const int t = min((int)threadIdx.x + m * NT, OS - 1);
        const int yy = t / OX, xx = t - yy * OX;
        const int jj = ty0 + yy - 1, ii = tx0 + xx - 1;
        off[m] = nx * (jj < 0 ? 0 : (jj >= ny ? ny - 1 : jj)) + (ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii));
    }
    const size_t plane = (size_t)nx * ny;
    const size_t own = (size_t)nx * gy + gx;
    // the iteration's scalars (written by the kernels before this one)
    const double zmean = DIR ? st->zmean : 0.0, gamma = DIR ? st->gamma : 0.0, alpha = DIR ? st->alpha : 0.0;
    const bool init = DIR && od.init;
    auto modz = [&](int k) { return ((k % nz) + nz) % nz; };
    auto slot = [&](int k) { return (k - k0 + 4) & 3; };
    // a register set holds one plane: d (or z and d_old) and beta; with DIR
    // also psi and d_old of this thread's own cell on the plane before (o2)
    // (plane k1 is the last one a chunk reads; the prefetch of the last steps
    // loads it again instead of the planes after it: a cache hit, not HBM traffic)
    auto fetch = [&](double(&rd)[ON], TZ(&rz)[ON], double(&rb)[ON], double(&o2)[2], int kp) {
        kp = min(kp, k1);
        const size_t o = plane * modz(kp);
#pragma unroll
        for (int m = 0; m < ON; ++m) {
            if (DIR) {
                rz[m] = od.z[o + off[m]];
                rd[m] = init ? 0.0 : od.dold[o + off[m]];
            } else {
                rd[m] = d[o + off[m]];
            }
            rb[m] = beta[o + off[m]];
        }
        if (DIR) {
            const size_t oo = plane * modz(kp - 1) + own;
            o2[0] = init ? 0.0 : od.psi[oo];
            o2[1] = init ? 0.0 : od.dold[oo];
        }
    };
    auto stash = [&](const double(&rd)[ON], const TZ(&rz)[ON], const double(&rb)[ON], int k) {
        double* D = sD + slot(k) * OS;
        double* B = sB + slot(k) * OS;
#pragma unroll
        for (int m = 0; m < ON; ++m) {
            const int t = min((int)threadIdx.x + m * NT, OS - 1);
            double dv = rd[m];
            if (DIR) {  // k_pcg_dir's arithmetic
                const double zc = (double)rz[m] - zmean;
                dv = init ? zc : zc + gamma * rd[m];
            }
            D[t] = dv;
            B[t] = rb[m];
        }
    };
    double acc[1] = {0.0};
    if (k0 < k1) {
        const double ipx = L.a[0].ip[gx], imx = L.a[0].im[gx];
        const double ipy = L.a[1].ip[gy], imy = L.a[1].im[gy];
        const int ic = (ty + 1) * OX + tx + 1;
        {
            double rd[ON], rb[ON], o2[2];
            TZ rz[ON];
            fetch(rd, rz, rb, o2, k0 - 1);
            stash(rd, rz, rb, k0 - 1);
            fetch(rd, rz, rb, o2, k0);
            stash(rd, rz, rb, k0);
        }
        double rds[PD][ON], rbs[PD][ON], o2s[PD][2];
        TZ rzs[PD][ON];
#pragma unroll
        for (int u = 0; u < PD; ++u) {
            fetch(rds[u], rzs[u], rbs[u], o2s[u], k0 + 1 + u);
            __builtin_amdgcn_sched_barrier(0);
        }
        double* out = qv + own;
        auto step = [&](int k, double(&rd)[ON], TZ(&rz)[ON], double(&rb)[ON], double(&o2)[2]) {
            stash(rd, rz, rb, k + 1);
            const double opsi = o2[0], odold = o2[1];  // the own cell on plane k
            fetch(rd, rz, rb, o2, k + 1 + PD);  // (past the chunk end: plane k1 again, not used)
            __syncthreads();
            if (k < k1) {
                const double* D0 = sD + slot(k) * OS;
                const double* Dp = sD + slot(k + 1) * OS;
                const double* Dm = sD + slot(k - 1) * OS;
                const double* B0 = sB + slot(k) * OS;
                const double* Bp = sB + slot(k + 1) * OS;
                const double* Bm = sB + slot(k - 1) * OS;
                // face coefficients (rbgs.f90:220-226) and (A d)_c (pcg.f90:171-176), as apply_cell
                const double bc = B0[ic];
                const double aR = 0.5 * (B0[ic + 1] + bc) * ipx;
                const double aL = 0.5 * (bc + B0[ic - 1]) * imx;
                const double aT = 0.5 * (B0[ic + OX] + bc) * ipy;
                const double aB = 0.5 * (bc + B0[ic - OX]) * imy;
                const double aF = 0.5 * (Bp[ic] + bc) * sTz[k - k0];
                const double aK = 0.5 * (bc + Bm[ic]) * sTz[kOpMaxChunk + k - k0];
                const double dc = D0[ic];
                const double v = aR * (D0[ic + 1] - dc) - aL * (dc - D0[ic - 1]) + aT * (D0[ic + OX] - dc) -
                                 aB * (dc - D0[ic - OX]) + aF * (Dp[ic] - dc) - aK * (dc - Dm[ic]);
                out[plane * k] = v;
                acc[0] += dc * v;
                if (DIR) {
                    od.dnew[plane * k + own] = dc;
                    if (!init) od.psi[plane * k + own] = opsi + alpha * odold;  // k_pcg_update's arithmetic
                }
            }
        };
        for (int k = k0; k < k1; k += PD) {
#pragma unroll
            for (int u = 0; u < PD; ++u) step(k + u, rds[u], rzs[u], rbs[u], o2s[u]);
        }
    }
    block_partial<1>(acc, red);
}

// psi += alpha d of the last iteration that ran (its direction is in buffer
// iter & 1 of the pair); nothing after a non-finite value or before any iteration
__global__ void k_psi_last(size_t ncell, double* __restrict__ psi, const double* __restrict__ d0,
                           const double* __restrict__ d1, const PcgState* st)
{
    const int it = st->iter;
    if (it == 0 || st->status == TBFG_ENUMERIC) return;
    const double alpha = st->alpha;
    const double* d = (it & 1) ? d1 : d0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) psi[t] = psi[t] + alpha * d[t];
}

__global__ void k_fin_alpha(unsigned nblocks, PcgState* st, Reducer red)
{
    if (st->done) return;
    double tot[1];
    final_sum<1>(red, nblocks, tot);
    if (threadIdx.x == 0) {
        const double den = tot[0];
        if (!(den != 0.0) || !isfinite(den) || !isfinite(st->delta)) {
            st->done = 1;
            st->status = TBFG_ENUMERIC;
            st->alpha = 0.0;
        } else {
            st->alpha = st->delta / den;
        }
    }
}

// r = q - A psi with ||r||^2, ||q||^2 (computeResiduals, rbgs.f90:276-333, used at pcg.f90:116).
// r32 (optional): an FP32 copy of r, the right-hand side of the FP32 multigrid.
template <class TR>
__global__ void k_residual_init(LevelDev<double> L, const double* __restrict__ beta, const double* __restrict__ q,
                                const double* __restrict__ psi, double* __restrict__ r, TR* __restrict__ r32,
                                PcgState* st, Reducer red)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    double v[2] = {0.0, 0.0};
    if (i < L.nx && j < L.ny) {
        const int kc = (L.nz + gridDim.z - 1) / gridDim.z;
        const int k1 = min(L.nz, (int)(blockIdx.z + 1) * kc);
        for (int k = blockIdx.z * kc; k < k1; ++k) {
            size_t c, e, w, n, s, f, b;
            neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, n, s, f, b);
            Coef<double> a = coefs<double>(L, beta, i, j, k, c, e, w, n, s, f, b);
            const double qc = q[c];
            const double rc = qc - apply_cell<double>(a, psi, c, e, w, n, s, f, b);
            r[c] = rc;
            if (r32) r32[c] = store_as<TR>((float)rc);
            v[0] += rc * rc;
            v[1] += qc * qc;
        }
    }
    block_partial<2>(v, red);
}

// The same residual with k_op_march's plane march (psi and beta of a plane
// staged once in LDS with a one-cell ring, two plane steps of loads in
// flight), so in-plane neighbours cost no HBM traffic: the column form above
// re-reads them from HBM when they miss in L2.
template <int TX, int NT, class TR>
__global__ void __launch_bounds__(NT) k_resid_march(LevelDev<double> L, const double* __restrict__ beta,
                                                     const double* __restrict__ q, const double* __restrict__ psi,
                                                     double* __restrict__ r, TR* __restrict__ r32, Reducer red,
                                                     int kc, int tiles_x, int ntiles)
{
    using Sh = OpTile<TX, NT>;
    constexpr int TY = Sh::TY, OX = Sh::OX, OS = Sh::OS, ON = Sh::ON;
    constexpr int PD = 2;
    __shared__ double sD[4 * OS];
    __shared__ double sB[4 * OS];
    __shared__ double sTz[2 * kOpMaxChunk];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    unsigned id = blockIdx.x;
    if ((gridDim.x & 7u) == 0) id = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int tile = (int)(id % (unsigned)ntiles);
    const int chunk = (int)(id / (unsigned)ntiles);
    const int k0 = min(chunk * kc, nz);
    const int k1 = min(nz, k0 + kc);
    const int tx0 = (tile % tiles_x) * TX;
    const int ty0 = (tile / tiles_x) * TY;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int gx = tx0 + tx, gy = ty0 + ty;  // TX divides nx, TY divides ny
    for (int t = threadIdx.x; t < k1 - k0; t += NT) {
        sTz[t] = L.a[2].ip[k0 + t];
        sTz[kOpMaxChunk + t] = L.a[2].im[k0 + t];
    }
    int off[ON];
#pragma unroll
    for (int m = 0; m < ON; ++m) {
        const int t = min((int)threadIdx.x + m * NT, OS - 1);
        const int yy = t / OX, xx = t - yy * OX;
        const int jj = ty0 + yy - 1, ii = tx0 + xx - 1;
        off[m] = nx * (jj < 0 ? 0 : (jj >= ny ? ny - 1 : jj)) + (ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii));
    }
    const size_t plane = (size_t)nx * ny;
    const size_t own = (size_t)nx * gy + gx;
    auto modz = [&](int k) { return ((k % nz) + nz) % nz; };
    auto slot = [&](int k) { return (k - k0 + 4) & 3; };
    auto fetch = [&](double(&rd)[ON], double(&rb)[ON], double& rq, int kp) {
        kp = min(kp, k1);  // past the chunk's last plane: plane k1 again (a cache hit, unused)
        const size_t o = plane * modz(kp);
#pragma unroll
        for (int m = 0; m < ON; ++m) {
            rd[m] = psi[o + off[m]];
            rb[m] = beta[o + off[m]];
        }
        rq = q[plane * modz(kp - 1) + own];  // q of this thread's cell on the plane before
    };
    auto stash = [&](const double(&rd)[ON], const double(&rb)[ON], int k) {
        double* D = sD + slot(k) * OS;
        double* B = sB + slot(k) * OS;
#pragma unroll
        for (int m = 0; m < ON; ++m) {
            const int t = min((int)threadIdx.x + m * NT, OS - 1);
            D[t] = rd[m];
            B[t] = rb[m];
        }
    };
    double acc[2] = {0.0, 0.0};
    if (k0 < k1) {
        const double ipx = L.a[0].ip[gx], imx = L.a[0].im[gx];
        const double ipy = L.a[1].ip[gy], imy = L.a[1].im[gy];
        const int ic = (ty + 1) * OX + tx + 1;
        {
            double rd[ON], rb[ON], rq;
            fetch(rd, rb, rq, k0 - 1);
            stash(rd, rb, k0 - 1);
            fetch(rd, rb, rq, k0);
            stash(rd, rb, k0);
        }
        double rds[PD][ON], rbs[PD][ON], rqs[PD];
#pragma unroll
        for (int u = 0; u < PD; ++u) {
            fetch(rds[u], rbs[u], rqs[u], k0 + 1 + u);
            __builtin_amdgcn_sched_barrier(0);
        }
        auto step = [&](int k, double(&rd)[ON], double(&rb)[ON], double& rq) {
            stash(rd, rb, k + 1);
            const double qc = rq;  // q of the own cell on plane k
            fetch(rd, rb, rq, k + 1 + PD);
            __syncthreads();
            if (k < k1) {
                const double* D0 = sD + slot(k) * OS;
                const double* Dp = sD + slot(k + 1) * OS;
                const double* Dm = sD + slot(k - 1) * OS;
                const double* B0 = sB + slot(k) * OS;
                const double* Bp = sB + slot(k + 1) * OS;
                const double* Bm = sB + slot(k - 1) * OS;
                // coefficients and (A psi)_c as apply_cell (rbgs.f90:300-315, pcg.f90:171-176)
                const double bc = B0[ic];
                const double aR = 0.5 * (B0[ic + 1] + bc) * ipx;
                const double aL = 0.5 * (bc + B0[ic - 1]) * imx;
                const double aT = 0.5 * (B0[ic + OX] + bc) * ipy;
                const double aB = 0.5 * (bc + B0[ic - OX]) * imy;
                const double aF = 0.5 * (Bp[ic] + bc) * sTz[k - k0];
                const double aK = 0.5 * (bc + Bm[ic]) * sTz[kOpMaxChunk + k - k0];
                const double dc = D0[ic];
                const double v = aR * (D0[ic + 1] - dc) - aL * (dc - D0[ic - 1]) + aT * (D0[ic + OX] - dc) -
                                 aB * (dc - D0[ic - OX]) + aF * (Dp[ic] - dc) - aK * (dc - Dm[ic]);
                const double rc = qc - v;
                r[plane * k + own] = rc;
                if (r32) r32[plane * k + own] = store_as<TR>((float)rc);
                acc[0] += rc * rc;
                acc[1] += qc * qc;
            }
        };
        for (int k = k0; k < k1; k += PD) {
#pragma unroll
            for (int u = 0; u < PD; ++u) step(k + u, rds[u], rbs[u], rqs[u]);
        }
    }
    block_partial<2>(acc, red);
}

// a solve's state before its first kernel: every field zero but the parameters
// (a kernel rather than a copy from host memory, so that a solve can be captured
// in a graph and replayed)
__global__ void k_state_init(PcgState* st, double tol, int32_t max_iter, int32_t flexible)
{
    if (threadIdx.x != 0) return;
    PcgState z{};
    z.tol = tol;
    z.max_iter = max_iter;
    z.flexible = flexible;
    *st = z;
}

__global__ void k_fin_init(unsigned nblocks, PcgState* st, Reducer red)
{
    double tot[2];
    final_sum<2>(red, nblocks, tot);
    if (threadIdx.x == 0) {
        st->rr = tot[0];
        st->qq = tot[1];
        st->iter = 0;
        st->status = TBFG_OK;
        st->done = 0;
        const double res = sqrt(tot[0]) / sqrt(tot[1] + 2.2250738585072014e-308);
        st->res = res;
        if (!isfinite(res)) {
            st->done = 1;
            st->status = TBFG_ENUMERIC;
        } else if (res <= st->tol) {
            st->done = 1;
        } else if (st->max_iter <= 0) {
            st->done = 1;
            st->status = TBFG_ENOTCONVERGED;
        }
    }
}

// psi += alpha d; r -= alpha qv; ||r|| and the stop test (pcg.f90:228-262,331-355,359-429).
template <class TR>
__global__ void k_pcg_update(size_t ncell, double* __restrict__ psi, double* __restrict__ r, TR* __restrict__ r32,
                             const double* __restrict__ d, const double* __restrict__ qv, PcgState* st, Reducer red)
{
    if (st->done) return;
    const double alpha = st->alpha;
    double v[1] = {0.0};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) {
        if (psi) psi[t] = psi[t] + alpha * d[t];  // (NULL: deferred to the next operator pass)
        const double rn = r[t] - alpha * qv[t];
        r[t] = rn;
        if (r32) r32[t] = store_as<TR>((float)rn);
        v[0] += rn * rn;
    }
    block_partial<1>(v, red);
}

__global__ void k_fin_update(unsigned nblocks, PcgState* st, Reducer red)
{
    if (st->done) return;
    double tot[1];
    final_sum<1>(red, nblocks, tot);
    if (threadIdx.x == 0) {
        st->rr = tot[0];
        st->iter += 1;
        const double res = sqrt(tot[0]) / sqrt(st->qq + 2.2250738585072014e-308);
        st->res = res;
        if (!isfinite(res)) {
            st->done = 1;
            st->status = TBFG_ENUMERIC;
        } else if (res <= st->tol) {
            st->done = 1;
            st->status = TBFG_OK;
        } else if (st->iter >= st->max_iter) {
            st->done = 1;
            st->status = TBFG_ENOTCONVERGED;
        }
    }
}

// r.z, qv.z, sum z, sum r, sum qv in one pass; the preconditioned residual is
// used with its arithmetic mean removed (rbgs.f90:403-446 leaves z mean-free):
// z' = z - m, r.z' = r.z - m sum r, qv.z' = qv.z - m sum qv.
template <class TZ>
__global__ void k_pcg_dots(size_t ncell, const double* __restrict__ r, const TZ* __restrict__ z,
                           const double* __restrict__ qv, int init, PcgState* st, Reducer red)
{
    if (st->done) return;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) {
        const double zc = (double)z[t];
        const double rc = r[t];
        v[0] += rc * zc;
        v[2] += zc;
        v[3] += rc;
        if (!init) {
            const double qc = qv[t];
            v[1] += qc * zc;
            v[4] += qc;
        }
    }
    block_partial<5>(v, red);
}

__global__ void k_fin_dots(unsigned nblocks, size_t ncell, int init, PcgState* st, Reducer red)
{
    if (st->done) return;
    double tot[5];
    final_sum<5>(red, nblocks, tot);
    if (threadIdx.x == 0) {
        const double m = tot[2] / (double)ncell;
        const double rz = tot[0] - m * tot[3];
        st->zmean = m;
        if (init) {
            st->delta = rz;
            st->gamma = 0.0;
        } else {
            const double qz = tot[1] - m * tot[4];
            const double g = st->flexible ? (-st->alpha * qz / st->delta) : (rz / st->delta);
            st->gamma = g;
            st->delta = rz;
            if (!isfinite(g) || !isfinite(rz)) {
                st->done = 1;
                st->status = TBFG_ENUMERIC;
            }
        }
    }
}

// d = (z - m) + gamma d (pcg.f90:266-290 with d0 <- d, pcg.f90:127,133-134)
template <class TZ>
__global__ void k_pcg_dir(size_t ncell, const TZ* __restrict__ z, double* __restrict__ d, int init,
                          const PcgState* st)
{
    if (st->done) return;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const double zc = (double)z[t] - st->zmean;
    d[t] = init ? zc : zc + st->gamma * d[t];
}

// ------------------------------------------------------------- MG kernels

// one red or black half-sweep of iterateGS (rbgs.f90:205-259) on a level.
// FIRST: the level starts from p = 0 (rbgs.f90:196-198 / pcg.f90:118), so the
// red pass reads no p and the black pass uses 0 as its own wall ghost.
template <class T, class TQ, int COLOUR, bool FIRST, class LV, class PB, class PQ, class PP>
__device__ __forceinline__ void rbgs_cell(const LV& L, PB __restrict__ beta, PQ __restrict__ q, PP __restrict__ p,
                                          int h, int j, int k)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int i = ((COLOUR + j + k + 1) & 1) + 2 * h;
    typename IdxOf<LV>::type c, e, w, n, s, f, b;
    neighbours(i, j, k, nx, ny, nz, c, e, w, n, s, f, b);
    Coef<T> a = coefs<T>(L, beta, i, j, k, c, e, w, n, s, f, b);
    const T den = a.aR + a.aL + a.aT + a.aB + a.aF + a.aK;
    const T qc = (T)q[c];
    T num;
    if (FIRST && COLOUR == RED) {
        num = -qc;
    } else {
        const T self = FIRST ? (T)0 : p[c];
        const T pn = (n == c) ? self : p[n];
        const T ps = (s == c) ? self : p[s];
        num = -qc + (a.aR * p[e] + a.aL * p[w]) + (a.aT * pn + a.aB * ps) + (a.aF * p[f] + a.aK * p[b]);
    }
    p[c] = smooth_div(num, den);
}

template <class T, class TQ, int COLOUR, bool FIRST>
__global__ void k_rbgs(LevelDev<T> L, const T* __restrict__ beta, const TQ* __restrict__ q, T* __restrict__ p,
                       const int* done)
{
    if (done && *done) return;
    const Blk3 bk = xcd_block3(((L.nx >> 1) + blockDim.x - 1) / blockDim.x, (L.ny + blockDim.y - 1) / blockDim.y);
    const int h = bk.x * blockDim.x + threadIdx.x;
    const int j = bk.y * blockDim.y + threadIdx.y;
    const int k = bk.z;
    if (h >= (L.nx >> 1) || j >= L.ny) return;
    rbgs_cell<T, TQ, COLOUR, FIRST>(L, beta, q, p, h, j, k);
}

// fused residual (rbgs.f90:300-315) and restriction (interpolation.f90:97-118):
// one thread per coarse cell, qc = sum_{children} wx wy wz (q - A p).  The
// plane loop stays rolled: all eight children in flight need more registers
// than a 1024-thread workgroup has, and the spill costs more than the overlap.
// resid_restrict_plane: the four children in fine plane 2K + dk, added to acc.
template <class T, class TQ, class LV, class PB, class PQ, class PP>
__device__ __forceinline__ T resid_restrict_plane(const LV& L, PB __restrict__ beta, PQ __restrict__ q,
                                                  PP __restrict__ p, int I, int J, int K, int dk, T acc)
{
    const int k = 2 * K + dk;
    const T wz = dk ? L.a[2].rhi[K] : L.a[2].rlo[K];
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
        const int j = 2 * J + dj;
        const T wy = dj ? L.a[1].rhi[J] : L.a[1].rlo[J];
#pragma unroll
        for (int di = 0; di < 2; ++di) {
            const int i = 2 * I + di;
            const T wx = di ? L.a[0].rhi[I] : L.a[0].rlo[I];
            typename IdxOf<LV>::type c, e, w, n, s, f, b;
            neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, n, s, f, b);
            Coef<T> a = coefs<T>(L, beta, i, j, k, c, e, w, n, s, f, b);
            const T res = (T)q[c] - apply_cell<T>(a, p, c, e, w, n, s, f, b);
            acc += (wx * wy * wz) * res;
        }
    }
    return acc;
}
template <class T, class TQ, class LV, class PB, class PQ, class PP>
__device__ __forceinline__ T resid_restrict_cell(const LV& L, PB __restrict__ beta, PQ __restrict__ q,
                                                 PP __restrict__ p, int I, int J, int K)
{
    T acc = 0;
#pragma unroll 1
    for (int dk = 0; dk < 2; ++dk) acc = resid_restrict_plane<T, TQ>(L, beta, q, p, I, J, K, dk, acc);
    return acc;
}

// The same for a column segment of kRrPlanes coarse cells per thread: the 2x2
// fine columns under the coarse column are loaded once for the segment (and
// the fine planes either side), the eight in-plane neighbours of the 2x2 block
// once per fine plane -- about 7 loads per fine cell instead of 20.
constexpr int kRrPlanes = 2;

template <class T>
__device__ __forceinline__ T cell_residual(T qc, T bc, T be, T bw, T bn, T bs, T bf, T bb, T pc, T pe, T pw, T pn,
                                           T ps, T pf, T pb, T ipx, T imx, T ipy, T imy, T ipz, T imz)
{
    // coefficients as coefs(), (A p)_c as apply_cell()
    const T aR = (T)0.5 * (be + bc) * ipx;
    const T aL = (T)0.5 * (bc + bw) * imx;
    const T aT = (T)0.5 * (bn + bc) * ipy;
    const T aB = (T)0.5 * (bc + bs) * imy;
    const T aF = (T)0.5 * (bf + bc) * ipz;
    const T aK = (T)0.5 * (bc + bb) * imz;
    return qc - (aR * (pe - pc) - aL * (pc - pw) + aT * (pn - pc) - aB * (pc - ps) + aF * (pf - pc) - aK * (pc - pb));
}

template <class T, class TQ, class TB>
__global__ void k_resid_restrict(LevelDev<T> L, const TB* __restrict__ beta, const TQ* __restrict__ q,
                                 const T* __restrict__ p, T* __restrict__ qc, const int* done)
{
    if (done && *done) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int ncx = nx >> 1, ncy = ny >> 1, ncz = nz >> 1;
    const Blk3 bk = xcd_block3((ncx + blockDim.x - 1) / blockDim.x, (ncy + blockDim.y - 1) / blockDim.y);
    const int I = bk.x * blockDim.x + threadIdx.x;
    const int J = bk.y * blockDim.y + threadIdx.y;
    const int K0 = bk.z * kRrPlanes;
    if (I >= ncx || J >= ncy) return;
    const int i0 = 2 * I, i1 = i0 + 1, j0 = 2 * J, j1 = j0 + 1;
    const int iw = i0 > 0 ? i0 - 1 : nx - 1, ie = i1 + 1 < nx ? i1 + 1 : 0;
    const int js = j0 > 0 ? j0 - 1 : j0, jn = j1 + 1 < ny ? j1 + 1 : j1;  // wall ghost = the cell itself
    const size_t plane = (size_t)nx * ny;
    // the 2x2 block (c_ij) and its eight in-plane neighbours
    const size_t c00 = (size_t)nx * j0 + i0, c10 = c00 + 1, c01 = c00 + nx, c11 = c01 + 1;
    const size_t w0 = (size_t)nx * j0 + iw, w1 = (size_t)nx * j1 + iw;
    const size_t e0 = (size_t)nx * j0 + ie, e1 = (size_t)nx * j1 + ie;
    const size_t s0 = (size_t)nx * js + i0, s1 = (size_t)nx * js + i1;
    const size_t n0 = (size_t)nx * jn + i0, n1 = (size_t)nx * jn + i1;
    const T ipx0 = L.a[0].ip[i0], imx0 = L.a[0].im[i0], ipx1 = L.a[0].ip[i1], imx1 = L.a[0].im[i1];
    const T ipy0 = L.a[1].ip[j0], imy0 = L.a[1].im[j0], ipy1 = L.a[1].ip[j1], imy1 = L.a[1].im[j1];
    const T wx0 = L.a[0].rlo[I], wx1 = L.a[0].rhi[I], wy0 = L.a[1].rlo[J], wy1 = L.a[1].rhi[J];
    constexpr int NF = 2 * kRrPlanes + 2;  // fine planes 2K0-1 .. 2K0+2kRrPlanes
    T P[NF][4], B[NF][4];
#pragma unroll
    for (int u = 0; u < NF; ++u) {
        int k = 2 * K0 - 1 + u;
        k = k < 0 ? k + nz : (k >= nz ? k - nz : k);
        const size_t pk = plane * k;
        P[u][0] = p[pk + c00], P[u][1] = p[pk + c10], P[u][2] = p[pk + c01], P[u][3] = p[pk + c11];
        B[u][0] = beta[pk + c00], B[u][1] = beta[pk + c10], B[u][2] = beta[pk + c01], B[u][3] = beta[pk + c11];
    }
#pragma unroll
    for (int kk = 0; kk < kRrPlanes; ++kk) {
        const int K = K0 + kk;
        if (K >= ncz) break;
        T acc = 0;
#pragma unroll
        for (int dk = 0; dk < 2; ++dk) {
            const int k = 2 * K + dk;
            const int u = 2 * kk + dk + 1;
            const size_t pk = plane * k;
            const T wz = dk ? L.a[2].rhi[K] : L.a[2].rlo[K];
            const T ipz = L.a[2].ip[k], imz = L.a[2].im[k];
            const T pw0 = p[pk + w0], pw1 = p[pk + w1], pe0 = p[pk + e0], pe1 = p[pk + e1];
            const T ps0 = p[pk + s0], ps1 = p[pk + s1], pn0 = p[pk + n0], pn1 = p[pk + n1];
            const T bw0 = beta[pk + w0], bw1 = beta[pk + w1], be0 = beta[pk + e0], be1 = beta[pk + e1];
            const T bs0 = beta[pk + s0], bs1 = beta[pk + s1], bn0 = beta[pk + n0], bn1 = beta[pk + n1];
            const T q00 = (T)q[pk + c00], q10 = (T)q[pk + c10], q01 = (T)q[pk + c01], q11 = (T)q[pk + c11];
            const T* Pc = P[u];
            const T* Bc = B[u];
            const T* Pf = P[u + 1];
            const T* Bf = B[u + 1];
            const T* Pb = P[u - 1];
            const T* Bb = B[u - 1];
            // children in the reference's order: dj outer, di inner
            const T r00 = cell_residual<T>(q00, Bc[0], Bc[1], bw0, Bc[2], bs0, Bf[0], Bb[0], Pc[0], Pc[1], pw0, Pc[2],
                                           ps0, Pf[0], Pb[0], ipx0, imx0, ipy0, imy0, ipz, imz);
            const T r10 = cell_residual<T>(q10, Bc[1], be0, Bc[0], Bc[3], bs1, Bf[1], Bb[1], Pc[1], pe0, Pc[0], Pc[3],
                                           ps1, Pf[1], Pb[1], ipx1, imx1, ipy0, imy0, ipz, imz);
            const T r01 = cell_residual<T>(q01, Bc[2], Bc[3], bw1, bn0, Bc[0], Bf[2], Bb[2], Pc[2], Pc[3], pw1, pn0,
                                           Pc[0], Pf[2], Pb[2], ipx0, imx0, ipy1, imy1, ipz, imz);
            const T r11 = cell_residual<T>(q11, Bc[3], be1, Bc[2], bn1, Bc[1], Bf[3], Bb[3], Pc[3], pe1, Pc[2], pn1,
                                           Pc[1], Pf[3], Pb[3], ipx1, imx1, ipy1, imy1, ipz, imz);
            acc += (wx0 * wy0 * wz) * r00;
            acc += (wx1 * wy0 * wz) * r10;
            acc += (wx0 * wy1 * wz) * r01;
            acc += (wx1 * wy1 * wz) * r11;
        }
        qc[I + (size_t)ncx * (J + (size_t)ncy * K)] = acc;
    }
}

// plain restriction of a field (beta, restrictField at multiGrid.f90:210)
template <class TF, class T>
__global__ void k_restrict(LevelDev<T> L, const TF* __restrict__ fine, T* __restrict__ coarse)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    const int nx = L.nx, ny = L.ny;
    const int ncx = nx >> 1, ncy = ny >> 1;
    if (I >= ncx || J >= ncy) return;
    T acc = 0;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const T w = (di ? L.a[0].rhi[I] : L.a[0].rlo[I]) * (dj ? L.a[1].rhi[J] : L.a[1].rlo[J]) *
                            (dk ? L.a[2].rhi[K] : L.a[2].rlo[K]);
                acc += w * (T)fine[(2 * I + di) + (size_t)nx * ((2 * J + dj) + (size_t)ny * (2 * K + dk))];
            }
    coarse[I + (size_t)ncx * (J + (size_t)ncy * K)] = acc;
}

// prolongation with coarse ghosts and correction p += P pc (multiGrid.f90:222-223).
template <class T, class LV, class PC, class PP>
__device__ __forceinline__ void prolong_cell(const LV& L, int ncx, int ncy, int ncz, PC __restrict__ pc,
                                             PP __restrict__ p, int i, int j, int k)
{
    using I = typename IdxOf<LV>::type;
    int il = ((i + 1) >> 1) - 1, jl = ((j + 1) >> 1) - 1, kl = ((k + 1) >> 1) - 1;
    int ih = il + 1, jh = jl + 1, kh = kl + 1;
    if (il < 0) il = ncx - 1;
    if (ih >= ncx) ih = 0;
    if (jl < 0) jl = 0;
    if (jh >= ncy) jh = ncy - 1;
    if (kl < 0) kl = ncz - 1;
    if (kh >= ncz) kh = 0;
    const T wxl = L.a[0].plo[i], wxh = L.a[0].phi[i];
    const T wyl = L.a[1].plo[j], wyh = L.a[1].phi[j];
    const T wzl = L.a[2].plo[k], wzh = L.a[2].phi[k];
    const I sxy = (I)ncx * ncy;
    const I r_ll = (I)ncx * jl + sxy * kl, r_hl = (I)ncx * jh + sxy * kl;
    const I r_lh = (I)ncx * jl + sxy * kh, r_hh = (I)ncx * jh + sxy * kh;
    const T v = wzl * (wyl * (wxl * pc[r_ll + il] + wxh * pc[r_ll + ih]) + wyh * (wxl * pc[r_hl + il] + wxh * pc[r_hl + ih])) +
                wzh * (wyl * (wxl * pc[r_lh + il] + wxh * pc[r_lh + ih]) + wyh * (wxl * pc[r_hh + il] + wxh * pc[r_hh + ih]));
    const I c = i + (I)L.nx * (j + (I)L.ny * k);
    p[c] = p[c] + v;
}

template <class T>
struct Pair;
template <>
struct Pair<float> {
    using type = float2;
};
template <>
struct Pair<double> {
    using type = double2;
};

// One thread per coarse cell: its 2x2x2 fine children all interpolate from the
// 3x3x3 coarse neighbourhood, loaded once (the weights and the order of the
// arithmetic are prolong_cell's), and each pair of x-children is one vector
// access.
template <class T>
__global__ void k_prolong_add(LevelDev<T> L, const T* __restrict__ pc, T* __restrict__ p, const int* done)
{
    if (done && *done) return;
    const int nx = L.nx, ny = L.ny;
    const int ncx = L.nx >> 1, ncy = L.ny >> 1, ncz = L.nz >> 1;
    const Blk3 bk = xcd_block3((ncx + blockDim.x - 1) / blockDim.x, (ncy + blockDim.y - 1) / blockDim.y);
    const int I = bk.x * blockDim.x + threadIdx.x;
    const int J = bk.y * blockDim.y + threadIdx.y;
    const int K = bk.z;
    if (I >= ncx || J >= ncy) return;
    // coarse neighbours: x, z periodic; y walls (the ghost is the wall cell)
    const int xs[3] = {I > 0 ? I - 1 : ncx - 1, I, I + 1 < ncx ? I + 1 : 0};
    const int ys[3] = {J > 0 ? J - 1 : 0, J, J + 1 < ncy ? J + 1 : ncy - 1};
    const int zs[3] = {K > 0 ? K - 1 : ncz - 1, K, K + 1 < ncz ? K + 1 : 0};
    T c[3][3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int e = 0; e < 3; ++e) c[a][b][e] = pc[xs[e] + (size_t)ncx * (ys[b] + (size_t)ncy * zs[a])];
    const T wxl0 = L.a[0].plo[2 * I], wxh0 = L.a[0].phi[2 * I];
    const T wxl1 = L.a[0].plo[2 * I + 1], wxh1 = L.a[0].phi[2 * I + 1];
#pragma unroll
    for (int dk = 0; dk < 2; ++dk) {
        const int k = 2 * K + dk;
        const T wzl = L.a[2].plo[k], wzh = L.a[2].phi[k];
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const int j = 2 * J + dj;
            const T wyl = L.a[1].plo[j], wyh = L.a[1].phi[j];
            // children (2I, j, k) and (2I+1, j, k): coarse x cells (I-1, I) and (I, I+1)
            const T (&zl)[3][3] = c[dk], (&zh)[3][3] = c[dk + 1];
            const T v0 = wzl * (wyl * (wxl0 * zl[dj][0] + wxh0 * zl[dj][1]) + wyh * (wxl0 * zl[dj + 1][0] + wxh0 * zl[dj + 1][1])) +
                         wzh * (wyl * (wxl0 * zh[dj][0] + wxh0 * zh[dj][1]) + wyh * (wxl0 * zh[dj + 1][0] + wxh0 * zh[dj + 1][1]));
            const T v1 = wzl * (wyl * (wxl1 * zl[dj][1] + wxh1 * zl[dj][2]) + wyh * (wxl1 * zl[dj + 1][1] + wxh1 * zl[dj + 1][2])) +
                         wzh * (wyl * (wxl1 * zh[dj][1] + wxh1 * zh[dj][2]) + wyh * (wxl1 * zh[dj + 1][1] + wxh1 * zh[dj + 1][2]));
            typename Pair<T>::type* q = reinterpret_cast<typename Pair<T>::type*>(p + 2 * I + (size_t)nx * (j + (size_t)ny * k));
            typename Pair<T>::type f = *q;
            f.x = f.x + v0;
            f.y = f.y + v1;
            *q = f;
        }
    }
}

template <class TS, class TD>
__global__ void k_convert(size_t n, const TS* __restrict__ src, TD* __restrict__ dst)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = store_as<TD>(src[t]);
}

// ------------------------------------------------------------ tail kernel

// The levels below ~32K cells are latency bound: every phase is a few
// microseconds of work, less than a kernel launch.  One workgroup runs all of
// them back to back with workgroup barriers; their arrays stay in L2, or in
// LDS (LDS = true: the level's pointers point into the workgroup's LDS and are
// used as LDS addresses).
// floor(x / d) by one multiply-high with m = ceil(2^32 / d): exact when x * d
// < 2^32 (the error term x (m d - 2^32) / (d 2^32) stays below 1/d); d = 1 is
// the identity
__device__ __forceinline__ int mdiv(int x, int d, uint32_t m) { return d == 1 ? x : (int)__umulhi((uint32_t)x, m); }

template <class T, bool LDS>
struct TailView;
template <class T>
struct TailView<T, false> {
    LevelDev<T> geo;
    T* p;
    T* q;
    const T* beta;
    uint32_t dm_nx, dm_ny, dm_hx, dm_hy;
    __device__ explicit TailView(const TailLevel<T>& lv)
        : geo(lv.geo), p(lv.p), q(lv.q), beta(lv.beta), dm_nx(lv.dm_nx), dm_ny(lv.dm_ny), dm_hx(lv.dm_hx),
          dm_hy(lv.dm_hy)
    {
    }
};
template <class T>
struct TailView<T, true> {
    LevelLds<T> geo;
    lds_ptr<T> p;
    lds_ptr<T> q;
    lds_ptr<const T> beta;
    uint32_t dm_nx, dm_ny, dm_hx, dm_hy;
    __device__ explicit TailView(const TailLevel<T>& lv)
    {
        dm_nx = lv.dm_nx;
        dm_ny = lv.dm_ny;
        dm_hx = lv.dm_hx;
        dm_hy = lv.dm_hy;
        geo.nx = lv.geo.nx;
        geo.ny = lv.geo.ny;
        geo.nz = lv.geo.nz;
        for (int a = 0; a < 3; ++a) {
            geo.a[a].ip = (lds_ptr<const T>)lv.geo.a[a].ip;
            geo.a[a].im = (lds_ptr<const T>)lv.geo.a[a].im;
            geo.a[a].rlo = (lds_ptr<const T>)lv.geo.a[a].rlo;
            geo.a[a].rhi = (lds_ptr<const T>)lv.geo.a[a].rhi;
            geo.a[a].plo = (lds_ptr<const T>)lv.geo.a[a].plo;
            geo.a[a].phi = (lds_ptr<const T>)lv.geo.a[a].phi;
        }
        p = (lds_ptr<T>)lv.p;
        q = (lds_ptr<T>)lv.q;
        beta = (lds_ptr<const T>)lv.beta;
    }
};

template <class T, bool LDS = false>
__device__ void tail_smooth(const TailLevel<T>& lvl, int sweeps, bool first, bool reverse = false)
{
    const TailView<T, LDS> lv(lvl);
    const auto& L = lv.geo;
    const int half = L.nx >> 1;
    const int rows = L.ny * L.nz;
    const int work = half * rows;
    if (first && sweeps == 0) {
        const int cells = L.nx * rows;
        for (int t = threadIdx.x; t < cells; t += blockDim.x) lv.p[t] = (T)0;
        __syncthreads();
        return;
    }
    for (int sw = 0; sw < sweeps; ++sw) {
        const bool f0 = first && sw == 0;
        for (int half_sweep = 0; half_sweep < 2; ++half_sweep) {
            const bool red = (half_sweep == 0) != reverse;
            for (int t = threadIdx.x; t < work; t += blockDim.x) {
                const int row = mdiv(t, half, lv.dm_hx), h = t - row * half;
                const int k = mdiv(row, L.ny, lv.dm_ny), j = row - k * L.ny;
                if (red) {
                    if (f0)
                        rbgs_cell<T, T, RED, true>(L, lv.beta, lv.q, lv.p, h, j, k);
                    else
                        rbgs_cell<T, T, RED, false>(L, lv.beta, lv.q, lv.p, h, j, k);
                } else {
                    if (f0)
                        rbgs_cell<T, T, BLACK, true>(L, lv.beta, lv.q, lv.p, h, j, k);
                    else
                        rbgs_cell<T, T, BLACK, false>(L, lv.beta, lv.q, lv.p, h, j, k);
                }
            }
            __syncthreads();
        }
    }
}

template <class T, bool LDS = false>
__device__ void tail_resid_restrict(const TailLevel<T>& lvl, const TailLevel<T>& lcl)
{
    const TailView<T, LDS> lv(lvl), lc(lcl);
    const auto& L = lv.geo;
    const int ncx = L.nx >> 1, ncy = L.ny >> 1, ncz = L.nz >> 1;
    const int cells = ncx * ncy * ncz;
    // two neighbouring lanes per coarse cell, one per fine plane (the chains
    // of dependent loads are half as long); the cell's value is the sum of the
    // two plane sums, the same on both lanes
    for (int t2 = threadIdx.x; t2 < 2 * cells; t2 += blockDim.x) {
        const int t = t2 >> 1, dk = t2 & 1;
        const int r = mdiv(t, ncx, lv.dm_hx), I = t - r * ncx;
        const int K = mdiv(r, ncy, lv.dm_hy), J = r - K * ncy;
        const T part = resid_restrict_plane<T, T>(L, lv.beta, lv.q, lv.p, I, J, K, dk, (T)0);
        const T other = __shfl_xor(part, 1, 64);
        if (dk == 0) lc.q[t] = part + other;
    }
    __syncthreads();
}

template <class T, bool LDS = false>
__device__ void tail_prolong(const TailLevel<T>& lvl, const TailLevel<T>& lcl)
{
    const TailView<T, LDS> lv(lvl), lc(lcl);
    const auto& L = lv.geo;
    const int cells = L.nx * L.ny * L.nz;
    for (int t = threadIdx.x; t < cells; t += blockDim.x) {
        const int r = mdiv(t, L.nx, lv.dm_nx), i = t - r * L.nx;
        const int k = mdiv(r, L.ny, lv.dm_ny), j = r - k * L.ny;
        prolong_cell<T>(L, lc.geo.nx, lc.geo.ny, lc.geo.nz, lc.p, lv.p, i, j, k);
    }
    __syncthreads();
}

// x = G q on the bottom level: S threads per row (S a power of two, at most
// 64, as many as the workgroup allows) each sum every S-th product of the row,
// then a butterfly over the S lanes; a row's loads are coalesced and in flight
// together instead of one dependent load per column.
template <class T, bool LDS = false>
__device__ void tail_dense(const TailPlan<T>& P, const TailLevel<T>& lvl)
{
    const TailView<T, LDS> lv(lvl);
    const int nb = lv.geo.nx * lv.geo.ny * lv.geo.nz;
    int S = 1;
    while (S < 64 && nb * (S * 2) <= (int)blockDim.x) S *= 2;
    const int rows = (int)blockDim.x / S;
    const int l = (int)threadIdx.x & (S - 1);
    for (int r0 = 0; r0 < nb; r0 += rows) {
        const int r = r0 + (int)threadIdx.x / S;
        double acc = 0.0;
        if (r < nb) {
            const glb_ptr<const double> g = (glb_ptr<const double>)(P.G + (size_t)r * nb);
            for (int c = l; c < nb; c += S) acc += g[c] * (double)lv.q[c];
        }
        for (int off = S >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (r < nb && l == 0) lv.p[r] = (T)acc;
    }
    __syncthreads();
}

// workgroup-wide deterministic sum of NV values (one workgroup, <= 1024 threads)
template <int NV>
__device__ void wg_sum(double (&v)[NV])
{
    __shared__ double part[16 * 4];
    __shared__ double res[4];
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < NV; ++q) part[wave * 4 + q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += part[w * 4 + q];
            res[q] = s;
        }
    }
    __syncthreads();
    for (int q = 0; q < NV; ++q) v[q] = res[q];
    __syncthreads();
}

// z = M_c r on the reference's coarsest level U: one V-cycle from zero over the
// sub-hierarchy U..bottom (red/black pre-smoothing, black/red post-smoothing,
// dense exact solve at the bottom).  Input in lv[U].q, output in lv[U].p.
template <class T>
__device__ void tail_inner_vcycle(const TailPlan<T>& P)
{
    const int top = P.user_last;
    for (int l = top; l < P.bottom; ++l) {
        tail_smooth<T>(P.lv[l], P.n_pre, true);
        tail_resid_restrict<T>(P.lv[l], P.lv[l + 1]);
    }
    tail_dense<T>(P, P.lv[P.bottom]);
    for (int l = P.bottom - 1; l >= top; --l) {
        tail_prolong<T>(P.lv[l], P.lv[l + 1]);
        tail_smooth<T>(P.lv[l], P.n_post, false, true);
    }
}

// Coarsest-level solve replacing the reference's MUMPS factor+solve
// (poissMat.f90:289-312): the Lagrange-constrained system (poissMat.f90:246-259)
// is solved by flexible conjugate gradients in the V-weighted inner product
// (V A is symmetric), preconditioned by the inner V-cycle, to the relative
// residual coarse_tol.  The right-hand side is first projected onto the range
// of A as the Lagrange multiplier does: q' = q - (sum V q)/(N V).
template <class T>
__device__ void tail_coarse_pcg(const TailPlan<T>& P)
{
    const TailLevel<T>& U = P.lv[P.user_last];
    const LevelDev<T>& L = U.geo;
    const int n = L.nx * L.ny * L.nz;
    T *x = P.cx, *r = P.cr, *d = P.cd, *ad = P.cad;
    const T* vol = P.cvol;
    double s1[1] = {0.0};
    for (int t = threadIdx.x; t < n; t += blockDim.x) s1[0] += (double)vol[t] * (double)U.q[t];
    wg_sum<1>(s1);
    const double shift = s1[0] / (double)n;
    double s2[1] = {0.0};
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const T rv = U.q[t] - (T)(shift / (double)vol[t]);
        r[t] = rv;
        x[t] = (T)0;
        U.q[t] = rv;
        s2[0] += (double)vol[t] * (double)rv * (double)rv;
    }
    wg_sum<1>(s2);
    const double rr0 = s2[0];
    if (!(rr0 > 0.0)) {
        for (int t = threadIdx.x; t < n; t += blockDim.x) U.p[t] = (T)0;
        __syncthreads();
        return;
    }
    tail_inner_vcycle<T>(P);
    double s3[1] = {0.0};
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        d[t] = U.p[t];
        s3[0] += (double)vol[t] * (double)r[t] * (double)U.p[t];
    }
    wg_sum<1>(s3);
    double delta = s3[0];
    const double tol2 = (double)P.ctol * (double)P.ctol;
    int it = 0;
    double rlast = rr0;
    for (it = 0; it < P.cmaxit; ++it) {
        double dq[1] = {0.0};
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const int i = t % L.nx, rr = t / L.nx, j = rr % L.ny, k = rr / L.ny;
            size_t c, e, w, nn, s, f, b;
            neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, nn, s, f, b);
            Coef<T> a = coefs<T>(L, U.beta, i, j, k, c, e, w, nn, s, f, b);
            const T v = apply_cell<T>(a, d, c, e, w, nn, s, f, b);
            ad[t] = v;
            dq[0] += (double)vol[t] * (double)d[t] * (double)v;
        }
        wg_sum<1>(dq);
        if (!(dq[0] != 0.0)) break;
        const double alpha = delta / dq[0];
        double rn[1] = {0.0};
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            x[t] = x[t] + (T)alpha * d[t];
            const T rv = r[t] - (T)alpha * ad[t];
            r[t] = rv;
            U.q[t] = rv;
            rn[0] += (double)vol[t] * (double)rv * (double)rv;
        }
        wg_sum<1>(rn);
        rlast = rn[0];
        if (rn[0] <= tol2 * rr0) {
            ++it;
            break;
        }
        tail_inner_vcycle<T>(P);
        double dz[2] = {0.0, 0.0};
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
            const double z = (double)U.p[t], vv = (double)vol[t];
            dz[0] += vv * (double)r[t] * z;
            dz[1] += vv * (double)ad[t] * z;
        }
        wg_sum<2>(dz);
        const double gmm = -alpha * dz[1] / delta;  // Polak-Ribiere
        delta = dz[0];
        for (int t = threadIdx.x; t < n; t += blockDim.x) d[t] = U.p[t] + (T)gmm * d[t];
        __syncthreads();
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) U.p[t] = x[t];
    if (threadIdx.x == 0 && P.st) {
        P.st->coarse_calls += 1;
        P.st->coarse_iters += it;
        if (it > P.st->coarse_iters_max) P.st->coarse_iters_max = it;
        const double rel = sqrt(rlast / rr0);
        if (rel > P.st->coarse_res_max) P.st->coarse_res_max = rel;
    }
    __syncthreads();
}

// ---- solveMG's stop test (multiGrid.f90:236-279): after V-cycle `cyc` of a
// preconditioner application, ||q - A p|| / (||q|| + tiny) on the finest level
// (computeResiduals, rbgs.f90:276-333, in the multigrid's precision, summed in
// FP64); the application stops at tolMG or after max_cycles cycles.
template <class T, class TQ, class TB>
__global__ void k_mg_res(LevelDev<T> L, const TB* __restrict__ beta, const TQ* __restrict__ q,
                         const T* __restrict__ p, const PcgState* st, Reducer red)
{
    double v[2] = {0.0, 0.0};
    if (!st->done) {
        const int n = L.nx * L.ny * L.nz;
        for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
            const int i = t % L.nx, rr = t / L.nx, j = rr % L.ny, k = rr / L.ny;
            size_t c, e, w, nn, s, f, b;
            neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, nn, s, f, b);
            Coef<T> a = coefs<T>(L, beta, i, j, k, c, e, w, nn, s, f, b);
            const T qc = (T)q[c];
            const T r = qc - apply_cell<T>(a, p, c, e, w, nn, s, f, b);
            v[0] += (double)r * (double)r;
            v[1] += (double)qc * (double)qc;
        }
    }
    block_partial<2>(v, red);
}
__global__ void k_fin_mgres(unsigned nb, double tol, int cyc, int max_cycles, PcgState* st, Reducer red)
{
    double tot[2];
    final_sum<2>(red, nb, tot);
    if (threadIdx.x == 0) {
        if (st->done) {
            st->mg_stop = 1;
            return;
        }
        const double res = sqrt(tot[0]) / (sqrt(tot[1]) + 2.2250738585072014e-308);
        const int cycles = cyc + 1;
        st->mg_cyc_total += 1;
        if (cycles > st->mg_cyc_max) st->mg_cyc_max = cycles;
        st->mg_stop = (cycles >= max_cycles || !(res > tol)) ? 1 : 0;
    }
}

template <class T, class T0>
int launch_mg_res(const LevelDev<T>& L, const T0* beta, const T0* q, const T* p, double tol, int cyc, int max_cycles,
                  PcgState* st, const Reducer& red, void* stream)
{
    const size_t n = (size_t)L.nx * L.ny * L.nz;
    // grid-stride: at most 2048 blocks (the reduction scratch holds 5 sums for that many)
    const unsigned nb = std::min<unsigned>(ceil_div(n, 256), 2048u);
    hipLaunchKernelGGL((k_mg_res<T, T0, T0>), dim3(nb), dim3(256), 0, (hipStream_t)stream, L, beta, q, p, st, red);
    hipLaunchKernelGGL(k_fin_mgres, dim3(1), dim3(256), 0, (hipStream_t)stream, nb, tol, cyc, max_cycles, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}
template int launch_mg_res<float, float>(const LevelDev<float>&, const float*, const float*, const float*, double, int, int,
                                  PcgState*, const Reducer&, void*);
template int launch_mg_res<float, bf16s>(const LevelDev<float>&, const bf16s*, const bf16s*, const float*, double, int,
                                         int, PcgState*, const Reducer&, void*);
template int launch_mg_res<double, double>(const LevelDev<double>&, const double*, const double*, const double*, double, int,
                                   int, PcgState*, const Reducer&, void*);

// ---- the coarsest-level solve by multi-workgroup kernels (tail_coarse_pcg's
// algorithm, for a coarsest level too large for one workgroup): one thread per
// cell, block partial sums, a one-block finalize per reduction.  Every kernel
// but the last returns at once when the inner solve has stopped.

// sum V q; the finalize also stops the inner solve when the outer one has
template <class T>
__global__ void k_in_sum_vq(int n, const T* __restrict__ vol, const T* __restrict__ q, const PcgState* st,
                            Reducer red)
{
    double v[1] = {0.0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (!st->done && t < n) v[0] = (double)vol[t] * (double)q[t];
    block_partial<1>(v, red);
}
__global__ void k_in_fin_shift(unsigned nb, int n, InnerState* is, const PcgState* st, Reducer red)
{
    double tot[1];
    final_sum<1>(red, nb, tot);
    if (threadIdx.x == 0) {
        is->shift = tot[0] / (double)n;
        is->done = st->done ? 1 : 0;
        is->it = 0;
        is->rlast = 0.0;
    }
}
// r = q - shift / V (the right-hand side projected as the Lagrange row does),
// x = 0, q = r (the right-hand side of the inner V-cycle); sum V r^2
template <class T>
__global__ void k_in_init(int n, const T* __restrict__ vol, T* __restrict__ q, T* __restrict__ r, T* __restrict__ x,
                          const InnerState* is, Reducer red)
{
    double v[1] = {0.0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (!is->done && t < n) {
        const T rv = q[t] - (T)(is->shift / (double)vol[t]);
        r[t] = rv;
        x[t] = (T)0;
        q[t] = rv;
        v[0] = (double)vol[t] * (double)rv * (double)rv;
    }
    block_partial<1>(v, red);
}
__global__ void k_in_fin_rr0(unsigned nb, InnerState* is, Reducer red)
{
    if (is->done) return;
    double tot[1];
    final_sum<1>(red, nb, tot);
    if (threadIdx.x == 0) {
        is->rr0 = tot[0];
        is->rlast = tot[0];
        if (!(tot[0] > 0.0)) is->done = 1;  // zero right-hand side: x = 0
    }
}
// d = z, sum V r z (the first direction)
template <class T>
__global__ void k_in_first_dir(int n, const T* __restrict__ vol, const T* __restrict__ r, const T* __restrict__ z,
                               T* __restrict__ d, const InnerState* is, Reducer red)
{
    if (is->done) return;
    double v[1] = {0.0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        d[t] = z[t];
        v[0] = (double)vol[t] * (double)r[t] * (double)z[t];
    }
    block_partial<1>(v, red);
}
__global__ void k_in_fin_delta(unsigned nb, InnerState* is, Reducer red)
{
    if (is->done) return;
    double tot[1];
    final_sum<1>(red, nb, tot);
    if (threadIdx.x == 0) is->delta = tot[0];
}
// ad = A d, sum V d ad
template <class T>
__global__ void k_in_op(LevelDev<T> L, const T* __restrict__ beta, const T* __restrict__ vol, const T* __restrict__ d,
                        T* __restrict__ ad, const InnerState* is, Reducer red)
{
    if (is->done) return;
    double v[1] = {0.0};
    const int n = L.nx * L.ny * L.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const int i = t % L.nx, rr = t / L.nx, j = rr % L.ny, k = rr / L.ny;
        size_t c, e, w, nn, s, f, b;
        neighbours(i, j, k, L.nx, L.ny, L.nz, c, e, w, nn, s, f, b);
        Coef<T> a = coefs<T>(L, beta, i, j, k, c, e, w, nn, s, f, b);
        const T av = apply_cell<T>(a, d, c, e, w, nn, s, f, b);
        ad[t] = av;
        v[0] = (double)vol[t] * (double)d[t] * (double)av;
    }
    block_partial<1>(v, red);
}
__global__ void k_in_fin_alpha(unsigned nb, InnerState* is, Reducer red)
{
    if (is->done) return;
    double tot[1];
    final_sum<1>(red, nb, tot);
    if (threadIdx.x == 0) {
        if (!(tot[0] != 0.0))
            is->done = 1;
        else
            is->alpha = is->delta / tot[0];
    }
}
// x += alpha d, r -= alpha ad, q = r; sum V r^2
template <class T>
__global__ void k_in_update(int n, const T* __restrict__ vol, T* __restrict__ x, T* __restrict__ r, T* __restrict__ q,
                            const T* __restrict__ d, const T* __restrict__ ad, const InnerState* is, Reducer red)
{
    if (is->done) return;
    double v[1] = {0.0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const T alpha = (T)is->alpha;
        x[t] = x[t] + alpha * d[t];
        const T rv = r[t] - alpha * ad[t];
        r[t] = rv;
        q[t] = rv;
        v[0] = (double)vol[t] * (double)rv * (double)rv;
    }
    block_partial<1>(v, red);
}
__global__ void k_in_fin_update(unsigned nb, double tol2, InnerState* is, Reducer red)
{
    if (is->done) return;
    double tot[1];
    final_sum<1>(red, nb, tot);
    if (threadIdx.x == 0) {
        is->rlast = tot[0];
        is->it += 1;
        if (tot[0] <= tol2 * is->rr0) is->done = 1;
    }
}
// sum V r z, sum V ad z of the new preconditioned residual z
template <class T>
__global__ void k_in_dots(int n, const T* __restrict__ vol, const T* __restrict__ r, const T* __restrict__ ad,
                          const T* __restrict__ z, const InnerState* is, Reducer red)
{
    if (is->done) return;
    double v[2] = {0.0, 0.0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        const double zz = (double)z[t], vv = (double)vol[t];
        v[0] = vv * (double)r[t] * zz;
        v[1] = vv * (double)ad[t] * zz;
    }
    block_partial<2>(v, red);
}
__global__ void k_in_fin_gamma(unsigned nb, InnerState* is, Reducer red)
{
    if (is->done) return;
    double tot[2];
    final_sum<2>(red, nb, tot);
    if (threadIdx.x == 0) {
        is->gamma = -is->alpha * tot[1] / is->delta;  // Polak-Ribiere
        is->delta = tot[0];
    }
}
template <class T>
__global__ void k_in_dir(int n, const T* __restrict__ z, T* __restrict__ d, const InnerState* is)
{
    if (is->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = z[t] + (T)is->gamma * d[t];
}
// the solution into the level's p, and the diagnostics of tbfg_pcg_coarse_stats
template <class T>
__global__ void k_in_final(int n, const T* __restrict__ x, T* __restrict__ p, const InnerState* is, PcgState* st)
{
    if (st->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = x[t];
    if (blockIdx.x == 0 && threadIdx.x == 0 && is->rr0 > 0.0) {  // a zero right-hand side counts no call
        st->coarse_calls += 1;
        st->coarse_iters += is->it;
        if (is->it > st->coarse_iters_max) st->coarse_iters_max = is->it;
        const double rel = sqrt(is->rlast / is->rr0);
        if (rel > st->coarse_res_max) st->coarse_res_max = rel;
    }
}

template <class T>
int launch_inner(InnerOp op, const LevelDev<T>& L, const InnerArgs<T>& a, InnerState* is, PcgState* st,
                 const Reducer& red, void* stream)
{
    const int n = L.nx * L.ny * L.nz;
    const unsigned nb = ceil_div((size_t)n, 256);
    hipStream_t s = (hipStream_t)stream;
    switch (op) {
    case InnerOp::Start:
        hipLaunchKernelGGL(k_in_sum_vq<T>, dim3(nb), dim3(256), 0, s, n, a.vol, a.q, st, red);
        hipLaunchKernelGGL(k_in_fin_shift, dim3(1), dim3(256), 0, s, nb, n, is, st, red);
        hipLaunchKernelGGL(k_in_init<T>, dim3(nb), dim3(256), 0, s, n, a.vol, a.q, a.r, a.x, is, red);
        hipLaunchKernelGGL(k_in_fin_rr0, dim3(1), dim3(256), 0, s, nb, is, red);
        break;
    case InnerOp::FirstDir:
        hipLaunchKernelGGL(k_in_first_dir<T>, dim3(nb), dim3(256), 0, s, n, a.vol, a.r, a.z, a.d, is, red);
        hipLaunchKernelGGL(k_in_fin_delta, dim3(1), dim3(256), 0, s, nb, is, red);
        break;
    case InnerOp::OpUpdate:
        hipLaunchKernelGGL(k_in_op<T>, dim3(nb), dim3(256), 0, s, L, a.beta, a.vol, a.d, a.ad, is, red);
        hipLaunchKernelGGL(k_in_fin_alpha, dim3(1), dim3(256), 0, s, nb, is, red);
        hipLaunchKernelGGL(k_in_update<T>, dim3(nb), dim3(256), 0, s, n, a.vol, a.x, a.r, a.q, a.d, a.ad, is, red);
        hipLaunchKernelGGL(k_in_fin_update, dim3(1), dim3(256), 0, s, nb, a.tol2, is, red);
        break;
    case InnerOp::NextDir:
        hipLaunchKernelGGL(k_in_dots<T>, dim3(nb), dim3(256), 0, s, n, a.vol, a.r, a.ad, a.z, is, red);
        hipLaunchKernelGGL(k_in_fin_gamma, dim3(1), dim3(256), 0, s, nb, is, red);
        hipLaunchKernelGGL(k_in_dir<T>, dim3(nb), dim3(256), 0, s, n, a.z, a.d, is);
        break;
    case InnerOp::Finish:
        hipLaunchKernelGGL(k_in_final<T>, dim3(nb), dim3(256), 0, s, n, a.x, a.p, is, st);
        break;
    }
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}
template int launch_inner<float>(InnerOp, const LevelDev<float>&, const InnerArgs<float>&, InnerState*, PcgState*,
                                 const Reducer&, void*);
template int launch_inner<double>(InnerOp, const LevelDev<double>&, const InnerArgs<double>&, InnerState*,
                                  PcgState*, const Reducer&, void*);

// V-cycle over tail levels top..bot (multiGrid.f90:186-232) with the coarsest
// level solved by tail_coarse_pcg (or the dense solve when it is the bottom;
// COARSE_CG false: always, and the inner CG is not compiled in).
template <class T, bool COARSE_CG, bool LDS = false>
__device__ void tail_vcycle(const TailPlan<T>& P, int top, int bot, bool reset_top)
{
    static_assert(!(COARSE_CG && LDS), "the inner coarsest-level CG keeps its vectors in HBM");
    for (int l = top; l < bot; ++l) {
        tail_smooth<T, LDS>(P.lv[l], P.n_pre, (l > top) || reset_top);
        tail_resid_restrict<T, LDS>(P.lv[l], P.lv[l + 1]);
    }
    if constexpr (COARSE_CG) {
        if (P.bottom > P.user_last)
            tail_coarse_pcg<T>(P);
        else
            tail_dense<T>(P, P.lv[bot]);
    } else {
        tail_dense<T, LDS>(P, P.lv[bot]);
    }
    for (int l = bot - 1; l >= top; --l) {
        tail_prolong<T, LDS>(P.lv[l], P.lv[l + 1]);
        tail_smooth<T, LDS>(P.lv[l], P.n_post, false, P.rev_post != 0);
    }
}

template <class T>
__global__ void __launch_bounds__(1024) k_tail(const TailPlan<T>* __restrict__ plan, const int* done)
{
    if (done && *done) return;
    const TailPlan<T>& P = *plan;
    tail_vcycle<T, true>(P, 0, P.user_last, true);
}

// The same V-cycle with p, q, beta and the metric tables of every tail level
// held in LDS (dynamic shared memory sized by the host): every phase then
// costs LDS latency instead of an L2 round trip.  beta, the tables and q of the
// first level come in at the start, p of the first level goes back to HBM at
// the end (the next level up prolongs from it).
template <class T>
__global__ void __launch_bounds__(1024) k_tail_lds(const TailPlan<T>* __restrict__ plan, const int* done)
{
    extern __shared__ __align__(16) unsigned char tail_smem[];
    __shared__ TailPlan<T> Ps;
    if (done && *done) return;
    {
        static_assert(sizeof(TailPlan<T>) % 4 == 0, "plan copied as 32-bit words");
        const unsigned* src = reinterpret_cast<const unsigned*>(plan);
        unsigned* dst = reinterpret_cast<unsigned*>(&Ps);
        for (unsigned t = threadIdx.x; t < sizeof(TailPlan<T>) / 4; t += blockDim.x) dst[t] = src[t];
    }
    __syncthreads();
    T* const base = reinterpret_cast<T*>(tail_smem);
    // level m's pointers into LDS (thread m), the table pointers moved by the
    // offset of the block
    if ((int)threadIdx.x < Ps.nlev) {
        TailLevel<T>& lv = Ps.lv[threadIdx.x];
        const T* g0 = lv.geo.a[0].ip;
        T* tb = base + lv.lds_tab;
        for (int a = 0; a < 3; ++a) {
            AxisDev<T>& ax = lv.geo.a[a];
            ax.ip = tb + (ax.ip - g0);
            ax.im = tb + (ax.im - g0);
            ax.rlo = tb + (ax.rlo - g0);
            ax.rhi = tb + (ax.rhi - g0);
            ax.plo = tb + (ax.plo - g0);
            ax.phi = tb + (ax.phi - g0);
        }
        lv.p = base + lv.lds_p;
        lv.q = base + lv.lds_q;
        lv.beta = base + lv.lds_beta;
    }
    {
        // every thread issues the loads of up to kStage elements of the list
        // before it stores any, so the HBM/L2 latency is paid once per round
        // instead of once per array
        constexpr int kStage = sizeof(T) == 4 ? 12 : 6;
        const lds_ptr<T> sm = (lds_ptr<T>)base;
        const int total = Ps.seg_first[Ps.nseg];
        for (int e0 = 0; e0 < total; e0 += kStage * (int)blockDim.x) {
            T v[kStage];
            int dst[kStage];
            int sg = 0;
#pragma unroll
            for (int u = 0; u < kStage; ++u) {
                const int e = e0 + u * (int)blockDim.x + (int)threadIdx.x;
                dst[u] = -1;
                if (e < total) {
                    while (e >= Ps.seg_first[sg + 1]) ++sg;  // e grows with u
                    const int o = e - Ps.seg_first[sg];
                    v[u] = ((glb_ptr<const T>)Ps.seg_src[sg])[o];
                    dst[u] = Ps.seg_dst[sg] + o;
                }
            }
#pragma unroll
            for (int u = 0; u < kStage; ++u)
                if (dst[u] >= 0) sm[dst[u]] = v[u];
        }
    }
    __syncthreads();
    tail_vcycle<T, false, true>(Ps, 0, Ps.user_last, true);
    const int n0 = Ps.lv[0].geo.nx * Ps.lv[0].geo.ny * Ps.lv[0].geo.nz;
    const glb_ptr<T> gp = (glb_ptr<T>)plan->lv[0].p;
    const lds_ptr<T> p0 = (lds_ptr<T>)Ps.lv[0].p;
    for (int t = threadIdx.x; t < n0; t += blockDim.x) gp[t] = p0[t];
}

// Dense bottom operator: G = (M + c 1 1^T)^-1 (I - 1 1^T / n) (-V), with
// M = -V A the V-scaled symmetric matrix of poissMat.f90:746-763.  x = G q is the
// Lagrange-constrained exact solve (poissMat.f90:246-259,467-503, sum x = 0).
// Gauss-Jordan on the symmetric positive definite M + c 1 1^T, one workgroup.
// A, B: the n x n work matrices, in HBM (W) or in LDS (bottoms of at most
// kDenseLdsCells cells); int indices (n <= 512).
template <class T, class PA>
__device__ void dense_setup_body(const LevelDev<double>& L, const T* __restrict__ beta_t,
                                 const double* __restrict__ vol, double* __restrict__ bet, PA A, PA B,
                                 double* __restrict__ G)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int nb = nx * ny * nz;
    const int nn = nb * nb;
    __shared__ double s_trace;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) bet[t] = (double)beta_t[t];
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        A[t] = 0.0;
        B[t] = ((t / nb) == (t % nb)) ? 1.0 : 0.0;
    }
    __syncthreads();
    // rows of M: for each face with neighbour nb != c: M[c][nb] -= V a_f, M[c][c] += V a_f
    for (int c = threadIdx.x; c < nb; c += blockDim.x) {
        const int i = c % nx, r = c / nx, j = r % ny, k = r / ny;
        int cc, e, w, n, s, f, b;
        neighbours(i, j, k, nx, ny, nz, cc, e, w, n, s, f, b);
        Coef<double> a = coefs<double>(L, bet, i, j, k, cc, e, w, n, s, f, b);
        const double v = vol[c];
        const int nbr[6] = {e, w, n, s, f, b};
        const double af[6] = {a.aR, a.aL, a.aT, a.aB, a.aF, a.aK};
        for (int q = 0; q < 6; ++q) {
            if (nbr[q] == cc) continue;
            A[c * nb + nbr[q]] -= v * af[q];
            A[c * nb + c] += v * af[q];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0;
        for (int c = 0; c < nb; ++c) tr += A[c * nb + c];
        s_trace = (tr > 0.0 ? tr : 1.0) / ((double)nb * nb);
    }
    __syncthreads();
    const double cshift = s_trace;
    for (int t = threadIdx.x; t < nn; t += blockDim.x) A[t] += cshift;
    __syncthreads();
    // Gauss-Jordan without pivoting (SPD)
    for (int piv = 0; piv < nb; ++piv) {
        const double inv = 1.0 / A[piv * nb + piv];
        __syncthreads();
        for (int c = threadIdx.x; c < nb; c += blockDim.x) {
            A[piv * nb + c] *= inv;
            B[piv * nb + c] *= inv;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const int r = t / nb, c = t - r * nb;
            if (r == piv) continue;
            const double fct = A[r * nb + piv];
            if (c != piv) A[t] -= fct * A[piv * nb + c];
            B[t] -= fct * B[piv * nb + c];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < nb; r += blockDim.x)
            if (r != piv) A[r * nb + piv] = 0.0;
        __syncthreads();
    }
    // G[i][j] = -V_j (B[i][j] - rowmean_i(B))
    for (int r = threadIdx.x; r < nb; r += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < nb; ++c) s += B[r * nb + c];
        s /= (double)nb;
        for (int c = 0; c < nb; ++c) G[(size_t)r * nb + c] = -vol[c] * (B[r * nb + c] - s);
    }
}

template <class T>
__global__ void __launch_bounds__(1024) k_dense_setup(LevelDev<double> L, const T* __restrict__ beta_t,
                                                     const double* __restrict__ vol, double* __restrict__ W,
                                                     double* __restrict__ G)
{
    const int nb = L.nx * L.ny * L.nz;
    double* A = W;                        // n x n: M + c 1 1^T, reduced to the identity
    double* B = W + (size_t)nb * nb;      // n x n: identity, becomes the inverse
    dense_setup_body<T>(L, beta_t, vol, W + 2 * (size_t)nb * nb, A, B, G);
}

// the same with A and B in LDS (dynamic, 16 n^2 bytes)
constexpr int kDenseLdsCells = 63;
template <class T>
__global__ void __launch_bounds__(1024) k_dense_setup_lds(LevelDev<double> L, const T* __restrict__ beta_t,
                                                         const double* __restrict__ vol, double* __restrict__ W,
                                                         double* __restrict__ G)
{
    extern __shared__ double dense_smem[];
    const int nb = L.nx * L.ny * L.nz;
    const lds_ptr<double> A = (lds_ptr<double>)dense_smem;
    dense_setup_body<T>(L, beta_t, vol, W + 2 * (size_t)nb * nb, A, A + nb * nb, G);
}

// ------------------------------------------------------------ host launches

namespace {
dim3 block2d(int nx_threads)
{
    int bx = next_pow2(nx_threads < 1 ? 1 : nx_threads);
    if (bx > 64) bx = 64;
    return dim3(bx, 256 / bx, 1);
}
inline hipStream_t S(void* s) { return (hipStream_t)s; }
}  // namespace

int launch_apply_op(const LevelDev<double>& L, const double* beta, const double* p, double* q, void* stream)
{
    dim3 blk = block2d(L.nx);
    dim3 grd(ceil_div(L.nx, blk.x), ceil_div(L.ny, blk.y), L.nz);
    hipLaunchKernelGGL(k_apply_op, grd, blk, 0, S(stream), L, beta, p, q);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

// reduction kernels: a few thousand blocks, each reducing once (per-block
// reduction latency times tens of thousands of blocks would dominate)
constexpr unsigned kRedBlocks = 2048;

static dim3 red_grid3(const LevelDev<double>& L, dim3 blk)
{
    const unsigned gx = ceil_div(L.nx, blk.x), gy = ceil_div(L.ny, blk.y);
    unsigned gz = kRedBlocks / (gx * gy);
    if (gz < 1) gz = 1;
    if (gz > (unsigned)L.nz) gz = (unsigned)L.nz;
    return dim3(gx, gy, gz);
}

static unsigned red_blocks1(size_t n)
{
    const unsigned b = ceil_div(n, 256);
    return b < kRedBlocks ? b : kRedBlocks;
}

// operator blocks: whole rows (up to 256 cells, so the x neighbours of a row
// come from the same block), two rows, kOpPlanes planes per thread
constexpr int kOpPlanes = 16;

static dim3 op_block(const LevelDev<double>& L)
{
    int bx = next_pow2(L.nx < 1 ? 1 : L.nx);
    if (bx > 256) bx = 256;
    return dim3(bx, 2, 1);
}

static dim3 op_grid(const LevelDev<double>& L, dim3 blk)
{
    return dim3(ceil_div(L.nx, blk.x), ceil_div(L.ny, blk.y), ceil_div(L.nz, kOpPlanes));
}

OpPlan op_plan(const LevelDev<double>& L)
{
    OpPlan m;
    const char* env = std::getenv("TBFG_OP_TILE");
    // 256-thread tiles first: with long marches they measured fastest (256x192x256:
    // 64x256 tiles 136 us, 32x256 138 us, 128x512 152 us, all with 64-plane chunks)
    std::vector<std::pair<int, int>> cand = {{64, 256}, {32, 256}, {128, 512}, {64, 512}};
    int tuned[3];
    if (env) {
        int a = 0, b = 0;
        if (std::sscanf(env, "%dx%d", &a, &b) != 2) return m;
        cand = {{a, b}};
    } else if (tune_lookup(TBFG_TUNE_OP, L.nx, L.ny, L.nz, 8, tuned)) {
        if (tuned[0] < 0) return m;  // columns
        const std::pair<int, int> t{tuned[0], tuned[1]};
        const bool known = t == std::pair<int, int>{128, 512} || t == std::pair<int, int>{64, 512} ||
                           t == std::pair<int, int>{32, 256} || t == std::pair<int, int>{64, 256};
        if (known) cand.insert(cand.begin(), t);
    }
    for (auto [tx, nt] : cand) {
        if (L.nx % tx == 0 && L.ny % (nt / tx) == 0 && L.nz >= 2) {
            m.tx = tx;
            m.nt = nt;
            break;
        }
    }
    if (!m.tx) return m;
    m.tiles_x = L.nx / m.tx;
    m.ntiles = m.tiles_x * (L.ny / (m.nt / m.tx));
    // about 3 x 256 threads per CU, chunks a multiple of 4 planes: fewer, longer
    // marches re-read fewer halo planes (256x192x256, 64x256 tiles: 64-plane
    // chunks 136 us, 32-plane chunks 149 us)
    const int chunks = std::max(1, (768 * 256 / m.nt + m.ntiles - 1) / m.ntiles);
    int kc = (L.nz + chunks - 1) / chunks;
    kc = std::min(kOpMaxChunk, std::max(4, (kc + 3) / 4 * 4));
    if (const char* e = std::getenv("TBFG_OP_KC")) kc = std::min(kOpMaxChunk, std::max(4, std::atoi(e) / 4 * 4));
    m.kc = kc;
    m.nblocks = (unsigned)m.ntiles * (unsigned)((L.nz + kc - 1) / kc);
    return m;
}

size_t reduction_slots(const LevelDev<double>& L, const OpPlan& plan)
{
    const dim3 g3 = red_grid3(L, block2d(L.nx));
    const dim3 go = op_grid(L, op_block(L));
    const size_t chunked = (size_t)g3.x * g3.y * g3.z * 5;  // dots: 5 sums per block
    const size_t flat = (size_t)kRedBlocks * 5;
    const size_t op = std::max((size_t)go.x * go.y * go.z, (size_t)plan.nblocks);
    // the sweep with the dot products on any tile shape: at most one workgroup
    // per 32x8 tile and 6-plane chunk
    const size_t sweep = (size_t)(L.nx / 32 + 1) * (L.ny / 8 + 1) * (L.nz / 6 + 1) * 5;
    return std::max(std::max(chunked, flat), std::max(op, sweep));
}

int launch_op_dot(const LevelDev<double>& L, const OpPlan& m, const double* beta, const double* d, double* qv,
                  PcgState* st, const Reducer& red, void* stream)
{
    if (m.tx) {
#define TBFG_OPM(TX_, NT_)                                                                                          \
    if (m.tx == TX_ && m.nt == NT_)                                                                                 \
        hipLaunchKernelGGL((k_op_march<TX_, NT_, 4, double, false>), dim3(m.nblocks), dim3(NT_), 0, S(stream), L,   \
                           beta, d, qv, st, red, m.kc, m.tiles_x, m.ntiles, OpDir<double>{});
        TBFG_OPM(128, 512) else TBFG_OPM(64, 512) else TBFG_OPM(32, 256) else TBFG_OPM(64, 256) else {
            set_error("no operator tile %dx%d", m.tx, m.nt);
            return TBFG_EINVAL;
        }
#undef TBFG_OPM
        hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(1024), 0, S(stream), m.nblocks, st, red);
        TBFG_HIP(hipGetLastError());
        return TBFG_OK;
    }
    dim3 blk = op_block(L);
    dim3 grd = op_grid(L, blk);
    const unsigned nb = grd.x * grd.y * grd.z;
    hipLaunchKernelGGL(k_op_dot<kOpPlanes>, dim3(nb), blk, 0, S(stream), L, beta, d, qv, st, red);
    hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(1024), 0, S(stream), nb, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TZ>
int launch_op_dir(const LevelDev<double>& L, const OpPlan& m, const double* beta, const TZ* z, const double* dold,
                  double* dnew, double* psi, int init, double* qv, PcgState* st, const Reducer& red, void* stream,
                  const KernelEvents* kev)
{
    const OpDir<TZ> od{z, dold, dnew, psi, init};
    // (kev: events recorded by the dispatch of the pass itself, for timing it alone)
#define TBFG_OPD(TX_, NT_)                                                                                         \
    if (m.tx == TX_ && m.nt == NT_) {                                                                              \
        if (kev)                                                                                                   \
            hipExtLaunchKernelGGL((k_op_march<TX_, NT_, 4, TZ, true>), dim3(m.nblocks), dim3(NT_), 0, S(stream),  \
                                  kev->start, kev->stop, 0, L, beta, nullptr, qv, st, red, m.kc, m.tiles_x,        \
                                  m.ntiles, od);                                                                   \
        else                                                                                                       \
            hipLaunchKernelGGL((k_op_march<TX_, NT_, 4, TZ, true>), dim3(m.nblocks), dim3(NT_), 0, S(stream), L,  \
                               beta, nullptr, qv, st, red, m.kc, m.tiles_x, m.ntiles, od);                        \
    }
    TBFG_OPD(128, 512) else TBFG_OPD(64, 512) else TBFG_OPD(32, 256) else TBFG_OPD(64, 256) else {
        set_error("launch_op_dir: no operator tile %dx%d", m.tx, m.nt);
        return TBFG_EINVAL;
    }
#undef TBFG_OPD
    hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(1024), 0, S(stream), m.nblocks, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}
template int launch_op_dir<float>(const LevelDev<double>&, const OpPlan&, const double*, const float*, const double*,
                                  double*, double*, int, double*, PcgState*, const Reducer&, void*,
                                  const KernelEvents*);
template int launch_op_dir<double>(const LevelDev<double>&, const OpPlan&, const double*, const double*,
                                   const double*, double*, double*, int, double*, PcgState*, const Reducer&, void*,
                                   const KernelEvents*);

int launch_state_init(PcgState* st, double tol, int32_t max_iter, int32_t flexible, void* stream)
{
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, S(stream), st, tol, max_iter, flexible);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int launch_psi_last(size_t n, double* psi, const double* d0, const double* d1, const PcgState* st, void* stream)
{
    hipLaunchKernelGGL(k_psi_last, dim3(red_blocks1(n)), dim3(256), 0, S(stream), n, psi, d0, d1, st);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TR>
int launch_residual_init(const LevelDev<double>& L, const OpPlan& m, const double* beta, const double* q,
                         const double* psi, double* r, TR* r32, PcgState* st, const Reducer& red, void* stream)
{
    if (m.tx && 2 * (size_t)m.nblocks <= reduction_slots(L, m)) {  // the operator's plane-march tiles
#define TBFG_RSM(TX_, NT_)                                                                                        \
    if (m.tx == TX_ && m.nt == NT_)                                                                               \
        hipLaunchKernelGGL((k_resid_march<TX_, NT_, TR>), dim3(m.nblocks), dim3(NT_), 0, S(stream), L, beta, q,   \
                           psi, r, r32, red, m.kc, m.tiles_x, m.ntiles);
        TBFG_RSM(128, 512) else TBFG_RSM(64, 512) else TBFG_RSM(32, 256) else TBFG_RSM(64, 256) else {
            set_error("launch_residual_init: no operator tile %dx%d", m.tx, m.nt);
            return TBFG_EINVAL;
        }
#undef TBFG_RSM
        hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(256), 0, S(stream), m.nblocks, st, red);
        TBFG_HIP(hipGetLastError());
        return TBFG_OK;
    }
    dim3 blk = block2d(L.nx);
    dim3 grd = red_grid3(L, blk);
    hipLaunchKernelGGL(k_residual_init<TR>, grd, blk, 0, S(stream), L, beta, q, psi, r, r32, st, red);
    hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(256), 0, S(stream), grd.x * grd.y * grd.z, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TR>
int launch_pcg_update(size_t n, double* psi, double* r, TR* r32, const double* d, const double* qv,
                      PcgState* st, const Reducer& red, void* stream)
{
    const unsigned nb = red_blocks1(n);
    hipLaunchKernelGGL(k_pcg_update<TR>, dim3(nb), dim3(256), 0, S(stream), n, psi, r, r32, d, qv, st, red);
    hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(256), 0, S(stream), nb, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TZ>
int launch_pcg_dots(size_t n, const double* r, const TZ* z, const double* qv, int init, PcgState* st,
                    const Reducer& red, void* stream)
{
    const unsigned nb = red_blocks1(n);
    hipLaunchKernelGGL(k_pcg_dots<TZ>, dim3(nb), dim3(256), 0, S(stream), n, r, z, qv, init, st, red);
    hipLaunchKernelGGL(k_fin_dots, dim3(1), dim3(256), 0, S(stream), nb, n, init, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int launch_fin_dots(unsigned nblocks, size_t ncell, int init, PcgState* st, const Reducer& red, void* stream)
{
    hipLaunchKernelGGL(k_fin_dots, dim3(1), dim3(1024), 0, S(stream), nblocks, ncell, init, st, red);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TZ>
int launch_pcg_dir(size_t n, const TZ* z, double* d, int init, const PcgState* st, void* stream)
{
    hipLaunchKernelGGL(k_pcg_dir<TZ>, dim3(ceil_div(n, 256)), dim3(256), 0, S(stream), n, z, d, init, st);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class T, class TQ>
int launch_rbgs(const LevelDev<T>& L, const T* beta, const TQ* q, T* p, int colour, bool first, const int* done,
                void* stream)
{
    dim3 blk = block2d(L.nx >> 1);
    const dim3 grd(ceil_div(L.nx >> 1, blk.x) * ceil_div(L.ny, blk.y) * L.nz);
    if (colour == RED) {
        if (first)
            hipLaunchKernelGGL((k_rbgs<T, TQ, RED, true>), grd, blk, 0, S(stream), L, beta, q, p, done);
        else
            hipLaunchKernelGGL((k_rbgs<T, TQ, RED, false>), grd, blk, 0, S(stream), L, beta, q, p, done);
    } else {
        if (first)
            hipLaunchKernelGGL((k_rbgs<T, TQ, BLACK, true>), grd, blk, 0, S(stream), L, beta, q, p, done);
        else
            hipLaunchKernelGGL((k_rbgs<T, TQ, BLACK, false>), grd, blk, 0, S(stream), L, beta, q, p, done);
    }
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class T, class TQ, class TB>
int launch_resid_restrict(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* p, T* qc, const int* done,
                          void* stream)
{
    dim3 blk = block2d(L.nx >> 1);
    const dim3 grd(ceil_div(L.nx >> 1, blk.x) * ceil_div(L.ny >> 1, blk.y) * ceil_div(L.nz >> 1, kRrPlanes));
    hipLaunchKernelGGL((k_resid_restrict<T, TQ, TB>), grd, blk, 0, S(stream), L, beta, q, p, qc, done);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TF, class T>
int launch_restrict(const LevelDev<T>& L, const TF* fine, T* coarse, void* stream)
{
    dim3 blk = block2d(L.nx >> 1);
    dim3 grd(ceil_div(L.nx >> 1, blk.x), ceil_div(L.ny >> 1, blk.y), L.nz >> 1);
    hipLaunchKernelGGL((k_restrict<TF, T>), grd, blk, 0, S(stream), L, fine, coarse);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class T>
int launch_prolong_add(const LevelDev<T>& L, const T* pc, T* p, const int* done, void* stream)
{
    dim3 blk = block2d(L.nx >> 1);
    const dim3 grd(ceil_div(L.nx >> 1, blk.x) * ceil_div(L.ny >> 1, blk.y) * (L.nz >> 1));
    hipLaunchKernelGGL(k_prolong_add<T>, grd, blk, 0, S(stream), L, pc, p, done);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class TS, class TD>
int launch_convert(size_t n, const TS* src, TD* dst, void* stream)
{
    hipLaunchKernelGGL((k_convert<TS, TD>), dim3(ceil_div(n, 256)), dim3(256), 0, S(stream), n, src, dst);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class T>
int launch_tail(const TailPlan<T>* plan_dev, const int* done, size_t lds_bytes, void* stream)
{
    if (lds_bytes > 0) {
        hipLaunchKernelGGL(k_tail_lds<T>, dim3(1), dim3(1024), lds_bytes, S(stream), plan_dev, done);
        TBFG_HIP(hipGetLastError());
        return TBFG_OK;
    }
    hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(1024), 0, S(stream), plan_dev, done);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

template <class T>
int launch_dense_setup(const LevelDev<double>& L, const T* beta, const double* vol, double* work, double* G,
                       void* stream)
{
    const int nb = L.nx * L.ny * L.nz;
    if (nb <= kDenseLdsCells)
        hipLaunchKernelGGL(k_dense_setup_lds<T>, dim3(1), dim3(1024), (size_t)nb * nb * 2 * sizeof(double), S(stream),
                           L, beta, vol, work, G);
    else
        hipLaunchKernelGGL(k_dense_setup<T>, dim3(1), dim3(1024), 0, S(stream), L, beta, vol, work, G);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

// explicit instantiations
template int launch_pcg_dots<double>(size_t, const double*, const double*, const double*, int, PcgState*,
                                     const Reducer&, void*);
template int launch_pcg_dots<float>(size_t, const double*, const float*, const double*, int, PcgState*,
                                    const Reducer&, void*);
template int launch_pcg_dir<double>(size_t, const double*, double*, int, const PcgState*, void*);
template int launch_pcg_dir<float>(size_t, const float*, double*, int, const PcgState*, void*);
template int launch_rbgs<double, double>(const LevelDev<double>&, const double*, const double*, double*, int, bool,
                                         const int*, void*);
template int launch_rbgs<float, double>(const LevelDev<float>&, const float*, const double*, float*, int, bool,
                                        const int*, void*);
template int launch_rbgs<float, float>(const LevelDev<float>&, const float*, const float*, float*, int, bool,
                                       const int*, void*);
template int launch_resid_restrict<double, double, double>(const LevelDev<double>&, const double*, const double*,
                                                           const double*, double*, const int*, void*);
template int launch_resid_restrict<float, double, float>(const LevelDev<float>&, const float*, const double*,
                                                         const float*, float*, const int*, void*);
template int launch_resid_restrict<float, float, float>(const LevelDev<float>&, const float*, const float*,
                                                        const float*, float*, const int*, void*);
template int launch_resid_restrict<float, bf16s, bf16s>(const LevelDev<float>&, const bf16s*, const bf16s*,
                                                        const float*, float*, const int*, void*);
template int launch_restrict<double, double>(const LevelDev<double>&, const double*, double*, void*);
template int launch_restrict<float, float>(const LevelDev<float>&, const float*, float*, void*);
template int launch_restrict<double, float>(const LevelDev<float>&, const double*, float*, void*);
template int launch_prolong_add<double>(const LevelDev<double>&, const double*, double*, const int*, void*);
template int launch_prolong_add<float>(const LevelDev<float>&, const float*, float*, const int*, void*);
template int launch_convert<double, float>(size_t, const double*, float*, void*);
template int launch_convert<float, double>(size_t, const float*, double*, void*);
template int launch_convert<double, bf16s>(size_t, const double*, bf16s*, void*);
template int launch_residual_init<float>(const LevelDev<double>&, const OpPlan&, const double*, const double*,
                                         const double*, double*, float*, PcgState*, const Reducer&, void*);
template int launch_residual_init<bf16s>(const LevelDev<double>&, const OpPlan&, const double*, const double*,
                                         const double*, double*, bf16s*, PcgState*, const Reducer&, void*);
template int launch_pcg_update<float>(size_t, double*, double*, float*, const double*, const double*, PcgState*,
                                      const Reducer&, void*);
template int launch_pcg_update<bf16s>(size_t, double*, double*, bf16s*, const double*, const double*, PcgState*,
                                      const Reducer&, void*);
template int launch_tail<double>(const TailPlan<double>*, const int*, size_t, void*);
template int launch_tail<float>(const TailPlan<float>*, const int*, size_t, void*);

template <class T>
size_t tail_lds_capacity()
{
    int dev = 0, maxb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    const size_t avail = maxb > (int)(sizeof(TailPlan<T>) + 2048) ? (size_t)maxb - sizeof(TailPlan<T>) - 2048 : 0;
    if (avail == 0) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail_lds<T>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return avail;
}
template size_t tail_lds_capacity<float>();
template size_t tail_lds_capacity<double>();
template int launch_dense_setup<double>(const LevelDev<double>&, const double*, const double*, double*, double*,
                                        void*);
template int launch_dense_setup<float>(const LevelDev<double>&, const float*, const double*, double*, double*,
                                       void*);

}  // namespace tbfg
