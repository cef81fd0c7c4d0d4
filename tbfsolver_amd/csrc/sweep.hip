// sweep.hip -- one full red/black Gauss-Seidel sweep per launch (rbgs.f90:205-259:
// the red half-sweep, then the black one), out of place: pin holds the level's
// values before the sweep and is only read, pout receives the values after it.
//
// A workgroup owns a TX x TY tile of cells in the (x, y) plane and a chunk of
// planes in z, and marches through the chunk one plane at a time.  At plane
// step k it computes the red cells of plane k over the tile plus a one-cell
// ring and then the black cells of plane k-1 over the tile, which need the new
// red values of planes k-2, k-1, k.  The ring and the two extra red planes at
// the chunk ends are recomputed by every workgroup that needs them, from the
// old values in pin, so no workgroup waits on another and the result equals
// the two half-sweep launches exactly.
//
// Every input plane is staged once in LDS (p and beta with a two-cell ring, q
// with a one-cell ring), loaded into registers one step ahead so that the HBM
// latency of plane k+2 overlaps the arithmetic of plane k.  HBM traffic per
// cell: p in, beta, q, p out -- once per sweep, where two half-sweep launches
// stream every array twice.
#include "mg_kernels.hpp"
#include "mg_launch.hpp"
#include "reduce.hpp"
#include "smooth.hpp"
#include "tune.hpp"

#include <hip/hip_ext.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>

namespace tbfg {

namespace {

constexpr int kSweepMaxChunk = 58;  // planes per workgroup (z face factors staged in LDS)
constexpr int kZT = kSweepMaxChunk + 6;

template <class T, int TX, int TY, int NT>
struct SweepShape {
    static constexpr int PX = TX + 4, PY = TY + 4, PS = PX * PY;  // p, beta planes: tile + 2-cell ring
    static constexpr int QX = TX + 2, QY = TY + 2, QS = QX * QY;  // q, new-red planes: tile + 1-cell ring
    static constexpr int NP = (PS + NT - 1) / NT;
    static constexpr int NQ = (QS + NT - 1) / NT;
    static constexpr int NBLACK = TY * (TX / 2);                    // black cells of a tile plane
    // ring depth 4 for every staged plane (the plane steps run in groups of 4,
    // so the slot of a plane is a constant of the unrolled step): a slot is
    // refilled at least one full step after its last reader, so two barriers
    // per plane step suffice
    static constexpr int RING = 4;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int clampi(int j, int n) { return j < 0 ? 0 : (j >= n ? n - 1 : j); }

template <int NT, class T, int N>
__device__ __forceinline__ void stash_plane(const T (&reg)[N], T* __restrict__ dst, int S)
{
#pragma unroll
    for (int m = 0; m < N; ++m) dst[min((int)threadIdx.x + m * NT, S - 1)] = reg[m];
}

// raw buffer loads: a plane's loads share one scalar byte offset (the plane
// base) and differ by a fixed per-thread 32-bit offset, so no 64-bit address
// arithmetic per load; a load outside the array's bytes returns 0
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                             (int)(bytes < 0x7fffffffu ? bytes : 0x7fffffffu), 0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ bf16s buf_load<bf16s>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return bf16s{__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0)};
}

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

// Per-thread column history: the values a thread used at its tile cell two
// plane steps earlier (same colour parity, same column) that the next step of
// the same parity needs again.
template <class T>
struct ColHist {
    T bf;  // beta(k-1)
    T bc;  // beta(k-2)
    T pf;  // p_old(k-1)
    T r;   // new red value at k-2
};

// The tile height TY divides ny, so every thread stores on every step: no
// branch around the stores, which lets the compiler wait for the loads of a
// plane without waiting for the stores issued after them.
//
// Thread t owns the pair of columns (tx0 + 2h, tx0 + 2h + 1) of tile row
// t / (TX/2).  At plane step k its red cell of plane k and its black cell of
// plane k-1 share one column (both have gx + gy + k odd), which alternates
// between the two columns of the pair from step to step; the beta, p_old and
// red values of that column from two steps earlier stay in registers.  The
// red cells of the ring around the tile are a second, smaller round.
// DOTS (the finest level's last sweep of a preconditioner application): the
// sums of k_pcg_dots -- r.z, qv.z, sum z, sum r, sum qv of the new values z --
// accumulate while the plane is written, one partial per workgroup, instead of
// a separate pass over r, z and qv.  r and qv of a plane are loaded with the
// plane registers, two steps ahead.
// TQ, TB: the storage types of q and beta (bf16s on a bf16-stored finest
// level: the loads convert to T, LDS and arithmetic stay in T).
template <class T, class TQ, class TB, int TX, int TY, int NT, bool FIRST, bool DOTS, int PD>
__global__ void __launch_bounds__(NT)
    k_sweep(LevelDev<T> L, const TB* __restrict__ beta, const TQ* __restrict__ q, const T* __restrict__ pin,
            T* __restrict__ pout, int kc, int tiles_x, int ntiles, const int* done, SweepDots dots)
{
    using Sh = SweepShape<T, TX, TY, NT>;
    static_assert(Sh::NBLACK == NT, "one tile cell pair per thread");
    static_assert(PD % 2 == 0, "the register sets alternate with the colour parity of the steps");
    constexpr int PX = Sh::PX, PS = Sh::PS, QX = Sh::QX, QS = Sh::QS, QY = Sh::QY;
    constexpr int RING = Sh::RING;
    constexpr int NRING = QX + TY;  // red cells of the one-cell ring
    __shared__ T sP[(FIRST ? 1 : RING) * PS];
    __shared__ T sB[RING * PS];
    __shared__ T sQ[RING * QS];
    // new red values: two slots (the black cells of step k read those of step
    // k-1, and the slot is next written after the first barrier of step k+1)
    __shared__ T sR[2 * QS];
    // face factors times 1/2 (exact), so a face coefficient is (beta + beta') * f
    __shared__ T sTx[2 * QX];    // x, tile columns -1..TX
    __shared__ T sTy[2 * QY];    // y, tile rows -1..TY
    __shared__ T sTz[2 * kZT];  // z, planes k0-1.. of the chunk's plane steps

    if (done && *done) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    // consecutive tiles of a chunk on one XCD (blocks go round-robin to the 8 XCDs)
    unsigned id = blockIdx.x;
    if ((gridDim.x & 7u) == 0) id = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int tile = (int)(id % (unsigned)ntiles);
    const int chunk = (int)(id / (unsigned)ntiles);
    const int k0 = chunk * kc;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // DOTS
    if (k0 >= nz) {  // (the launch covers nz exactly; an empty chunk adds nothing)
        if (DOTS) block_partial<5>(acc, dots.red);
        return;
    }
    const int k1 = min(nz, k0 + kc);
    const int tx0 = (tile % tiles_x) * TX;
    const int ty0 = (tile / tiles_x) * TY;

    // plane steps k0-1 .. k1 (and up to 3 more, which compute red cells only)
    const int nsteps = (k1 - k0 + 2 + 3) & ~3;
    auto slot = [&](int k) { return (k - k0 + RING) & (RING - 1); };
    // the planes touched lie in [k0 - 2, k1 + 9] and nz >= 16
    auto modz = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };

    const T half = (T)0.5;
    for (int t = threadIdx.x; t < QX; t += NT) {
        const int gx = wrapi(tx0 + t - 1, nx);
        sTx[t] = half * L.a[0].ip[gx];
        sTx[QX + t] = half * L.a[0].im[gx];
    }
    for (int t = threadIdx.x; t < QY; t += NT) {
        const int gy = clampi(ty0 + t - 1, ny);
        sTy[t] = half * L.a[1].ip[gy];
        sTy[QY + t] = half * L.a[1].im[gy];
    }
    // (vector loads of these inside the plane loop would wait behind the
    // prefetch of the next planes)
    for (int t = threadIdx.x; t < nsteps + 1; t += NT) {
        const int kw = modz(k0 - 1 + t);
        sTz[t] = half * L.a[2].ip[kw];
        sTz[kZT + t] = half * L.a[2].im[kw];
    }

    // in-plane byte offsets of the cells this thread loads (fixed for the
    // chunk; unsigned, so a load is a uniform plane base plus this offset)
    unsigned offp[Sh::NP], offq[Sh::NQ];
#pragma unroll
    for (int m = 0; m < Sh::NP; ++m) {
        const int t = min((int)threadIdx.x + m * NT, PS - 1);
        const int yy = t / PX, xx = t - yy * PX;
        offp[m] = (unsigned)(nx * clampi(ty0 + yy - 2, ny) + wrapi(tx0 + xx - 2, nx)) * (unsigned)sizeof(T);
    }
#pragma unroll
    for (int m = 0; m < Sh::NQ; ++m) {
        const int t = min((int)threadIdx.x + m * NT, QS - 1);
        const int yy = t / QX, xx = t - yy * QX;
        offq[m] = (unsigned)(nx * clampi(ty0 + yy - 1, ny) + wrapi(tx0 + xx - 1, nx)) * (unsigned)sizeof(TQ);
    }
    const size_t plane_sz = (size_t)nx * ny;
    const size_t level_sz = plane_sz * nz;
    const __amdgpu_buffer_rsrc_t rs_p = plane_rsrc(pin, level_sz * sizeof(T));
    const __amdgpu_buffer_rsrc_t rs_b = plane_rsrc(beta, level_sz * sizeof(TB));
    const __amdgpu_buffer_rsrc_t rs_q = plane_rsrc(q, level_sz * sizeof(TQ));
    // this thread's tile cell pair and the column used at steps of each parity
    const int row = threadIdx.x / (TX / 2), h = threadIdx.x - row * (TX / 2);
    const int gy = ty0 + row;
    const bool wall_s = gy == 0, wall_n = gy == ny - 1;
    const size_t pair_off = (size_t)nx * gy + tx0 + 2 * h;

    // planes of p and beta at kp, of q at kp - 1, into registers.  The last
    // planes a chunk needs are p and beta at k1 + 1 and q at k1; a fetch for a
    // later plane step (the prefetch of the last steps, the red-only steps
    // that round the march up to a group of 4) loads that plane again, a
    // cache hit, where a load of the next planes would cost HBM traffic for
    // values no cell uses.
    const int klast = k1 + 1;
    auto fetch_pb = [&](T(&rp)[Sh::NP], T(&rb)[Sh::NP], int kp) {
        kp = min(kp, klast);
        const unsigned so = (unsigned)(plane_sz * modz(kp) * sizeof(T));
        const unsigned sob = (unsigned)(plane_sz * modz(kp) * sizeof(TB));
#pragma unroll
        for (int m = 0; m < Sh::NP; ++m) {
            if (!FIRST) rp[m] = buf_load<T>(rs_p, offp[m], so);
            rb[m] = (T)buf_load<TB>(rs_b, offp[m] / (unsigned)sizeof(T) * (unsigned)sizeof(TB), sob);
        }
    };
    // ... and, for DOTS, r and qv of this thread's pair on plane kp - 2 (the
    // black plane of the step that stashes these registers)
    auto fetch = [&](T(&rp)[Sh::NP], T(&rb)[Sh::NP], T(&rq)[Sh::NQ], double2& dr, double2& dq, int kp) {
        fetch_pb(rp, rb, kp);
        kp = min(kp, klast);
        const unsigned so = (unsigned)(plane_sz * modz(kp - 1) * sizeof(TQ));
#pragma unroll
        for (int m = 0; m < Sh::NQ; ++m) rq[m] = (T)buf_load<TQ>(rs_q, offq[m], so);
        if (DOTS) {
            const size_t o = plane_sz * modz(kp - 2) + pair_off;
            dr = *reinterpret_cast<const double2*>(dots.r + o);
            if (!dots.init) dq = *reinterpret_cast<const double2*>(dots.qv + o);
        }
    };
    auto stash = [&](const T(&rp)[Sh::NP], const T(&rb)[Sh::NP], int sl) {
        if (!FIRST) stash_plane<NT, T, Sh::NP>(rp, sP + sl * PS, PS);
        stash_plane<NT, T, Sh::NP>(rb, sB + sl * PS, PS);
    };

    // prologue: planes k0-2 and k0-1 of p and beta into LDS; the next two
    // planes into the two register sets (loads run two plane steps ahead)
    {
        T rp[Sh::NP], rb[Sh::NP];
        fetch_pb(rp, rb, k0 - 2);
        stash(rp, rb, slot(k0 - 2));
        fetch_pb(rp, rb, k0 - 1);
        stash(rp, rb, slot(k0 - 1));
    }
    // PD register sets: the loads of a plane are issued PD steps before the
    // step that stashes them
    T rps[PD][Sh::NP], rbs[PD][Sh::NP], rqs[PD][Sh::NQ];
    double2 drs[PD], dqs[PD];
#pragma unroll
    for (int u = 0; u < PD; ++u) {
        drs[u] = double2{0.0, 0.0};
        dqs[u] = double2{0.0, 0.0};
        fetch(rps[u], rbs[u], rqs[u], drs[u], dqs[u], k0 + u);
        // keep each set's loads ahead of the next set's: the wait before a
        // plane step then covers only the loads of the set it stashes
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // tables and planes k0-2, k0-1 in LDS

    const T ipy = sTy[row + 1], imy = sTy[QY + row + 1];

    // column of the steps with the parity of k: red (gx + gy + k) odd (nz is
    // even, so the parity of a plane index is that of its periodic image)
    auto colx = [&](int k) { return 2 * h + ((tx0 + gy + k + 1) & 1); };
    ColHist<T> hA, hB;  // steps k0-1, k0+1, ... and k0, k0+2, ...
    {
        const int xa = colx(k0 - 1), xb = colx(k0);
        hA.bf = sB[slot(k0 - 2) * PS + (row + 2) * PX + xa + 2];
        hB.bf = sB[slot(k0 - 1) * PS + (row + 2) * PX + xb + 2];
        hA.pf = FIRST ? (T)0 : sP[(FIRST ? 0 : slot(k0 - 2)) * PS + (row + 2) * PX + xa + 2];
        hB.pf = FIRST ? (T)0 : sP[(FIRST ? 0 : slot(k0 - 1)) * PS + (row + 2) * PX + xb + 2];
        hA.bc = hB.bc = hA.r = hB.r = (T)0;  // first read two steps later
    }

    // S: the ring slot of plane k (a constant of the unrolled step)
    auto step = [&](int k, int S, T(&rp)[Sh::NP], T(&rb)[Sh::NP], T(&rq)[Sh::NQ], double2& dr, double2& dq,
                    ColHist<T>& H) {
        const int Sp = (S + 1) & (RING - 1), Sm = (S + RING - 1) & (RING - 1);
        const int SpP = FIRST ? 0 : Sp, SP0 = FIRST ? 0 : S, SmP = FIRST ? 0 : Sm;
        // planes k+1 (p, beta) and k (q) into LDS, then start loading planes
        // k+PD+1 and k+PD into the same registers (past the chunk end: unused)
        stash(rp, rb, Sp);
        stash_plane<NT, T, Sh::NQ>(rq, sQ + S * QS, QS);
        const double2 r_pair = dr, qv_pair = dq;  // plane k-1
        fetch(rp, rb, rq, dr, dq, k + PD + 1);
        __syncthreads();

        const T* B0 = sB + S * PS;
        const T* Bp = sB + Sp * PS;
        const T* Bm = sB + Sm * PS;
        const T* P0 = sP + SP0 * PS;
        const T* Pp = sP + SpP * PS;
        const T* Pm = sP + SmP * PS;
        const T* Q0 = sQ + S * QS;
        T* R0 = sR + (S & 1) * QS;

        // ---- red cell of plane k at this thread's column
        const int xr = colx(k);
        const int ipc = (row + 2) * PX + xr + 2;
        const int iqc = (row + 1) * QX + xr + 1;
        const T ipx = sTx[xr + 1], imx = sTx[QX + xr + 1];
        const T p_self = H.pf;  // p_old(k-1) at this column
        T red_c, b_c, b_f;
        {
            const T azp = sTz[k - k0 + 1], azm = sTz[kZT + k - k0 + 1];
            const T bc = B0[ipc];
            const T bf = Bp[ipc];
            const T aR = (B0[ipc + 1] + bc) * ipx;
            const T aL = (bc + B0[ipc - 1]) * imx;
            const T aT = (B0[ipc + PX] + bc) * ipy;
            const T aB = (bc + B0[ipc - PX]) * imy;
            const T aF = (bf + bc) * azp;
            const T aK = (bc + H.bf) * azm;
            const T den = aR + aL + aT + aB + aF + aK;
            const T qc = Q0[iqc];
            T num;
            T pf = (T)0;
            if (FIRST) {
                num = -qc;
            } else {
                // rows outside the walls hold the wall row itself: the
                // Neumann ghost of the reference equals the cell's own value
                pf = Pp[ipc];
                num = -qc + (aR * P0[ipc + 1] + aL * P0[ipc - 1]) + (aT * P0[ipc + PX] + aB * P0[ipc - PX]) +
                      (aF * pf + aK * H.pf);
            }
            red_c = smooth_div(num, den);
            R0[iqc] = red_c;
            b_c = bc;
            b_f = bf;
            H.pf = pf;
        }
        // ---- red cells of the ring around the tile
        for (int t = threadIdx.x; t < NRING; t += NT) {
            int qrow, xx;
            if (t < QX) {  // rows -1 and TY, every second cell
                qrow = t < QX / 2 ? 0 : QY - 1;
                const int hh = t < QX / 2 ? t : t - QX / 2;
                xx = 2 * hh + ((tx0 + ty0 + qrow - 1 + k) & 1);
            } else if (t < QX + TY / 2) {  // column -1, every second row
                xx = 0;
                qrow = 1 + 2 * (t - QX) + ((tx0 + ty0 + k) & 1);
            } else {  // column TX
                xx = QX - 1;
                qrow = 1 + 2 * (t - QX - TY / 2) + ((tx0 + TX + ty0 + k + 1) & 1);
            }
            const int gyr = ty0 + qrow - 1;
            if (gyr < 0 || gyr >= ny) continue;
            const int ip = (qrow + 1) * PX + xx + 1;
            const T bc = B0[ip];
            const T aR = (B0[ip + 1] + bc) * sTx[xx];
            const T aL = (bc + B0[ip - 1]) * sTx[QX + xx];
            const T aT = (B0[ip + PX] + bc) * sTy[qrow];
            const T aB = (bc + B0[ip - PX]) * sTy[QY + qrow];
            const T aF = (Bp[ip] + bc) * sTz[k - k0 + 1];
            const T aK = (bc + Bm[ip]) * sTz[kZT + k - k0 + 1];
            const T den = aR + aL + aT + aB + aF + aK;
            const T qc = Q0[qrow * QX + xx];
            T num;
            if (FIRST) {
                num = -qc;
            } else {
                num = -qc + (aR * P0[ip + 1] + aL * P0[ip - 1]) + (aT * P0[ip + PX] + aB * P0[ip - PX]) +
                      (aF * Pp[ip] + aK * Pm[ip]);
            }
            R0[qrow * QX + xx] = smooth_div(num, den);
        }
        __syncthreads();

        // ---- black cell of plane k-1 at the same column; write the pair
        if (k > k0 && k <= k1) {
            const int kb = k - 1;
            const T azp = sTz[kb - k0 + 1], azm = sTz[kZT + kb - k0 + 1];
            const T* B1 = sB + Sm * PS;
            const T* Q1 = sQ + Sm * QS;
            const T* R1 = sR + (Sm & 1) * QS;
            const T bc = H.bf;  // beta(k-1) at this column
            const T aR = (B1[ipc + 1] + bc) * ipx;
            const T aL = (bc + B1[ipc - 1]) * imx;
            const T aT = (B1[ipc + PX] + bc) * ipy;
            const T aB = (bc + B1[ipc - PX]) * imy;
            const T aF = (b_c + bc) * azp;  // beta(k)
            const T aK = (bc + H.bc) * azm; // beta(k-2)
            const T den = aR + aL + aT + aB + aF + aK;
            const T self = FIRST ? (T)0 : p_self;
            const T pn = wall_n ? self : R1[iqc + QX];
            const T ps = wall_s ? self : R1[iqc - QX];
            const T num = -Q1[iqc] + (aR * R1[iqc + 1] + aL * R1[iqc - 1]) + (aT * pn + aB * ps) +
                          (aF * red_c + aK * H.r);
            const T vb = smooth_div(num, den);
            const int par = xr - 2 * h;
            const T vr = R1[iqc + 1 - 2 * par];  // the red cell sharing the aligned pair
            typename Vec2<T>::type v;
            v.x = par ? vr : vb;
            v.y = par ? vb : vr;
            *reinterpret_cast<typename Vec2<T>::type*>(pout + plane_sz * modz(kb) + pair_off) = v;
            if (DOTS) {  // the order of k_pcg_dots, cell by cell
                const double z0 = (double)v.x, z1 = (double)v.y;
                acc[0] += r_pair.x * z0;
                acc[2] += z0;
                acc[3] += r_pair.x;
                if (!dots.init) {
                    acc[1] += qv_pair.x * z0;
                    acc[4] += qv_pair.x;
                }
                acc[0] += r_pair.y * z1;
                acc[2] += z1;
                acc[3] += r_pair.y;
                if (!dots.init) {
                    acc[1] += qv_pair.y * z1;
                    acc[4] += qv_pair.y;
                }
            }
        }
        H.bc = b_c;
        H.bf = b_f;
        H.r = red_c;
    };

    // steps k0-1 .. in groups of 4 (slot of step k0-1+4m+u: (u + 3) & 3)
    for (int k = k0 - 1; k < k0 - 1 + nsteps; k += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            step(k + u, (u + 3) & 3, rps[u % PD], rbs[u % PD], rqs[u % PD], drs[u % PD], dqs[u % PD],
                 (u & 1) ? hB : hA);
    }
    if (DOTS) block_partial<5>(acc, dots.red);
}

// One full red/black sweep of a small level per launch (rbgs.f90:205-259),
// out of place like k_sweep.  A workgroup stages a BX x BY x BZ brick of p and
// beta with a two-cell ring and of q with a one-cell ring in LDS (a level this
// small stays in L2, so the ring's re-reads cost little), computes the red
// cells of the brick and its one-cell ring, then the black cells of the brick,
// and writes the brick.  Ring cells are recomputed from pin by every workgroup
// that needs them, so the launch equals the red and the black half-sweep
// launches, with the same arithmetic as rbgs_cell.  Where a plane march would
// take a few steps per workgroup (latency bound), all of a brick's planes load
// at once.
constexpr int kBrickThreads = 256;

template <class T, class TQ, int BX, int BY, int BZ, bool FIRST>
__global__ void __launch_bounds__(kBrickThreads)
    k_brick_sweep(LevelDev<T> L, const T* __restrict__ beta, const TQ* __restrict__ q, const T* __restrict__ pin,
                  T* __restrict__ pout, const int* done)
{
    constexpr int NT = kBrickThreads;
    constexpr int PX = BX + 4, PY = BY + 4, PZ = BZ + 4, PS = PX * PY * PZ;
    constexpr int QX = BX + 2, QY = BY + 2, QZ = BZ + 2, QS = QX * QY * QZ;
    static_assert(BX % 2 == 0, "cell pairs along x");
    __shared__ T sP[FIRST ? 1 : PS];
    __shared__ T sB[PS];
    __shared__ T sQ[QS];
    __shared__ T sR[QS];  // new red values, brick + one-cell ring
    __shared__ T sT[2 * (QX + QY + QZ)];  // face factors ip, im of the ring-1 rows/columns/planes

    if (done && *done) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int nbx = nx / BX, nby = ny / BY;
    const int bid = (int)blockIdx.x;
    const int x0 = (bid % nbx) * BX, y0 = ((bid / nbx) % nby) * BY, z0 = (bid / (nbx * nby)) * BZ;
    const size_t plane = (size_t)nx * ny;

    // every load of the brick issued before the first LDS store (one wait)
    constexpr int NPL = (PS + NT - 1) / NT, NQL = (QS + NT - 1) / NT;
    T rb[NPL], rp[NPL], rq[NQL];
#pragma unroll
    for (int m = 0; m < NPL; ++m) {
        const int t = min((int)threadIdx.x + m * NT, PS - 1);
        const int xx = t % PX, r = t / PX, yy = r % PY, zz = r / PY;
        const size_t g = plane * wrapi(z0 + zz - 2, nz) + (size_t)nx * clampi(y0 + yy - 2, ny) + wrapi(x0 + xx - 2, nx);
        rb[m] = beta[g];
        if (!FIRST) rp[m] = pin[g];
    }
#pragma unroll
    for (int m = 0; m < NQL; ++m) {
        const int t = min((int)threadIdx.x + m * NT, QS - 1);
        const int xx = t % QX, r = t / QX, yy = r % QY, zz = r / QY;
        rq[m] = (T)q[plane * wrapi(z0 + zz - 1, nz) + (size_t)nx * clampi(y0 + yy - 1, ny) + wrapi(x0 + xx - 1, nx)];
    }
#pragma unroll
    for (int m = 0; m < NPL; ++m) {
        const int t = min((int)threadIdx.x + m * NT, PS - 1);
        sB[t] = rb[m];
        if (!FIRST) sP[t] = rp[m];
    }
#pragma unroll
    for (int m = 0; m < NQL; ++m) sQ[min((int)threadIdx.x + m * NT, QS - 1)] = rq[m];
    for (int t = threadIdx.x; t < QX + QY + QZ; t += NT) {
        const T *fp, *fm;
        int g;
        if (t < QX) {
            fp = L.a[0].ip;
            fm = L.a[0].im;
            g = wrapi(x0 + t - 1, nx);
        } else if (t < QX + QY) {
            fp = L.a[1].ip;
            fm = L.a[1].im;
            g = clampi(y0 + t - QX - 1, ny);
        } else {
            fp = L.a[2].ip;
            fm = L.a[2].im;
            g = wrapi(z0 + t - QX - QY - 1, nz);
        }
        sT[t] = fp[g];
        sT[QX + QY + QZ + t] = fm[g];
    }
    __syncthreads();
    constexpr int TM = QX + QY + QZ;  // offset of the im factors
    const T half = (T)0.5;

    // red cells (gx + gy + gz odd) of the brick and its ring; ghost rows
    // beyond the walls are not cells (their values are never read)
    constexpr int NRED = (QX / 2) * QY * QZ;
    for (int t = threadIdx.x; t < NRED; t += NT) {
        const int hh = t % (QX / 2), r = t / (QX / 2), yy = r % QY, zz = r / QY;
        const int gy = y0 + yy - 1, gz = z0 + zz - 1;
        if (gy < 0 || gy >= ny) continue;
        const int xx = 2 * hh + ((x0 + gy + gz) & 1);  // ring-1 coordinates (nx, nz even)
        const int ip = ((zz + 1) * PY + yy + 1) * PX + xx + 1;
        const T bc = sB[ip];
        const T aR = half * (sB[ip + 1] + bc) * sT[xx];
        const T aL = half * (bc + sB[ip - 1]) * sT[TM + xx];
        const T aT = half * (sB[ip + PX] + bc) * sT[QX + yy];
        const T aB = half * (bc + sB[ip - PX]) * sT[TM + QX + yy];
        const T aF = half * (sB[ip + PX * PY] + bc) * sT[QX + QY + zz];
        const T aK = half * (bc + sB[ip - PX * PY]) * sT[TM + QX + QY + zz];
        const T den = aR + aL + aT + aB + aF + aK;
        const T qc = sQ[(zz * QY + yy) * QX + xx];
        T num;
        if (FIRST) {
            num = -qc;
        } else {
            // rows beyond the walls hold the wall row: the ghost is the cell itself
            num = -qc + (aR * sP[ip + 1] + aL * sP[ip - 1]) + (aT * sP[ip + PX] + aB * sP[ip - PX]) +
                  (aF * sP[ip + PX * PY] + aK * sP[ip - PX * PY]);
        }
        sR[(zz * QY + yy) * QX + xx] = smooth_div(num, den);
    }
    __syncthreads();

    // black cells of the brick, written with their red neighbour in x as pairs
    constexpr int NBLK = (BX / 2) * BY * BZ;
    for (int t = threadIdx.x; t < NBLK; t += NT) {
        const int hh = t % (BX / 2), r = t / (BX / 2), yy = r % BY, zz = r / BY;
        const int gy = y0 + yy, gz = z0 + zz;
        const int pr = (x0 + gy + gz + 1) & 1;  // red offset in the pair
        const int xb = 2 * hh + (pr ^ 1);       // brick coordinates of the black cell
        const int ip = ((zz + 2) * PY + yy + 2) * PX + xb + 2;
        const int iq = ((zz + 1) * QY + yy + 1) * QX + xb + 1;
        const int tx = xb + 1, ty = QX + yy + 1, tz = QX + QY + zz + 1;
        const T bc = sB[ip];
        const T aR = half * (sB[ip + 1] + bc) * sT[tx];
        const T aL = half * (bc + sB[ip - 1]) * sT[TM + tx];
        const T aT = half * (sB[ip + PX] + bc) * sT[ty];
        const T aB = half * (bc + sB[ip - PX]) * sT[TM + ty];
        const T aF = half * (sB[ip + PX * PY] + bc) * sT[tz];
        const T aK = half * (bc + sB[ip - PX * PY]) * sT[TM + tz];
        const T den = aR + aL + aT + aB + aF + aK;
        const T self = FIRST ? (T)0 : sP[ip];
        const T pn = gy == ny - 1 ? self : sR[iq + QX];
        const T ps = gy == 0 ? self : sR[iq - QX];
        const T num = -sQ[iq] + (aR * sR[iq + 1] + aL * sR[iq - 1]) + (aT * pn + aB * ps) +
                      (aF * sR[iq + QX * QY] + aK * sR[iq - QX * QY]);
        const T vb = smooth_div(num, den);
        const T vr = sR[iq + 1 - 2 * (pr ^ 1)];  // the red cell of the pair
        typename Vec2<T>::type v;
        v.x = pr ? vb : vr;
        v.y = pr ? vr : vb;
        *reinterpret_cast<typename Vec2<T>::type*>(pout + plane * gz + (size_t)nx * gy + x0 + 2 * hh) = v;
    }
}

}  // namespace

template <class T>
static void brick_dims(int& bx, int& by, int& bz)
{
    bx = 16;
    by = 8;
    bz = sizeof(T) == 4 ? 8 : 4;
}

template <class T>
bool brick_supported(const LevelDev<T>& L)
{
    int bx, by, bz;
    brick_dims<T>(bx, by, bz);
    return L.nx % bx == 0 && L.ny % by == 0 && L.nz % bz == 0 && L.nz % 2 == 0;
}

template <class T, class TQ>
int launch_brick_sweep(const LevelDev<T>& L, const T* beta, const TQ* q, const T* pin, T* pout, bool first,
                       const int* done, void* stream)
{
    if (!brick_supported(L)) {
        set_error("launch_brick_sweep: level %dx%dx%d not supported", L.nx, L.ny, L.nz);
        return TBFG_EINVAL;
    }
    int bx, by, bz;
    brick_dims<T>(bx, by, bz);
    const dim3 grid((unsigned)((L.nx / bx) * (L.ny / by) * (L.nz / bz)));
    constexpr int BZ = sizeof(T) == 4 ? 8 : 4;
    if (first)
        hipLaunchKernelGGL((k_brick_sweep<T, TQ, 16, 8, BZ, true>), grid, dim3(kBrickThreads), 0, (hipStream_t)stream,
                           L, beta, q, pin, pout, done);
    else
        hipLaunchKernelGGL((k_brick_sweep<T, TQ, 16, 8, BZ, false>), grid, dim3(kBrickThreads), 0,
                           (hipStream_t)stream, L, beta, q, pin, pout, done);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

struct SweepCfg {
    int tx, ty, nt;
};

// tile shapes with an instantiated kernel, in order of preference: in FP32
// the plain sweep runs fastest on 32x16 tiles (a smaller ring per cell), the
// sweep from zero and the one with the dot products on 64x8 (measured)
constexpr SweepCfg kShapesF[] = {{64, 8, 256}, {32, 16, 256}, {64, 16, 512}, {32, 8, 128}};
constexpr SweepCfg kShapesFPlain[] = {{32, 16, 256}, {64, 8, 256}, {64, 16, 512}, {32, 8, 128}};
constexpr SweepCfg kShapesD[] = {{64, 4, 128}, {64, 8, 256}, {32, 8, 128}};

template <class T>
static bool pick_shape(const LevelDev<T>& L, SweepCfg& out, bool plain = false)
{
    // nz even (the colour of a plane is its index's parity), nz >= 16 (a
    // chunk's plane steps wrap around z at most once); the level's arrays
    // addressable by 32-bit buffer offsets (< 2 GiB)
    if (L.nz < 16 || L.nz % 2 != 0 || L.ny % 2 != 0) return false;
    if ((size_t)L.nx * L.ny * L.nz * sizeof(T) >= ((size_t)1 << 31)) return false;
    const SweepCfg* tab = sizeof(T) == 4 ? (plain ? kShapesFPlain : kShapesF) : kShapesD;
    const int n = sizeof(T) == 4 ? (int)(sizeof kShapesF / sizeof kShapesF[0]) : (int)(sizeof kShapesD / sizeof kShapesD[0]);
    int want_tx = 0, want_ty = 0, want_nt = 0;
    int tuned[3];
    if (const char* e = std::getenv("TBFG_SWEEP_SHAPE"))
        std::sscanf(e, "%dx%dx%d", &want_tx, &want_ty, &want_nt);
    else if (tune_lookup(plain ? TBFG_TUNE_SWEEP : TBFG_TUNE_SWEEP_EDGE, L.nx, L.ny, L.nz, (int)sizeof(T), tuned)) {
        want_tx = tuned[0];
        want_ty = tuned[1];
        want_nt = tuned[2];
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < n; ++i) {
            const SweepCfg c = tab[i];
            if (pass == 0 && !(c.tx == want_tx && c.ty == want_ty && c.nt == want_nt)) continue;
            if (L.nx % c.tx == 0 && L.ny % c.ty == 0) {
                out = c;
                return true;
            }
        }
    return false;
}

template <class T>
bool sweep_supported(const LevelDev<T>& L)
{
    SweepCfg c;
    return pick_shape(L, c);
}

// prefetch distance in plane steps: 4 in FP32 on long chunks; 2 in FP64, with
// the dot products (registers) and on short chunks (whose steps would not
// fill a group of 4)
template <class T, bool DOTS>
constexpr int sweep_pd_max()
{
    return sizeof(T) == 4 && !DOTS ? 4 : 2;
}

template <class T, class TQ, class TB, int TX, int TY, int NT, bool FIRST, bool DOTS>
static int launch_sweep_shape(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* pin, T* pout,
                              const int* done, const SweepDots& dots, unsigned* nblocks_out, void* stream,
                              const KernelEvents* kev)
{
    constexpr int pdm = sweep_pd_max<T, DOTS>();
    // resident workgroups per CU and CU count, queried once (threads may race
    // to fill it; they store the same value)
    static std::atomic<int64_t> occ{-1};
    int64_t o = occ.load(std::memory_order_acquire);
    if (o < 0) {
        int dev = 0, ncu_q = 0, nb = 0;
        TBFG_HIP(hipGetDevice(&dev));
        TBFG_HIP(hipDeviceGetAttribute(&ncu_q, hipDeviceAttributeMultiprocessorCount, dev));
        TBFG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<T, TQ, TB, TX, TY, NT, FIRST, DOTS, pdm>, NT, 0));
        o = ((int64_t)(nb < 1 ? 1 : nb) << 32) | (uint32_t)ncu_q;
        occ.store(o, std::memory_order_release);
    }
    const int per_cu = (int)(o >> 32), ncu = (int)(o & 0xffffffff);
    const int tiles_x = L.nx / TX;
    const int tiles_y = L.ny / TY;
    const int ntiles = tiles_x * tiles_y;
    // one resident wave of workgroups, chunks of at least 4 planes
    const int target = per_cu * (ncu > 0 ? ncu : 256);
    int chunks = (target + ntiles - 1) / ntiles;
    int kc = (L.nz + chunks - 1) / chunks;
    // kc + 2 plane steps run in groups of 4: kc = 4m + 2, from 6 to kSweepMaxChunk
    if (const char* e = std::getenv("TBFG_SWEEP_KC")) kc = std::atoi(e);  // A/B switch (planes per chunk)
    kc = ((kc + 2 + 3) / 4) * 4 - 2;
    if (kc < 6) kc = 6;
    if (kc > kSweepMaxChunk) kc = kSweepMaxChunk;
    const int pd = (pdm == 4 && kc >= 16) ? 4 : 2;
    chunks = (L.nz + kc - 1) / kc;
    const unsigned nblocks = (unsigned)ntiles * (unsigned)chunks;
    if (DOTS && (size_t)nblocks * 5 > dots.capacity) {
        set_error("launch_sweep: %u workgroups exceed the reduction scratch", nblocks);
        return TBFG_EINVAL;
    }
    auto k = pd == 4 ? k_sweep<T, TQ, TB, TX, TY, NT, FIRST, DOTS, pdm> : k_sweep<T, TQ, TB, TX, TY, NT, FIRST, DOTS, 2>;
    if (kev)
        hipExtLaunchKernelGGL(k, dim3(nblocks), dim3(NT), 0, (hipStream_t)stream, kev->start, kev->stop, 0, L, beta,
                              q, pin, pout, kc, tiles_x, ntiles, done, dots);
    else
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), 0, (hipStream_t)stream, L, beta, q, pin, pout, kc, tiles_x,
                           ntiles, done, dots);
    TBFG_HIP(hipGetLastError());
    if (nblocks_out) *nblocks_out = nblocks;
    return TBFG_OK;
}

template <class T, class TQ, class TB, int TX, int TY, int NT>
static int launch_sweep_first(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* pin, T* pout, bool first,
                              const int* done, const SweepDots* dots, unsigned* nblocks_out, void* stream,
                              const KernelEvents* kev)
{
    const SweepDots none{};
    if (dots)  // the dot products follow a post-smoothing sweep, never the first
        return launch_sweep_shape<T, TQ, TB, TX, TY, NT, false, true>(L, beta, q, pin, pout, done, *dots, nblocks_out,
                                                                  stream, kev);
    return first ? launch_sweep_shape<T, TQ, TB, TX, TY, NT, true, false>(L, beta, q, pin, pout, done, none,
                                                                      nblocks_out, stream, kev)
                 : launch_sweep_shape<T, TQ, TB, TX, TY, NT, false, false>(L, beta, q, pin, pout, done, none,
                                                                       nblocks_out, stream, kev);
}

template <class T, class TQ, class TB>
int launch_sweep(const LevelDev<T>& L, const TB* beta, const TQ* q, const T* pin, T* pout, bool first,
                 const int* done, void* stream, const SweepDots* dots, unsigned* nblocks_out,
                 const KernelEvents* kev)
{
    if (dots && first) {
        set_error("launch_sweep: the fused dot products need a sweep from a non-zero start");
        return TBFG_EINVAL;
    }
    SweepCfg c;
    if (!pick_shape(L, c, !first && !dots)) {
        set_error("launch_sweep: level %dx%dx%d not supported", L.nx, L.ny, L.nz);
        return TBFG_EINVAL;
    }
#define TBFG_SWEEP_CASE(X, Y, N)                                                             \
    if (c.tx == X && c.ty == Y && c.nt == N)                                                \
        return launch_sweep_first<T, TQ, TB, X, Y, N>(L, beta, q, pin, pout, first, done, dots, nblocks_out, stream, kev);
    if constexpr (sizeof(T) == 4) {
        TBFG_SWEEP_CASE(64, 8, 256)
        TBFG_SWEEP_CASE(64, 16, 512)
        TBFG_SWEEP_CASE(32, 16, 256)
        TBFG_SWEEP_CASE(32, 8, 128)
    } else {
        TBFG_SWEEP_CASE(64, 4, 128)
        TBFG_SWEEP_CASE(64, 8, 256)
        TBFG_SWEEP_CASE(32, 8, 128)
    }
#undef TBFG_SWEEP_CASE
    set_error("launch_sweep: no kernel for shape %dx%dx%d", c.tx, c.ty, c.nt);
    return TBFG_EINVAL;
}

template bool brick_supported<float>(const LevelDev<float>&);
template bool brick_supported<double>(const LevelDev<double>&);
template int launch_brick_sweep<float, float>(const LevelDev<float>&, const float*, const float*, const float*,
                                              float*, bool, const int*, void*);
template int launch_brick_sweep<double, double>(const LevelDev<double>&, const double*, const double*,
                                                const double*, double*, bool, const int*, void*);
template bool sweep_supported<float>(const LevelDev<float>&);
template bool sweep_supported<double>(const LevelDev<double>&);
template int launch_sweep<float, float, float>(const LevelDev<float>&, const float*, const float*, const float*,
                                               float*, bool, const int*, void*, const SweepDots*, unsigned*,
                                               const KernelEvents*);
template int launch_sweep<float, bf16s, bf16s>(const LevelDev<float>&, const bf16s*, const bf16s*, const float*,
                                               float*, bool, const int*, void*, const SweepDots*, unsigned*,
                                               const KernelEvents*);
template int launch_sweep<double, double, double>(const LevelDev<double>&, const double*, const double*, const double*,
                                          double*, bool, const int*, void*, const SweepDots*, unsigned*,
                                          const KernelEvents*);

}  // namespace tbfg
