// fps.hip -- FAST_MODE pressure solve on one GPU (fastPoissonSolver.f90:337-423).
//
// Five HBM passes:
//   K1 x-forward : real lines -> half spectrum (two real lines packed into one
//                  complex FFT), written as A[kx][j][k]            (r2c, :350-353)
//   K2 z-forward : A[kx][j][:] -> B[kx][kz][j]                     (:356-357)
//   K3 y-solve   : one wavefront per (kx,kz) column of B, partitioned Thomas
//                  + parallel cyclic reduction across the 64 lanes (:280-333)
//   K4 z-inverse : B -> A                                           (:366-367)
//   K5 x-inverse : A -> real lines, scaled by 1/(nx nz)             (:370-374)
// K1/K2/K4/K5 run two-level kernels (k_*_tl: in-register DFTs of the two
// factors of the length, one LDS exchange per line) for lines of 32-256 points
// (tl_split) and LDS-staged Stockham passes (fft_lines) for the other lengths.
// The pencil transposes of pencilDec.f90 become the layout changes folded into
// the loads and stores of K1/K2/K4/K5.  The volume-weighted mean removal
// (:388-423) is done exactly in spectral space on the (0,0) column, where
// sum_ijk V p = dx dz sum_j dyf_j P00(j).
//
// Mixed precision: K1-K5 in FP32, the (0,0) column in FP64, followed by FP64
// iterative-refinement sweeps p += S32(s - L p) with the exact FFT operator L.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <type_traits>
#include <vector>

#include "fps.hpp"
#include "reduce.hpp"

namespace tbfg {

template <class T>
struct C2 {
    T x, y;
};

template <class T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }

constexpr int kFftThreads = 256;
// butterflies per thread per pass: every butterfly's inputs are held in
// registers across the pass's barrier, so FP64 holds fewer (4 x 4 complex
// values, 64 registers) to keep the kernels clear of spills
template <class T>
constexpr int max_bfly() { return sizeof(T) == 8 ? 4 : 8; }
constexpr int kLdsElems = 4096;    // complex elements (lines + twiddles) per block: 64 KB in FP64
constexpr int kLdsElemsMax = 10240;  // lines longer than 1024 points, one per block: 160 KB in FP64

// Radices of each Stockham pass, and per pass the multipliers that divide by
// the pass's butterfly count N/R and its sub-transform length Ns with one
// multiply-high: floor(x/d) = mulhi(x, ceil(2^32/d)), exact for every divisor
// of these lengths (<= 1024) and x < 2^15 (checked exhaustively); d = 1 is the
// identity.  A runtime integer division costs ~20 instructions, and the index
// arithmetic of three of them per butterfly dominated the passes.
struct FftPlan {
    int n;
    int nrad;
    int rad[16];
    int swz;           // XOR mask of the LDS line swizzle (lds_pos), 0 when n % 8 != 0
    uint32_t mnb[16];  // ceil(2^32 / (n / rad[s]))
    uint32_t mns[16];  // ceil(2^32 / Ns), Ns = rad[0] * ... * rad[s-1]
};

static uint32_t div_magic(int d) { return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d); }

static int make_plan(int n, FftPlan& p)
{
    p.n = n;
    p.nrad = 0;
    int m = n;
    while (m % 4 == 0) {
        p.rad[p.nrad++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        p.rad[p.nrad++] = 2;
        m /= 2;
    }
    while (m % 3 == 0) {
        p.rad[p.nrad++] = 3;
        m /= 3;
    }
    p.swz = n % 8 == 0 ? 6 : 0;
    int Ns = 1;
    for (int st = 0; st < p.nrad; ++st) {
        p.mnb[st] = div_magic(n / p.rad[st]);
        p.mns[st] = div_magic(Ns);
        Ns *= p.rad[st];
    }
    return m == 1 ? TBFG_OK : TBFG_EINVAL;
}

__device__ __forceinline__ int fast_div(int x, int d, uint32_t magic)
{
    return d == 1 ? x : (int)__umulhi((uint32_t)x, magic);
}

// R-point DFT in registers, SIGN = -1 forward, +1 backward
template <class T, int SIGN>
__device__ __forceinline__ void dft2(C2<T>* v)
{
    const C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <class T, int SIGN>
__device__ __forceinline__ void dft4(C2<T>* v)
{
    const C2<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const C2<T> t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // SIGN*i*d: forward multiplies by -i, backward by +i
    const C2<T> t3 = (SIGN < 0) ? C2<T>{d.y, -d.x} : C2<T>{-d.y, d.x};
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <class T, int SIGN>
__device__ __forceinline__ void dft3(C2<T>* v)
{
    const T h = (T)0.8660254037844386;  // sqrt(3)/2
    const C2<T> t1 = cadd(v[1], v[2]);
    const C2<T> t2 = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    const C2<T> dd = csub(v[1], v[2]);
    // i * SIGN * h * dd
    const C2<T> t3 = (SIGN < 0) ? C2<T>{h * dd.y, -h * dd.x} : C2<T>{-h * dd.y, h * dd.x};
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

// LDS position of element i of line l.  Lines are N+1 elements apart, so the
// same index of consecutive lines -- the transposed side of the passes' loads
// and stores -- falls in different banks.  Within a line, index bits 1-2 are
// XORed with (bit 4 ^ bit 5) * swz: a permutation inside each aligned group of
// 8 (contiguous accesses keep their banks) that takes the bank conflicts out of
// the Stockham writes, whose lanes step by R*Ns elements while Ns < 8 (gfx950
// serves 16-byte accesses 8 lanes at a time over 64 four-byte banks).
__device__ __forceinline__ int lds_pos(int l, int LS, int i, int swz)
{
    return l * LS + (i ^ (((i >> 4) ^ (i >> 5)) & 1) * swz);
}

// Stockham autosort FFT of `nlines` lines of length N held in LDS at lds_pos.
// tw holds each pass's twiddles contiguously: pass s with sub-transform length
// Ns and radix R uses tw[Ns - 1 + (r-1)*Ns + k] = exp(-2 pi i r k / (Ns R)),
// r = 1..R-1, k = 0..Ns-1 (N-1 entries in all), so consecutive lanes read
// consecutive entries.  Each pass reads all its butterfly inputs into
// registers, synchronises, then writes.
template <class T, int SIGN>
__device__ void fft_lines(C2<T>* buf, int N, int nlines, const C2<T>* tw, const FftPlan& plan)
{
    const int LS = N + 1;
    const int swz = plan.swz;
    int Ns = 1;
    for (int s = 0; s < plan.nrad; ++s) {
        const int R = plan.rad[s];
        const int nb = N / R;
        const int total = nlines * nb;
        const C2<T>* tws = tw + Ns - 1;
        const uint32_t mnb = plan.mnb[s], mns = plan.mns[s];
        constexpr int kMaxBfly = max_bfly<T>();
        C2<T> v[kMaxBfly][4];
#pragma unroll
        for (int u = 0; u < kMaxBfly; ++u) {
            const int b = threadIdx.x + u * kFftThreads;
            if (b < total) {
                const int line = fast_div(b, nb, mnb), j = b - line * nb;
                const int k = j - fast_div(j, Ns, mns) * Ns;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (r < R) {
                        C2<T> x = buf[lds_pos(line, LS, j + r * nb, swz)];
                        if (r > 0 && k > 0) {
                            C2<T> w = tws[(r - 1) * Ns + k];
                            if (SIGN > 0) w.y = -w.y;
                            x = cmul(x, w);
                        }
                        v[u][r] = x;
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kMaxBfly; ++u) {
            const int b = threadIdx.x + u * kFftThreads;
            if (b < total) {
                const int line = fast_div(b, nb, mnb), j = b - line * nb;
                const int jq = fast_div(j, Ns, mns);
                const int k = j - jq * Ns;
                if (R == 4)
                    dft4<T, SIGN>(v[u]);
                else if (R == 2)
                    dft2<T, SIGN>(v[u]);
                else
                    dft3<T, SIGN>(v[u]);
                const int d0 = jq * Ns * R + k;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (r < R) buf[lds_pos(line, LS, d0 + r * Ns, swz)] = v[u][r];
            }
        }
        __syncthreads();
        Ns *= R;
    }
}

template <class T>
__device__ __forceinline__ void load_twiddles(C2<T>* dst, const C2<T>* __restrict__ src, int n)
{
    for (int t = threadIdx.x; t < n; t += kFftThreads) dst[t] = src[t];
}

// Staging loops issue the loads of kStageBatch elements per thread before the
// first LDS store, so the HBM latency is paid once per batch rather than once
// per element (a rolled loop of load-then-store waits on every load).  The
// loads are unconditional (a thread past the end re-reads the last element)
// and only the LDS stores are guarded: a guarded load is compiled as a branch
// that waits for its own load before the next one is issued.
template <class T>
constexpr int stage_batch() { return sizeof(T) == 8 ? 4 : 8; }

// K1: s(:, j, k0..k0+KT-1) -> A[kx][j][k], kx = 0..nx/2.
template <class T, class TIN>
__global__ void __launch_bounds__(kFftThreads)
    k_xfwd(int nx, int ny, int nz, int KT, FftPlan plan, const C2<T>* __restrict__ tw, const TIN* __restrict__ s,
           C2<T>* __restrict__ A)
{
    constexpr int kStageBatch = stage_batch<T>();
    extern __shared__ unsigned char smem[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(smem);
    const int L = KT >> 1;
    const int LS = nx + 1;  // padded line stride (fft_lines)
    C2<T>* twl = buf + L * LS;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * KT;
    load_twiddles(twl, tw, nx);
    for (int t0 = threadIdx.x; t0 < L * nx; t0 += kStageBatch * kFftThreads) {
        TIN v0[kStageBatch], v1[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = min(t0 + u * kFftThreads, L * nx - 1);
            const int l = t / nx, i = t - l * nx;
            const size_t o0 = i + (size_t)nx * (j + (size_t)ny * (k0 + 2 * l));
            v0[u] = s[o0];
            v1[u] = s[o0 + (size_t)nx * ny];
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < L * nx) {
                const int l = t / nx;
                buf[lds_pos(l, LS, t - l * nx, plan.swz)] = C2<T>{(T)v0[u], (T)v1[u]};
            }
        }
    }
    __syncthreads();
    fft_lines<T, -1>(buf, nx, L, twl, plan);
    const int nh = nx / 2 + 1;
    for (int t = threadIdx.x; t < nh * KT; t += kFftThreads) {
        const int m = t / KT, kk = t - m * KT;
        const int l = kk >> 1;
        const C2<T> Z = buf[lds_pos(l, LS, m, plan.swz)];
        const C2<T> Zn = buf[lds_pos(l, LS, m == 0 ? 0 : nx - m, plan.swz)];
        C2<T> val;
        if ((kk & 1) == 0)
            val = C2<T>{(T)0.5 * (Z.x + Zn.x), (T)0.5 * (Z.y - Zn.y)};  // (Z + conj Zn)/2
        else
            val = C2<T>{(T)0.5 * (Z.y + Zn.y), (T)0.5 * (Zn.x - Z.x)};  // (Z - conj Zn)/(2i)
        A[((size_t)m * ny + j) * nz + k0 + kk] = val;
    }
}

// A 1-D grid's workgroup index with consecutive workgroups on one XCD
// (workgroups go round-robin to the 8 XCDs, each with its own L2): the z passes
// number the y chunks of one kx consecutively, so the chunks that write (or
// read) the neighbouring runs of a column share an L2 instead of each XCD
// moving partial lines.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned nb = gridDim.x;
    const unsigned b = blockIdx.x;
    return (nb & 7u) == 0 ? (b & 7u) * (nb >> 3) + (b >> 3) : b;
}

// K2: A[kx][j0..j0+JT-1][:] -> B[kx][kz][j]
template <class T>
__global__ void __launch_bounds__(kFftThreads)
    k_zfwd(int ny, int nz, int JT, FftPlan plan, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ A,
           C2<T>* __restrict__ B)
{
    constexpr int kStageBatch = stage_batch<T>();
    extern __shared__ unsigned char smem[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(smem);
    const int LS = nz + 1;  // padded line stride (fft_lines)
    C2<T>* twl = buf + JT * LS;
    const unsigned id = xcd_block(), gj = (unsigned)(ny / JT);
    const int m = (int)(id / gj);
    const int j0 = (int)(id % gj) * JT;
    load_twiddles(twl, tw, nz);
    const C2<T>* src = A + ((size_t)m * ny + j0) * nz;
    for (int t0 = threadIdx.x; t0 < JT * nz; t0 += kStageBatch * kFftThreads) {
        // real and imaginary parts in scalar arrays: an array of C2<double> is
        // kept in scratch memory, each element loaded, waited for and spilled
        T vr[kStageBatch], vi[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = min(t0 + u * kFftThreads, JT * nz - 1);
            const C2<T> x = src[t];
            vr[u] = x.x;
            vi[u] = x.y;
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < JT * nz) {
                const int l = t / nz;
                buf[lds_pos(l, LS, t - l * nz, plan.swz)] = C2<T>{vr[u], vi[u]};
            }
        }
    }
    __syncthreads();
    fft_lines<T, -1>(buf, nz, JT, twl, plan);
    for (int t = threadIdx.x; t < JT * nz; t += kFftThreads) {
        const int kz = t / JT, l = t - kz * JT;
        B[((size_t)m * nz + kz) * ny + j0 + l] = buf[lds_pos(l, LS, kz, plan.swz)];
    }
}

// K4: B[kx][kz][j0..] -> A[kx][j][k]
template <class T>
__global__ void __launch_bounds__(kFftThreads)
    k_zinv(int ny, int nz, int JT, FftPlan plan, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ B,
           C2<T>* __restrict__ A)
{
    constexpr int kStageBatch = stage_batch<T>();
    extern __shared__ unsigned char smem[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(smem);
    const int LS = nz + 1;  // padded line stride (fft_lines)
    C2<T>* twl = buf + JT * LS;
    const unsigned id = xcd_block(), gj = (unsigned)(ny / JT);
    const int m = (int)(id / gj);
    const int j0 = (int)(id % gj) * JT;
    load_twiddles(twl, tw, nz);
    for (int t0 = threadIdx.x; t0 < JT * nz; t0 += kStageBatch * kFftThreads) {
        T vr[kStageBatch], vi[kStageBatch];  // scalar arrays, as in k_zfwd
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = min(t0 + u * kFftThreads, JT * nz - 1);
            const int kz = t / JT, l = t - kz * JT;
            const C2<T> x = B[((size_t)m * nz + kz) * ny + j0 + l];
            vr[u] = x.x;
            vi[u] = x.y;
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < JT * nz) {
                const int kz = t / JT, l = t - kz * JT;
                buf[lds_pos(l, LS, kz, plan.swz)] = C2<T>{vr[u], vi[u]};
            }
        }
    }
    __syncthreads();
    fft_lines<T, +1>(buf, nz, JT, twl, plan);
    C2<T>* dst = A + ((size_t)m * ny + j0) * nz;
    for (int t = threadIdx.x; t < JT * nz; t += kFftThreads) {
        const int l = t / nz;
        dst[t] = buf[lds_pos(l, LS, t - l * nz, plan.swz)];
    }
}

// K5: A[kx][j][k0..k0+KT-1] -> p(:, j, k) scaled by 1/(nx nz).
// mode 0: p = val; mode 1: p = p + val - shift[0] (refinement correction).
template <class T>
__global__ void __launch_bounds__(kFftThreads)
    k_xinv(int nx, int ny, int nz, int KT, FftPlan plan, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ A,
           double* __restrict__ p, int mode, const double* __restrict__ shift)
{
    constexpr int kStageBatch = stage_batch<T>();
    extern __shared__ unsigned char smem[];
    C2<T>* buf = reinterpret_cast<C2<T>*>(smem);
    const int L = KT >> 1;
    const int LS = nx + 1;  // padded line stride (fft_lines)
    C2<T>* twl = buf + L * LS;
    const int j = blockIdx.y;
    const int k0 = blockIdx.x * KT;
    const int nh = nx / 2 + 1;
    load_twiddles(twl, tw, nx);
    // Z_l[m] = X1[m] + i X2[m] with Hermitian extension X[nx-m] = conj X[m];
    // DC and Nyquist keep their real parts only, as a c2r transform does.
    constexpr int kB = kStageBatch / 2;  // two complex loads per element
    for (int t0 = threadIdx.x; t0 < nh * L; t0 += kB * kFftThreads) {
        T x1r[kB], x1i[kB], x2r[kB], x2i[kB];  // scalar arrays, as in k_zfwd
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = min(t0 + u * kFftThreads, nh * L - 1);
            const int m = t / L, l = t - m * L;
            const size_t o = ((size_t)m * ny + j) * nz + k0 + 2 * l;
            const C2<T> a = A[o], b = A[o + 1];
            x1r[u] = a.x;
            x1i[u] = a.y;
            x2r[u] = b.x;
            x2i[u] = b.y;
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < nh * L) {
                const int m = t / L, l = t - m * L;
                const C2<T> x1{x1r[u], x1i[u]}, x2{x2r[u], x2i[u]};
                if (m == 0 || 2 * m == nx) {
                    buf[lds_pos(l, LS, m, plan.swz)] = C2<T>{x1.x, x2.x};
                } else {
                    buf[lds_pos(l, LS, m, plan.swz)] = C2<T>{x1.x - x2.y, x1.y + x2.x};
                    buf[lds_pos(l, LS, nx - m, plan.swz)] = C2<T>{x1.x + x2.y, x2.x - x1.y};
                }
            }
        }
    }
    __syncthreads();
    fft_lines<T, +1>(buf, nx, L, twl, plan);
    const double scale = 1.0 / ((double)nx * (double)nz);
    const double sh = (mode == 1) ? shift[0] : 0.0;
    if (mode == 0) {
        for (int t = threadIdx.x; t < L * nx; t += kFftThreads) {
            const int l = t / nx, i = t - l * nx;
            const size_t o0 = i + (size_t)nx * (j + (size_t)ny * (k0 + 2 * l));
            const C2<T> z = buf[lds_pos(l, LS, i, plan.swz)];
            p[o0] = (double)z.x * scale;
            p[o0 + (size_t)nx * ny] = (double)z.y * scale;
        }
        return;
    }
    // the correction: p's old values loaded a batch at a time
    for (int t0 = threadIdx.x; t0 < L * nx; t0 += kStageBatch * kFftThreads) {
        double a0[kStageBatch], a1[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < L * nx) {
                const int l = t / nx, i = t - l * nx;
                const size_t o0 = i + (size_t)nx * (j + (size_t)ny * (k0 + 2 * l));
                a0[u] = p[o0];
                a1[u] = p[o0 + (size_t)nx * ny];
            }
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int t = t0 + u * kFftThreads;
            if (t < L * nx) {
                const int l = t / nx, i = t - l * nx;
                const size_t o0 = i + (size_t)nx * (j + (size_t)ny * (k0 + 2 * l));
                const C2<T> z = buf[lds_pos(l, LS, i, plan.swz)];
                p[o0] = a0[u] + (double)z.x * scale - sh;
                p[o0 + (size_t)nx * ny] = a1[u] + (double)z.y * scale - sh;
            }
        }
    }
}

// ------------------------------------------- two-level passes, N = N1 x N2
//
// Lines of N = N1 * N2 points (N1, N2 in {4, 8, 12, 16}; tl_split) skip the
// Stockham passes in LDS.  With n = N2 n1 + n2 and k = k1 + N1 k2,
//   X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) [ W_N^(n2 k1) sum_n1 x[N2 n1 + n2] W_N1^(n1 k1) ].
// Step A: thread n2 of a line holds x[N2 n1 + n2], n1 = 0..N1-1, in registers
// (loaded from HBM, or from the pre-processed line in LDS), runs the inner
// N1-point DFT, multiplies by W_N^(n2 k1) and writes Y[k1][n2] to LDS.
// Step B: thread k1 reads Y[k1][0..N2-1] and runs the outer N2-point DFT,
// which leaves X[k1 + N1 k2] in its registers, stored straight to HBM (z
// passes) or through one more LDS line (x passes, whose r2c/c2r packing pairs
// X[m] with X[N-m]).  One LDS exchange per line replaces the staging, the
// Stockham passes and the output round trip of fft_lines.  8 lines per
// workgroup of 8 max(N1, N2) threads (the threads of the smaller factor idle
// through the other step); Y rows of N2 + 1 elements and lines of
// N1 (N2 + 1) + 1, so that step A's stores (consecutive n2, or consecutive
// lines) and step B's loads (consecutive k1, or consecutive lines) fall in
// different banks.  Each step orders its lanes the way its side of HBM is
// laid out: along the line where the line is contiguous, across lines where
// the pass reads or writes the transposed layout.
constexpr int kTlLines = 8;

template <int N1, int N2>
struct TlShape {
    static constexpr int N = N1 * N2;
    static constexpr int NT = kTlLines * (N1 > N2 ? N1 : N2);  // threads
    static constexpr int ROW = N2 + 1;
    static constexpr int LINE = N1 * ROW + 1;
    static constexpr int LDS = kTlLines * LINE + N;            // Y (or the x passes' lines) + twiddles
    static_assert(LINE >= N + 1, "the x passes' lines of N + 1 fit in a Y line");
};

// cos and sin of 2 pi e / 48 for e = 0..12 (a quarter turn; the W_R of every
// supported factor R divides 48)
constexpr double kW48c[13] = {1.0, 0.991444861373810411144558, 0.965925826289068286749743,
                              0.923879532511286756128183, 0.866025403784438646763723, 0.793353340291235164579777,
                              0.707106781186547524400844, 0.608761429008720639416098, 0.5,
                              0.382683432365089771728460, 0.258819045102520762348899, 0.130526192220051591548406,
                              0.0};

// a * W_R^e, W_R = exp(SIGN 2 pi i / R); R divides 48; e a constant once the
// caller's loops are unrolled, so the table reads and branches fold away
template <class T, int SIGN, int R>
__device__ __forceinline__ C2<T> twr_mul(C2<T> a, int e)
{
    const int e48 = (e % R) * (48 / R);
    const int q = e48 / 12, r = e48 % 12;
    if (e48 == 0) return a;
    if (r == 0) {  // a quarter, half or three-quarter turn
        if (q == 2) return C2<T>{-a.x, -a.y};
        const bool minus_i = (q == 1) == (SIGN < 0);  // W = -i: forward quarter, backward three-quarter
        return minus_i ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x};
    }
    // cos and sin of the angle q * 90 + r * 7.5 degrees
    const double c = q == 0 ? kW48c[r] : (q == 1 ? -kW48c[12 - r] : (q == 2 ? -kW48c[r] : kW48c[12 - r]));
    const double s = q == 0 ? kW48c[12 - r] : (q == 1 ? kW48c[r] : (q == 2 ? -kW48c[12 - r] : -kW48c[r]));
    const T wr = (T)c, wi = (T)(SIGN * s);
    return C2<T>{a.x * wr - a.y * wi, a.x * wi + a.y * wr};
}

// R-point DFT in registers, natural order in and out: R = 2, 3, 4 directly,
// R = 8, 12, 16 as 4 x (R/4) Cooley-Tukey
template <class T, int SIGN, int R>
struct DftReg {
    static __device__ __forceinline__ void run(C2<T>* v)
    {
        static_assert(R % 4 == 0 && R > 4 && R <= 16, "supported: 2, 3, 4, 8, 12, 16");
        constexpr int P = 4, Q = R / 4;  // n = Q n1 + n2, k = k1 + P k2
        C2<T> a[Q][P];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) {
            C2<T> t[P];
#pragma unroll
            for (int n1 = 0; n1 < P; ++n1) t[n1] = v[Q * n1 + n2];
            dft4<T, SIGN>(t);
#pragma unroll
            for (int k1 = 0; k1 < P; ++k1) a[n2][k1] = twr_mul<T, SIGN, R>(t[k1], n2 * k1);
        }
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            C2<T> t[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; ++n2) t[n2] = a[n2][k1];
            DftReg<T, SIGN, Q>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
        }
    }
};
template <class T, int SIGN>
struct DftReg<T, SIGN, 2> {
    static __device__ __forceinline__ void run(C2<T>* v) { dft2<T, SIGN>(v); }
};
template <class T, int SIGN>
struct DftReg<T, SIGN, 3> {
    static __device__ __forceinline__ void run(C2<T>* v) { dft3<T, SIGN>(v); }
};
template <class T, int SIGN>
struct DftReg<T, SIGN, 4> {
    static __device__ __forceinline__ void run(C2<T>* v) { dft4<T, SIGN>(v); }
};

// v * w (SIGN = -1) or v * conj(w) (SIGN = +1)
template <class T, int SIGN>
__device__ __forceinline__ C2<T> tw_apply(C2<T> v, C2<T> w)
{
    if (SIGN > 0) w.y = -w.y;
    return cmul(v, w);
}

// The pass's twiddle table W_N^(n2 k1) (entry N2 k1 + n2): fetched into
// registers before the line loads are issued (clamped, unconditional loads), so
// that storing it into LDS waits for these loads only and leaves the line loads
// in flight
template <class T, int N1, int N2>
struct TlTwiddles {
    using S = TlShape<N1, N2>;
    static constexpr int U = (S::N + S::NT - 1) / S::NT;
    C2<T> w[U];
    __device__ __forceinline__ void fetch(const C2<T>* __restrict__ tw)
    {
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = tw[min((int)threadIdx.x + u * S::NT, S::N - 1)];
    }
    __device__ __forceinline__ void put(C2<T>* twl) const
    {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = (int)threadIdx.x + u * S::NT;
            if (S::N % S::NT == 0 || t < S::N) twl[t] = w[u];
        }
    }
};

// Step A's tail: twiddle W_N^(n2 k1) and store Y[k1][n2] of line l
template <class T, int SIGN, int N1, int N2>
__device__ __forceinline__ void tl_store_y(C2<T>* Y, const C2<T>* twl, int l, int n2, const C2<T> (&v)[N1])
{
    using S = TlShape<N1, N2>;
    C2<T>* y = Y + l * S::LINE + n2;
    y[0] = v[0];
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) y[k1 * S::ROW] = tw_apply<T, SIGN>(v[k1], twl[k1 * N2 + n2]);
}

// K2, two-level: A[kx][j][:] -> B[kx][kz][j]
template <class T, int N1, int N2>
__global__ void __launch_bounds__((TlShape<N1, N2>::NT))
    k_zfwd_tl(int ny, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ A, C2<T>* __restrict__ B)
{
    using S = TlShape<N1, N2>;
    constexpr int N = S::N;
    __shared__ C2<T> Y[S::LDS];
    C2<T>* twl = Y + kTlLines * S::LINE;
    const unsigned id = xcd_block(), gj = (unsigned)(ny / kTlLines);
    const int m = (int)(id / gj);
    const int j0 = (int)(id % gj) * kTlLines;
    const int tid = threadIdx.x;
    __builtin_assume(tid < S::NT);  // the launch's block size
    TlTwiddles<T, N1, N2> tws;
    tws.fetch(tw);
    {
        const int l = tid / N2, n2 = tid - l * N2;  // lanes along the line
        const bool act = tid < kTlLines * N2;
        // unconditional loads (idle threads read line 0), so no branch waits for them
        const C2<T>* src = A + ((size_t)m * ny + j0 + (act ? l : 0)) * N + n2;
        C2<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = src[N2 * n1];
        tws.put(twl);
        __syncthreads();
        if (act) {
            DftReg<T, -1, N1>::run(v);
            tl_store_y<T, -1, N1, N2>(Y, twl, l, n2, v);
        }
    }
    __syncthreads();
    if (tid < kTlLines * N1) {
        const int l = tid & (kTlLines - 1), k1 = tid / kTlLines;  // lanes across lines
        const C2<T>* y = Y + l * S::LINE + k1 * S::ROW;
        C2<T> v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = y[n2];
        DftReg<T, -1, N2>::run(v);
        C2<T>* dst = B + ((size_t)m * N + k1) * ny + j0 + l;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) dst[(size_t)N1 * k2 * ny] = v[k2];
    }
}

// K4, two-level: B[kx][kz][j] -> A[kx][j][:]
template <class T, int N1, int N2>
__global__ void __launch_bounds__((TlShape<N1, N2>::NT))
    k_zinv_tl(int ny, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ B, C2<T>* __restrict__ A)
{
    using S = TlShape<N1, N2>;
    constexpr int N = S::N;
    __shared__ C2<T> Y[S::LDS];
    C2<T>* twl = Y + kTlLines * S::LINE;
    const unsigned id = xcd_block(), gj = (unsigned)(ny / kTlLines);
    const int m = (int)(id / gj);
    const int j0 = (int)(id % gj) * kTlLines;
    const int tid = threadIdx.x;
    __builtin_assume(tid < S::NT);  // the launch's block size
    TlTwiddles<T, N1, N2> tws;
    tws.fetch(tw);
    {
        const int l = tid & (kTlLines - 1), n2 = tid / kTlLines;  // lanes across lines
        const bool act = tid < kTlLines * N2;
        const C2<T>* src = B + ((size_t)m * N + (act ? n2 : 0)) * ny + j0 + l;
        C2<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = src[(size_t)N2 * n1 * ny];
        tws.put(twl);
        __syncthreads();
        if (act) {
            DftReg<T, +1, N1>::run(v);
            tl_store_y<T, +1, N1, N2>(Y, twl, l, n2, v);
        }
    }
    __syncthreads();
    if (tid < kTlLines * N1) {
        const int l = tid / N1, k1 = tid - l * N1;  // lanes along the line
        const C2<T>* y = Y + l * S::LINE + k1 * S::ROW;
        C2<T> v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) v[n2] = y[n2];
        DftReg<T, +1, N2>::run(v);
        C2<T>* dst = A + ((size_t)m * ny + j0 + l) * N + k1;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) dst[N1 * k2] = v[k2];
    }
}

// K1, two-level: s(:, j, k0..k0+15) -> A[kx][j][k], kx = 0..N/2 (line l packs
// the real lines k0+2l and k0+2l+1 into one complex line, as k_xfwd)
template <class T, class TIN, int N1, int N2>
__global__ void __launch_bounds__((TlShape<N1, N2>::NT))
    k_xfwd_tl(int ny, int nz, const C2<T>* __restrict__ tw, const TIN* __restrict__ s, C2<T>* __restrict__ A)
{
    using S = TlShape<N1, N2>;
    constexpr int N = S::N, KT = 2 * kTlLines, LS = N + 1, nh = N / 2 + 1;
    __shared__ C2<T> Y[S::LDS];
    C2<T>* twl = Y + kTlLines * S::LINE;
    const int j = blockIdx.y, k0 = blockIdx.x * KT;
    const int tid = threadIdx.x;
    __builtin_assume(tid < S::NT);  // the launch's block size
    TlTwiddles<T, N1, N2> tws;
    tws.fetch(tw);
    {
        const int l = tid / N2, n2 = tid - l * N2;  // lanes along the line
        const bool act = tid < kTlLines * N2;
        const size_t plane = (size_t)N * ny;
        const TIN* s0 = s + n2 + (size_t)N * (j + (size_t)ny * (k0 + 2 * (act ? l : 0)));
        C2<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = C2<T>{(T)s0[N2 * n1], (T)s0[N2 * n1 + plane]};
        tws.put(twl);
        __syncthreads();
        if (act) {
            DftReg<T, -1, N1>::run(v);
            tl_store_y<T, -1, N1, N2>(Y, twl, l, n2, v);
        }
    }
    __syncthreads();
    {
        const int l = tid / N1, k1 = tid - l * N1;
        const bool act = tid < kTlLines * N1;
        C2<T> v[N2];
        if (act) {
            const C2<T>* y = Y + l * S::LINE + k1 * S::ROW;
#pragma unroll
            for (int n2 = 0; n2 < N2; ++n2) v[n2] = y[n2];
            DftReg<T, -1, N2>::run(v);
        }
        __syncthreads();  // every Y row read before the lines overwrite them
        if (act) {
            C2<T>* x = Y + l * LS + k1;
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) x[N1 * k2] = v[k2];
        }
    }
    __syncthreads();
    for (int t = tid; t < nh * KT; t += S::NT) {
        const int m = t / KT, kk = t - m * KT;
        const int l = kk >> 1;
        const C2<T> Z = Y[l * LS + m];
        const C2<T> Zn = Y[l * LS + (m == 0 ? 0 : N - m)];
        C2<T> val;
        if ((kk & 1) == 0)
            val = C2<T>{(T)0.5 * (Z.x + Zn.x), (T)0.5 * (Z.y - Zn.y)};  // (Z + conj Zn)/2
        else
            val = C2<T>{(T)0.5 * (Z.y + Zn.y), (T)0.5 * (Zn.x - Z.x)};  // (Z - conj Zn)/(2i)
        A[((size_t)m * ny + j) * nz + k0 + kk] = val;
    }
}

// K5, two-level: A[kx][j][k0..k0+15] -> p(:, j, k) scaled by 1/(nx nz);
// mode 0: p = val; mode 1: p = p + val - shift[0] (as k_xinv)
template <class T, int N1, int N2>
__global__ void __launch_bounds__((TlShape<N1, N2>::NT))
    k_xinv_tl(int ny, int nz, const C2<T>* __restrict__ tw, const C2<T>* __restrict__ A, double* __restrict__ p,
              int mode, const double* __restrict__ shift)
{
    using S = TlShape<N1, N2>;
    constexpr int N = S::N, L = kTlLines, KT = 2 * L, LS = N + 1, nh = N / 2 + 1;
    constexpr int kB = stage_batch<T>() / 2;  // two complex loads per element
    __shared__ C2<T> Y[S::LDS];
    C2<T>* twl = Y + kTlLines * S::LINE;
    const int j = blockIdx.y, k0 = blockIdx.x * KT;
    const int tid = threadIdx.x;
    __builtin_assume(tid < S::NT);  // the launch's block size
    TlTwiddles<T, N1, N2> tws;
    tws.fetch(tw);
    tws.put(twl);  // here: the compiler would otherwise issue the fetch after the staging loop
    // Z_l[m] = X1[m] + i X2[m] with Hermitian extension, into the lines (as k_xinv)
    for (int t0 = tid; t0 < nh * L; t0 += kB * S::NT) {
        T x1r[kB], x1i[kB], x2r[kB], x2i[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = min(t0 + u * S::NT, nh * L - 1);
            const int m = t / L, l = t - m * L;
            const size_t o = ((size_t)m * ny + j) * nz + k0 + 2 * l;
            const C2<T> a = A[o], b = A[o + 1];
            x1r[u] = a.x;
            x1i[u] = a.y;
            x2r[u] = b.x;
            x2i[u] = b.y;
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int t = t0 + u * S::NT;
            if (t < nh * L) {
                const int m = t / L, l = t - m * L;
                const C2<T> x1{x1r[u], x1i[u]}, x2{x2r[u], x2i[u]};
                if (m == 0 || 2 * m == N) {
                    Y[l * LS + m] = C2<T>{x1.x, x2.x};
                } else {
                    Y[l * LS + m] = C2<T>{x1.x - x2.y, x1.y + x2.x};
                    Y[l * LS + N - m] = C2<T>{x1.x + x2.y, x2.x - x1.y};
                }
            }
        }
    }
    __syncthreads();
    {
        const int l = tid / N2, n2 = tid - l * N2;
        const bool act = tid < kTlLines * N2;
        C2<T> v[N1];
        if (act) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) v[n1] = Y[l * LS + N2 * n1 + n2];
            DftReg<T, +1, N1>::run(v);
        }
        __syncthreads();  // every line read before Y overwrites them
        if (act) tl_store_y<T, +1, N1, N2>(Y, twl, l, n2, v);
    }
    __syncthreads();
    if (tid >= kTlLines * N1) return;
    const int l = tid / N1, k1 = tid - l * N1;  // lanes along the line
    const C2<T>* y = Y + l * S::LINE + k1 * S::ROW;
    C2<T> v[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) v[n2] = y[n2];
    DftReg<T, +1, N2>::run(v);
    const double scale = 1.0 / ((double)N * (double)nz);
    const size_t plane = (size_t)N * ny;
    double* p0 = p + k1 + (size_t)N * (j + (size_t)ny * (k0 + 2 * l));
    if (mode == 0) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            p0[N1 * k2] = (double)v[k2].x * scale;
            p0[N1 * k2 + plane] = (double)v[k2].y * scale;
        }
        return;
    }
    const double sh = shift[0];
    double a0[N2], a1[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
        a0[k2] = p0[N1 * k2];
        a1[k2] = p0[N1 * k2 + plane];
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
        p0[N1 * k2] = a0[k2] + (double)v[k2].x * scale - sh;
        p0[N1 * k2 + plane] = a1[k2] + (double)v[k2].y * scale - sh;
    }
}

// (N1, N2) of the two-level passes for a line length, or {0, 0} for the
// Stockham passes; N1 (the step-A DFT, held in registers per thread) takes the
// larger factor
struct TlSplit {
    int n1, n2;
};
static TlSplit tl_split(int n)
{
    switch (n) {
        case 32: return {8, 4};
        case 64: return {8, 8};
        case 96: return {12, 8};
        case 128: return {16, 8};
        case 192: return {16, 12};
        case 256: return {16, 16};
        default: return {0, 0};
    }
}

// f(integral_constant<N1>, integral_constant<N2>) for a length tl_split accepts
template <class F>
static void tl_dispatch(int n, F&& f)
{
    using std::integral_constant;
    switch (n) {
        case 32: f(integral_constant<int, 8>{}, integral_constant<int, 4>{}); break;
        case 64: f(integral_constant<int, 8>{}, integral_constant<int, 8>{}); break;
        case 96: f(integral_constant<int, 12>{}, integral_constant<int, 8>{}); break;
        case 128: f(integral_constant<int, 16>{}, integral_constant<int, 8>{}); break;
        case 192: f(integral_constant<int, 16>{}, integral_constant<int, 12>{}); break;
        case 256: f(integral_constant<int, 16>{}, integral_constant<int, 16>{}); break;
        default: break;
    }
}

// ------------------------------------------------------------ K3 y-solve

template <class T>
__device__ __forceinline__ T shfl_t(T v, int src) { return __shfl(v, src, 64); }

// The value of lane - 1 (lane 0 gets 0) or lane + 1 (lane 63 gets 0) of the
// wavefront by a DPP wave shift, a VALU move; a shuffle goes through the LDS
// crossbar, which the y solve's exchanges keep busy
__device__ __forceinline__ int dpp_prev(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_next(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true); }

// The values of lanes lane - S and lane + S of a 32-bit register; lanes whose
// neighbour lies outside the wavefront get an unspecified value.  Stride 1 by
// DPP wave shifts, the others by shuffles (permlane16/32 swaps at strides 16
// and 32, and two wave shifts at stride 2, measured slower than the shuffles).
template <int S>
__device__ __forceinline__ void nbr32(int u, int lane, int& prev, int& next)
{
    if constexpr (S == 1) {
        prev = dpp_prev(u);
        next = dpp_next(u);
    } else {
        prev = __shfl(u, lane - S, 64);
        next = __shfl(u, lane + S, 64);
    }
}
template <int S>
__device__ __forceinline__ void lane_nbrs(float v, int lane, float& prev, float& next)
{
    int p, n;
    nbr32<S>(__builtin_bit_cast(int, v), lane, p, n);
    prev = __builtin_bit_cast(float, p);
    next = __builtin_bit_cast(float, n);
}
template <int S>
__device__ __forceinline__ void lane_nbrs(double v, int lane, double& prev, double& next)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    int plo, nlo, phi, nhi;
    nbr32<S>((int)(unsigned)(b & 0xffffffffull), lane, plo, nlo);
    nbr32<S>((int)(unsigned)(b >> 32), lane, phi, nhi);
    prev = __builtin_bit_cast(double, ((unsigned long long)(unsigned)phi << 32) | (unsigned)plo);
    next = __builtin_bit_cast(double, ((unsigned long long)(unsigned)nhi << 32) | (unsigned)nlo);
}

// 1/x by the hardware reciprocal refined with Newton steps (two in FP64, one
// in FP32: about one ulp), without the scaling and fix-up of a correctly
// rounded division: the y systems are strictly diagonally dominant (or, for
// the pinned column, have a unit diagonal after scaling), so x is never zero,
// denormal or infinite
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float rcp_nr(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// One parallel-cyclic-reduction step at stride S over the 64 lanes' rows
// A x_{i-S} + x_i + C x_{i+S} = D (unit diagonal kept, so no B is exchanged).
template <int S, class TC>
__device__ __forceinline__ void pcr_step(TC& A, TC& C, C2<TC>& D, int lane)
{
    TC Am, Cm, Ap, Cp;
    C2<TC> Dm, Dp;
    lane_nbrs<S>(A, lane, Am, Ap);
    lane_nbrs<S>(C, lane, Cm, Cp);
    lane_nbrs<S>(D.x, lane, Dm.x, Dp.x);
    lane_nbrs<S>(D.y, lane, Dm.y, Dp.y);
    // Rows beyond the wavefront need no zeroing: at the start of the step at
    // stride S, A = 0 exactly on lanes below S (lane 0's row has no lower
    // coupling, the wall's a- being 0, and A' = -A(i-S) A inv keeps the band
    // of zeros doubling) and C = 0 on lanes from 64 - S (likewise from lane
    // 63), while the values fetched from outside are finite (another lane's
    // row, or 0): every term they enter is multiplied by that exact zero.
    const TC inv = rcp_nr((TC)1 - Cm * A - Ap * C);
    D = C2<TC>{(D.x - Dm.x * A - Dp.x * C) * inv, (D.y - Dm.y * A - Dp.y * C) * inv};
    const TC A2 = -Am * A * inv, C2v = -Cp * C * inv;
    A = A2;
    C = C2v;
}

// Solves am_j x_{j-1} + (lam - am_j - ap_j) x_j + ap_j x_{j+1} = d_j for one
// column with one wavefront.  Lane p owns rows p*m..p*m+m-1 (m <= MR).  Each
// lane eliminates its block to an affine map of the last unknowns of its own
// and the previous block; the resulting 64-unknown tridiagonal system is solved
// by parallel cyclic reduction with shuffles, then the block interiors are
// recovered.  Systems with lam != 0 are strictly diagonally dominant, so no
// pivoting is needed (LAPACK zgttrf at fastPoissonSolver.f90:316 never pivots
// them).  pinned: the singular (0,0) column, solved with x_{ny-1} = 0 and then
// made volume-weighted mean free (the reference patches U(n,n)=eps instead,
// :318-322, and removes the mean afterwards, :388-423).
// EXACT: ny = 64 MR, so every lane holds MR rows and the row guards fold away.
template <class TC, int MR, class TS, bool EXACT>
__device__ void tri_wave(C2<TS>* __restrict__ col, int ny, const TC* __restrict__ am, const TC* __restrict__ ap,
                         TC lam, bool pinned, const TC* __restrict__ dyf, TC inv_ly)
{
    const int lane = threadIdx.x & 63;
    const int m = EXACT ? MR : (ny + 63) / 64;
    TC a[MR], b[MR], c[MR];
    C2<TC> d[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        const int row = lane * m + r;
        if (r < m && (EXACT || row < ny) && !(pinned && row == ny - 1)) {
            a[r] = am[row];
            c[r] = ap[row];
            b[r] = lam - a[r] - c[r];
            const C2<TS> v = col[row];
            d[r] = C2<TC>{(TC)v.x, (TC)v.y};
        } else {
            a[r] = 0;
            c[r] = 0;
            b[r] = 1;
            d[r] = C2<TC>{0, 0};
        }
    }
    // local elimination of rows 0..m-2 (T' with external couplings moved right)
    C2<TC> al[MR];
    TC be[MR], ga[MR], cp[MR];
    C2<TC> fa = {0, 0}, la = {0, 0};
    TC fb = 0, fg = 1, lb = 1, lg = 0;  // first-row and previous-row maps for m == 1
    if (m >= 2) {
        const int mm = m - 1;
#pragma unroll
        for (int r = 0; r < MR - 1; ++r) {
            if (r < mm) {
                const TC cint = (r < mm - 1) ? c[r] : (TC)0;
                const TC rg = (r == mm - 1) ? -c[r] : (TC)0;
                const TC rb = (r == 0) ? -a[0] : (TC)0;
                // one division per row: its reciprocal scales the row's four results
                if (r == 0) {
                    const TC inv = rcp_nr(b[0]);
                    al[0] = C2<TC>{d[0].x * inv, d[0].y * inv};
                    be[0] = rb * inv;
                    ga[0] = rg * inv;
                    cp[0] = cint * inv;
                } else {
                    const TC inv = rcp_nr(b[r] - a[r] * cp[r - 1]);
                    al[r] = C2<TC>{(d[r].x - a[r] * al[r - 1].x) * inv, (d[r].y - a[r] * al[r - 1].y) * inv};
                    be[r] = (rb - a[r] * be[r - 1]) * inv;
                    ga[r] = (rg - a[r] * ga[r - 1]) * inv;
                    cp[r] = cint * inv;
                }
            }
        }
#pragma unroll
        for (int r = MR - 3; r >= 0; --r) {
            if (r < mm - 1) {
                al[r].x -= cp[r] * al[r + 1].x;
                al[r].y -= cp[r] * al[r + 1].y;
                be[r] -= cp[r] * be[r + 1];
                ga[r] -= cp[r] * ga[r + 1];
            }
        }
        fa = al[0];
        fb = be[0];
        fg = ga[0];
#pragma unroll
        for (int r = 0; r < MR - 1; ++r)
            if (r == mm - 1) {
                la = al[r];
                lb = be[r];
                lg = ga[r];
            }
    }
    // last row of the block
    TC ae = 0, bee = 1, ce = 0;
    C2<TC> de = {0, 0};
#pragma unroll
    for (int r = 0; r < MR; ++r)
        if (r == m - 1) {
            ae = a[r];
            bee = b[r];
            ce = c[r];
            de = d[r];
        }
    // first-row map of the next block (zeros on lane 63, from the wave shift:
    // no block above it; likewise ym below on lane 0)
    C2<TC> na, dummy;
    TC nb_, ng, dm;
    lane_nbrs<1>(fa.x, lane, dummy.x, na.x);
    lane_nbrs<1>(fa.y, lane, dummy.y, na.y);
    lane_nbrs<1>(fb, lane, dm, nb_);
    lane_nbrs<1>(fg, lane, dm, ng);
    // the reduced row, scaled to a unit diagonal
    const TC inv0 = rcp_nr(ae * lg + bee + ce * nb_);
    TC A = ae * lb * inv0;
    TC C = ce * ng * inv0;
    C2<TC> D = {(de.x - ae * la.x - ce * na.x) * inv0, (de.y - ae * la.y - ce * na.y) * inv0};
    // parallel cyclic reduction over the 64 lanes
    pcr_step<1>(A, C, D, lane);
    pcr_step<2>(A, C, D, lane);
    pcr_step<4>(A, C, D, lane);
    pcr_step<8>(A, C, D, lane);
    pcr_step<16>(A, C, D, lane);
    pcr_step<32>(A, C, D, lane);
    const C2<TC> y = D;
    C2<TC> ym;
    lane_nbrs<1>(y.x, lane, ym.x, dummy.x);
    lane_nbrs<1>(y.y, lane, ym.y, dummy.y);
    C2<TC> x[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        if (r < m - 1)
            x[r] = C2<TC>{al[r].x + be[r] * ym.x + ga[r] * y.x, al[r].y + be[r] * ym.y + ga[r] * y.y};
        else
            x[r] = y;
    }
    if (pinned) {
        // spectral volume-weighted mean: P00 -= sum_j dyf_j P00(j) / Ly
        TC sx = 0, sy = 0;
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            const int row = lane * m + r;
            if (r < m && (EXACT || row < ny)) {
                sx += dyf[row] * x[r].x;
                sy += dyf[row] * x[r].y;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            sx += __shfl_xor(sx, off, 64);
            sy += __shfl_xor(sy, off, 64);
        }
        sx *= inv_ly;
        sy *= inv_ly;
#pragma unroll
        for (int r = 0; r < MR; ++r) {
            x[r].x -= sx;
            x[r].y -= sy;
        }
    }
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        const int row = lane * m + r;
        if (r < m && (EXACT || row < ny)) col[row] = C2<TS>{(TS)x[r].x, (TS)x[r].y};
    }
}

struct TriTables {
    const double* lx;   // nh
    const double* lz;   // nz
    const double* am;   // ny
    const double* ap;
    const double* dyf;
    const float* amf;
    const float* apf;
    const float* dyff;
    double inv_ly;
};

template <class T, int MR, bool EXACT>
__global__ void __launch_bounds__(256) k_ysolve(int nh, int ny, int nz, TriTables tt, C2<T>* __restrict__ B)
{
    const int sys = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (sys >= nh * nz) return;
    const int m = sys / nz, kz = sys - m * nz;  // B is [kx][kz][j]
    C2<T>* col = B + (size_t)sys * ny;
    const double lam = tt.lx[m] + tt.lz[kz];
    const bool pinned = (m == 0 && kz == 0);
    if (sizeof(T) == 8 || pinned) {
        tri_wave<double, MR, T, EXACT>(col, ny, tt.am, tt.ap, lam, pinned, tt.dyf, tt.inv_ly);
    } else {
        tri_wave<float, MR, T, EXACT>(col, ny, tt.amf, tt.apf, (float)lam, false, tt.dyff, (float)tt.inv_ly);
    }
}

// ----------------------------------------------------- refinement residual

struct OpTables {
    int nx, ny, nz;
    double idx2, idz2;
    const double* am;
    const double* ap;
    const double* vol_y;  // dyf_j * dx * dz
};

// r = s - L p with L the operator the FFT solver inverts, and sum V p, as a
// march along z: a thread owns the column (i, j) of a chunk of kc planes and
// keeps p of planes k-1, k, k+1 in registers, so p is read once per chunk
// (plus its two halo planes) instead of once per z neighbour; the in-plane
// neighbours come from the rows of the 64 x 4 tile (L1) and their one-row
// ring (L2).
__global__ void __launch_bounds__(256)
    k_fft_residual(OpTables o, const double* __restrict__ s, const double* __restrict__ p,
                         double* __restrict__ r, int kc, Reducer red)
{
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const int i = blockIdx.x * 64 + (int)threadIdx.x;
    const int j = blockIdx.y * 4 + (int)threadIdx.y;
    const int k0 = (int)blockIdx.z * kc, k1 = min(nz, k0 + kc);
    double v[1] = {0.0};
    if (i < nx && j < ny && k0 < nz) {
        const size_t plane = (size_t)nx * ny;
        const size_t col = i + (size_t)nx * j;
        const int ie = i + 1 < nx ? i + 1 : 0, iw = i > 0 ? i - 1 : nx - 1;
        const double amj = o.am[j], apj = o.ap[j], vy = o.vol_y[j];
        double pb = p[col + plane * (k0 > 0 ? k0 - 1 : nz - 1)];
        double pc = p[col + plane * k0];
#pragma unroll 4
        for (int k = k0; k < k1; ++k) {
            const size_t c = col + plane * k;
            const double pf = p[col + plane * (k + 1 < nz ? k + 1 : 0)];
            const double pe = p[c - i + ie];
            const double pw = p[c - i + iw];
            const double pn = (j + 1 < ny) ? p[c + nx] : pc;
            const double ps = (j > 0) ? p[c - nx] : pc;
            const double lp = (pe - 2.0 * pc + pw) * o.idx2 + (pf - 2.0 * pc + pb) * o.idz2 + amj * (ps - pc) +
                              apj * (pn - pc);
            r[c] = s[c] - lp;
            v[0] += pc * vy;
            pb = pc;
            pc = pf;
        }
    }
    block_partial<1>(v, red);
}

__global__ void k_fin_mean(unsigned nblocks, double inv_vg, double* __restrict__ out_mean, Reducer red)
{
    double tot[1];
    final_sum<1>(red, nblocks, tot);
    if (threadIdx.x == 0) out_mean[0] = tot[0] * inv_vg;
}

__global__ void k_fft_operator(OpTables o, const double* __restrict__ p, double* __restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= o.nx) return;
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
    const double pc = p[c];
    const double pe = p[c - i + (i + 1 < nx ? i + 1 : 0)];
    const double pw = p[c - i + (i > 0 ? i - 1 : nx - 1)];
    const double pf = p[i + (size_t)nx * (j + (size_t)ny * (k + 1 < nz ? k + 1 : 0))];
    const double pb = p[i + (size_t)nx * (j + (size_t)ny * (k > 0 ? k - 1 : nz - 1))];
    const double pn = (j + 1 < ny) ? p[c + nx] : pc;
    const double ps = (j > 0) ? p[c - nx] : pc;
    q[c] = (pe - 2.0 * pc + pw) * o.idx2 + (pf - 2.0 * pc + pb) * o.idz2 + o.am[j] * (ps - pc) + o.ap[j] * (pn - pc);
}

}  // namespace tbfg

// ------------------------------------------------------------------ handle

using namespace tbfg;

struct tbfg_fps {
    tbfg_grid_desc g{};
    int precision = 0, refine = 0;
    int nx = 0, ny = 0, nz = 0, nh = 0;
    int KT = 0, JT = 0, MR = 1;
    bool tlx = false, tlz = false;  // two-level passes along x / z (tl_split)
    FftPlan px{}, pz{};
    DevBuf A, B;               // spectral work arrays (complex, T)
    DevBuf tab;                // lx, lz, am, ap, dyf, vol_y (double)
    DevBuf tabf;               // am, ap, dyf (float)
    DevBuf twx, twz, twxf, twzf;
    DevBuf r, partials, counter, mean;
    TriTables tt{};
    OpTables ot{};
    double inv_vg = 0;
    int64_t footprint = 0;
};

namespace tbfg {

// launch shape of the per-row FP64 kernels (refinement residual, operator)
static unsigned row_threads(int nx) { return nx >= 256 ? 256u : 64u; }
// grid and planes per chunk of k_fft_residual: 64 x 4 columns per block,
// chunks of planes for about 4096 blocks
static dim3 residual_grid(int nx, int ny, int nz, int& kc)
{
    const unsigned gx = ceil_div(nx, 64), gy = ceil_div(ny, 4);
    // (1024, 2048, 8192 and 16384 blocks measured 2-7 % slower at 256x192x256)
    const int chunks = std::max(1, std::min(nz, (int)(4096 / (gx * gy))));
    kc = (nz + chunks - 1) / chunks;
    return dim3(gx, gy, (unsigned)((nz + kc - 1) / kc));
}
static unsigned row_blocks(int nx) { return ceil_div(nx, row_threads(nx)); }

static int choose_lines(int n, int count, int pairs, int bfly, const FftPlan& plan)
{
    // lines longer than 1024 points: one line (pair) per workgroup, its LDS
    // beyond the default 64 KB (set_stockham_lds), if the plan's largest pass
    // (n over its smallest radix) fits the workgroup's butterflies
    if (n > 1024) {
        int nb = 0;
        for (int st = 0; st < plan.nrad; ++st) nb = std::max(nb, n / plan.rad[st]);
        if (2 * n + 1 > kLdsElemsMax || nb > bfly * kFftThreads) return 0;
        return count % (pairs ? 2 : 1) == 0 ? 1 : 0;
    }
    // largest L with L*(n+1)+n <= kLdsElems (padded lines plus twiddles) and L*n/2 <=
    // bfly * threads (the butterflies of a radix-2 pass), (pairs ? 2L : L)
    // dividing count, L <= 64 (preferring whole 128-byte lines in the z passes'
    // transposed runs, L = 8 instead of 12 at ny = 192, measured slower)
    int best = 0;
    int lmax = 64;
    if (const char* e = std::getenv("TBFG_FFT_LINES")) lmax = std::max(1, std::atoi(e));  // A/B switch
    for (int L = 1; L <= lmax; ++L) {
        if (L * (n + 1) + n > kLdsElems || L * n / 2 > bfly * kFftThreads) break;
        const int span = pairs ? 2 * L : L;
        if (count % span == 0) best = L;
    }
    return best;
}

// Dynamic LDS of the Stockham passes (padded lines plus twiddles) in bytes
template <class T>
static size_t stockham_lds_x(const tbfg_fps* h) { return ((size_t)(h->KT / 2) * (h->nx + 1) + h->nx) * sizeof(C2<T>); }
template <class T>
static size_t stockham_lds_z(const tbfg_fps* h) { return ((size_t)h->JT * (h->nz + 1) + h->nz) * sizeof(C2<T>); }

// Lines longer than 1024 points take more than the default 64 KB of dynamic
// LDS: raise the kernels' limit (gfx950 gives a workgroup up to 160 KB)
template <class T, class TIN>
static int set_stockham_lds(const tbfg_fps* h)
{
    const size_t lx = stockham_lds_x<T>(h), lz = stockham_lds_z<T>(h);
    auto raise = [](const void* f, size_t bytes) -> int {
        if (bytes <= 65536) return TBFG_OK;
        TBFG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        return TBFG_OK;
    };
    if (!h->tlx) {
        TBFG_TRY(raise(reinterpret_cast<const void*>(&k_xfwd<T, TIN>), lx));
        TBFG_TRY(raise(reinterpret_cast<const void*>(&k_xinv<T>), lx));
    }
    if (!h->tlz) {
        TBFG_TRY(raise(reinterpret_cast<const void*>(&k_zfwd<T>), lz));
        TBFG_TRY(raise(reinterpret_cast<const void*>(&k_zinv<T>), lz));
    }
    return TBFG_OK;
}

int fps_create(const tbfg_grid_desc* g, const tbfg_fps_opts* o, tbfg_fps** out)
{
    if (!out) {
        set_error("tbfg_fps_create: NULL output");
        return TBFG_EINVAL;
    }
    *out = nullptr;
    TBFG_TRY(validate_grid(g));
    std::unique_ptr<tbfg_fps> h(new tbfg_fps());
    h->g = *g;
    h->precision = o ? o->precision : TBFG_PREC_FP64;
    h->refine = o ? o->refine_sweeps : 2;
    if (h->precision != TBFG_PREC_FP64 && h->precision != TBFG_PREC_MIXED) {
        set_error("precision = %d", h->precision);
        return TBFG_EINVAL;
    }
    if (h->precision == TBFG_PREC_MIXED && (h->refine < 0 || h->refine > 8)) {
        set_error("refine_sweeps = %d (expected 0..8)", h->refine);
        return TBFG_EINVAL;
    }
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    if (nx % 2 || nz % 2 || nx > 4096 || nz > 4096 || ny > 1024 || nx < 2 || nz < 2) {
        set_error("FFT solver needs even nx, nz <= 4096 and ny <= 1024 (got %d x %d x %d)", nx, ny, nz);
        return TBFG_EINVAL;
    }
    if (make_plan(nx, h->px) || make_plan(nz, h->pz)) {
        set_error("FFT lengths must be 2^a 3^b (got nx=%d nz=%d)", nx, nz);
        return TBFG_EINVAL;
    }
    h->nx = nx;
    h->ny = ny;
    h->nz = nz;
    h->nh = nx / 2 + 1;
    const int bfly = h->precision == TBFG_PREC_MIXED ? max_bfly<float>() : max_bfly<double>();
    h->KT = 2 * choose_lines(nx, nz, 1, bfly, h->px);
    h->JT = choose_lines(nz, ny, 0, bfly, h->pz);
    if (h->KT < 2 || h->JT < 1) {
        set_error("no line tiling for %d x %d x %d (a line longer than 1024 points needs every FFT pass "
                  "within %d butterflies)", nx, ny, nz, bfly * kFftThreads);
        return TBFG_EINVAL;
    }
    const int m = (ny + 63) / 64;
    h->MR = m <= 1 ? 1 : (m <= 2 ? 2 : (m <= 3 ? 3 : (m <= 4 ? 4 : (m <= 8 ? 8 : 16))));
    HostLevel L = make_level(g, 0);
    // fastPoissonSolver.f90:215-251 (eigenvalues), :255-276 (y coefficients)
    std::vector<double> t;
    const double dx = L.ax[0].c(2) - L.ax[0].c(1), dz = L.ax[2].c(2) - L.ax[2].c(1);
    const double pi = 3.141592653589793;
    size_t o_lx = t.size();
    for (int i = 1; i <= h->nh; ++i) t.push_back(2.0 * (std::cos((i - 1) * 2.0 * pi / nx) - 1.0) / (dx * dx));
    size_t o_lz = t.size();
    for (int i = 1; i <= nz; ++i) t.push_back(2.0 * (std::cos((i - 1) * 2.0 * pi / nz) - 1.0) / (dz * dz));
    size_t o_am = t.size();
    for (int j = 1; j <= ny; ++j) t.push_back(j == 1 ? 0.0 : 1.0 / (L.ax[1].dc(j) * L.ax[1].df(j)));
    size_t o_ap = t.size();
    for (int j = 1; j <= ny; ++j) t.push_back(j == ny ? 0.0 : 1.0 / (L.ax[1].dc(j + 1) * L.ax[1].df(j)));
    size_t o_dyf = t.size();
    double ly = 0.0;
    for (int j = 1; j <= ny; ++j) {
        t.push_back(L.ax[1].df(j));
        ly += L.ax[1].df(j);
    }
    size_t o_vy = t.size();
    for (int j = 1; j <= ny; ++j) t.push_back(L.ax[1].df(j) * L.ax[0].df(1) * L.ax[2].df(1));
    TBFG_TRY(upload(h->tab, t.data(), t.size() * 8));
    std::vector<float> tf;
    for (size_t q = o_am; q < o_vy; ++q) tf.push_back((float)t[q]);
    TBFG_TRY(upload(h->tabf, tf.data(), tf.size() * 4));
    const double* tb = h->tab.as<double>();
    const float* tbf = h->tabf.as<float>();
    h->tt.lx = tb + o_lx;
    h->tt.lz = tb + o_lz;
    h->tt.am = tb + o_am;
    h->tt.ap = tb + o_ap;
    h->tt.dyf = tb + o_dyf;
    h->tt.amf = tbf;
    h->tt.apf = tbf + ny;
    h->tt.dyff = tbf + 2 * ny;
    h->tt.inv_ly = 1.0 / ly;
    h->ot.nx = nx;
    h->ot.ny = ny;
    h->ot.nz = nz;
    h->ot.idx2 = 1.0 / (dx * dx);
    h->ot.idz2 = 1.0 / (dz * dz);
    h->ot.am = h->tt.am;
    h->ot.ap = h->tt.ap;
    h->ot.vol_y = tb + o_vy;
    h->inv_vg = 1.0 / (ly * L.ax[0].df(1) * nx * L.ax[2].df(1) * nz);
    // two-level passes where the line length splits (tl_split) and the
    // workgroup's lines tile the other axis (TBFG_FFT_TWO_LEVEL=0 keeps
    // fft_lines: an A/B switch)
    const char* tl_env = std::getenv("TBFG_FFT_TWO_LEVEL");
    const bool tl_on = !(tl_env && tl_env[0] == '0');
    h->tlx = tl_on && tl_split(nx).n1 > 0 && nz % (2 * kTlLines) == 0;
    h->tlz = tl_on && tl_split(nz).n1 > 0 && ny % kTlLines == 0;
    TBFG_TRY(h->precision == TBFG_PREC_MIXED ? set_stockham_lds<float, double>(h.get())
                                             : set_stockham_lds<double, double>(h.get()));
    // twiddles in FP64 and rounded to FP32.  fft_lines: entry Ns-1 + (r-1)*Ns + k
    // of the pass with sub-transform length Ns and radix R is
    // exp(-2 pi i r k / (Ns R)) (entry n-1 unused); two-level passes (n = N1 N2):
    // entry N2 k1 + n2 is exp(-2 pi i n2 k1 / n)
    auto tw = [&](int n, bool two_level, DevBuf& d64, DevBuf& d32) -> int {
        FftPlan pl;
        make_plan(n, pl);
        std::vector<double> w(2 * (size_t)n, 0.0);
        std::vector<float> wf(2 * (size_t)n, 0.0f);
        if (two_level) {
            const TlSplit sp = tl_split(n);
            for (int k1 = 0; k1 < sp.n1; ++k1)
                for (int n2 = 0; n2 < sp.n2; ++n2) {
                    const size_t e = (size_t)sp.n2 * k1 + n2;
                    w[2 * e] = std::cos(2.0 * pi * n2 * k1 / n);
                    w[2 * e + 1] = -std::sin(2.0 * pi * n2 * k1 / n);
                }
        }
        int Ns = 1;
        for (int st = 0; st < pl.nrad && !two_level; ++st) {
            const int R = pl.rad[st];
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < Ns; ++k) {
                    const size_t e = (size_t)Ns - 1 + (size_t)(r - 1) * Ns + k;
                    const int q = r * k * (n / (Ns * R));  // exponent in units of 2 pi / n
                    w[2 * e] = std::cos(2.0 * pi * q / n);
                    w[2 * e + 1] = -std::sin(2.0 * pi * q / n);
                }
            Ns *= R;
        }
        for (size_t q = 0; q < w.size(); ++q) wf[q] = (float)w[q];
        TBFG_TRY(upload(d64, w.data(), w.size() * 8));
        TBFG_TRY(upload(d32, wf.data(), wf.size() * 4));
        return TBFG_OK;
    };
    TBFG_TRY(tw(nx, h->tlx, h->twx, h->twxf));
    TBFG_TRY(tw(nz, h->tlz, h->twz, h->twzf));
    const size_t es = h->precision == TBFG_PREC_MIXED ? 8 : 16;  // complex element size
    const size_t nspec = (size_t)h->nh * ny * nz;
    TBFG_TRY(h->A.alloc(nspec * es));
    TBFG_TRY(h->B.alloc(nspec * es));
    if (h->precision == TBFG_PREC_MIXED) {
        TBFG_TRY(h->r.alloc((size_t)nx * ny * nz * 8));
        int kc = 0;
        const dim3 rg = residual_grid(nx, ny, nz, kc);
        TBFG_TRY(h->partials.alloc((size_t)rg.x * rg.y * rg.z * 8 + 64));
        TBFG_TRY(h->counter.alloc(64));
        TBFG_TRY(memset_sync(h->counter.p, 0, 64));
        TBFG_TRY(h->mean.alloc(64));
    }
    h->footprint = (int64_t)(h->A.bytes + h->B.bytes + h->tab.bytes + h->tabf.bytes + h->twx.bytes + h->twz.bytes +
                             h->twxf.bytes + h->twzf.bytes + h->r.bytes + h->partials.bytes + h->counter.bytes +
                             h->mean.bytes);
    *out = h.release();
    return TBFG_OK;
}

template <class T, int MR>
static void launch_ysolve(tbfg_fps* h, C2<T>* B, hipStream_t s)
{
    const int nsys = h->nh * h->nz;
    if (h->ny == 64 * MR)
        hipLaunchKernelGGL((k_ysolve<T, MR, true>), dim3(ceil_div(nsys, 4)), dim3(256), 0, s, h->nh, h->ny, h->nz,
                           h->tt, B);
    else
        hipLaunchKernelGGL((k_ysolve<T, MR, false>), dim3(ceil_div(nsys, 4)), dim3(256), 0, s, h->nh, h->ny, h->nz,
                           h->tt, B);
}

// one spectral solve: in (double) -> p (double), mode 0 write / 1 accumulate
template <class T, class TIN>
static int spectral_solve(tbfg_fps* h, const TIN* in, double* p, int mode, hipStream_t s)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz, nh = h->nh;
    C2<T>* A = h->A.as<C2<T>>();
    C2<T>* B = h->B.as<C2<T>>();
    const bool f32 = sizeof(T) == 4;
    const C2<T>* twx = f32 ? h->twxf.as<C2<T>>() : h->twx.as<C2<T>>();
    const C2<T>* twz = f32 ? h->twzf.as<C2<T>>() : h->twz.as<C2<T>>();
    const size_t ldx = stockham_lds_x<T>(h);
    const size_t ldz = stockham_lds_z<T>(h);
    if (h->tlx)
        tl_dispatch(nx, [&](auto a, auto b) {
            constexpr int N1 = decltype(a)::value, N2 = decltype(b)::value;
            hipLaunchKernelGGL((k_xfwd_tl<T, TIN, N1, N2>), dim3(nz / (2 * kTlLines), ny), dim3(TlShape<N1, N2>::NT), 0,
                               s, ny, nz, twx, in, A);
        });
    else
        hipLaunchKernelGGL((k_xfwd<T, TIN>), dim3(nz / h->KT, ny), dim3(kFftThreads), ldx, s, nx, ny, nz, h->KT, h->px,
                           twx, in, A);
    if (h->tlz)
        tl_dispatch(nz, [&](auto a, auto b) {
            constexpr int N1 = decltype(a)::value, N2 = decltype(b)::value;
            hipLaunchKernelGGL((k_zfwd_tl<T, N1, N2>), dim3((ny / kTlLines) * nh), dim3(TlShape<N1, N2>::NT), 0, s, ny,
                               twz, A, B);
        });
    else
        hipLaunchKernelGGL((k_zfwd<T>), dim3((ny / h->JT) * nh), dim3(kFftThreads), ldz, s, ny, nz, h->JT, h->pz, twz,
                           A, B);
    switch (h->MR) {
        case 1: launch_ysolve<T, 1>(h, B, s); break;
        case 2: launch_ysolve<T, 2>(h, B, s); break;
        case 3: launch_ysolve<T, 3>(h, B, s); break;
        case 4: launch_ysolve<T, 4>(h, B, s); break;
        case 8: launch_ysolve<T, 8>(h, B, s); break;
        default: launch_ysolve<T, 16>(h, B, s); break;
    }
    if (h->tlz)
        tl_dispatch(nz, [&](auto a, auto b) {
            constexpr int N1 = decltype(a)::value, N2 = decltype(b)::value;
            hipLaunchKernelGGL((k_zinv_tl<T, N1, N2>), dim3((ny / kTlLines) * nh), dim3(TlShape<N1, N2>::NT), 0, s, ny,
                               twz, B, A);
        });
    else
        hipLaunchKernelGGL((k_zinv<T>), dim3((ny / h->JT) * nh), dim3(kFftThreads), ldz, s, ny, nz, h->JT, h->pz, twz,
                           B, A);
    if (h->tlx)
        tl_dispatch(nx, [&](auto a, auto b) {
            constexpr int N1 = decltype(a)::value, N2 = decltype(b)::value;
            hipLaunchKernelGGL((k_xinv_tl<T, N1, N2>), dim3(nz / (2 * kTlLines), ny), dim3(TlShape<N1, N2>::NT), 0, s,
                               ny, nz, twx, A, p, mode, h->mean.as<double>());
        });
    else
        hipLaunchKernelGGL((k_xinv<T>), dim3(nz / h->KT, ny), dim3(kFftThreads), ldx, s, nx, ny, nz, h->KT, h->px, twx,
                           A, p, mode, h->mean.as<double>());
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

int fps_solve_dev(tbfg_fps* h, const double* s, double* p, void* stream)
{
    if (!h || !s || !p) {
        set_error("tbfg_fps_solve_dev: NULL argument");
        return TBFG_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    if (h->precision == TBFG_PREC_FP64) return spectral_solve<double, double>(h, s, p, 0, st);
    TBFG_TRY((spectral_solve<float, double>(h, s, p, 0, st)));
    Reducer red{h->partials.as<double>()};
    for (int it = 0; it < h->refine; ++it) {
        int kc = 0;
        const dim3 grd = residual_grid(h->nx, h->ny, h->nz, kc);
        hipLaunchKernelGGL(k_fft_residual, grd, dim3(64, 4), 0, st, h->ot, s, p, h->r.as<double>(), kc, red);
        hipLaunchKernelGGL(k_fin_mean, dim3(1), dim3(1024), 0, st, grd.x * grd.y * grd.z, h->inv_vg,
                           h->mean.as<double>(), red);
        TBFG_TRY((spectral_solve<float, double>(h, h->r.as<double>(), p, 1, st)));
    }
    return TBFG_OK;
}

int fps_apply_operator_dev(tbfg_fps* h, const double* p, double* q, void* stream)
{
    const dim3 blk(row_threads(h->nx));
    hipLaunchKernelGGL(k_fft_operator, dim3(row_blocks(h->nx), h->ny, h->nz), blk, 0, (hipStream_t)stream, h->ot,
                       p, q);
    TBFG_HIP(hipGetLastError());
    return TBFG_OK;
}

void fps_destroy(tbfg_fps* h) { delete h; }

const tbfg_grid_desc* fps_grid(const tbfg_fps* h) { return &h->g; }

int64_t fps_footprint(const tbfg_fps* h) { return h ? h->footprint : 0; }

}  // namespace tbfg
