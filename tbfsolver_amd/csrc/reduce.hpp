// reduce.hpp -- deterministic block and grid reductions (device code).
#pragma once
#include "mg_kernels.hpp"

namespace tbfg {

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* >= 32*NV */)
{
    const int lane = threadIdx.x + blockDim.x * threadIdx.y;
    const int nthreads = blockDim.x * blockDim.y;
#pragma unroll
    for (int q = 0; q < NV; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
    const int wave = lane >> 6;
    const int nwaves = (nthreads + 63) >> 6;
    if ((lane & 63) == 0)
        for (int q = 0; q < NV; ++q) lds[wave * NV + q] = v[q];
    __syncthreads();
    if (lane == 0) {
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nwaves; ++w) s += lds[w * NV + q];
            v[q] = s;
        }
    }
}

// Deterministic grid reduction in two launches: every block of the producing
// kernel writes its NV partial sums (block_partial); a one-block finalize
// kernel adds them in block order (final_sum).  No device-scope fence is
// needed, which matters on a multi-XCD part where such a fence writes back L2.
template <int NV>
__device__ __forceinline__ void block_partial(double (&v)[NV], const Reducer& red)
{
    __shared__ double lds[32 * 8];
    const int lane = threadIdx.x + blockDim.x * threadIdx.y;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v, lds);
    if (lane == 0)
        for (int q = 0; q < NV; ++q) red.partials[(size_t)bid * NV + q] = v[q];
}

// one block; thread 0 receives the totals
template <int NV>
__device__ __forceinline__ void final_sum(const Reducer& red, unsigned nblocks, double (&tot)[NV])
{
    __shared__ double lds[32 * 8];
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int q = 0; q < NV; ++q) acc[q] += red.partials[(size_t)b * NV + q];
    block_sum<NV>(acc, lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < NV; ++q) tot[q] = acc[q];
}

}  // namespace tbfg
