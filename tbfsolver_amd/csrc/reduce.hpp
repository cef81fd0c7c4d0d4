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

// Writes the block's partials; returns true in every thread of the last block,
// whose thread 0 then holds the grand totals in `tot`.
template <int NV>
__device__ bool grid_sum(double (&v)[NV], const Reducer& red, double (&tot)[NV])
{
    __shared__ double lds[32 * 8];
    __shared__ int am_last;
    const int lane = threadIdx.x + blockDim.x * threadIdx.y;
    const int nthreads = blockDim.x * blockDim.y;
    const unsigned nblocks = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    block_sum<NV>(v, lds);
    if (lane == 0) {
        for (int q = 0; q < NV; ++q) red.partials[(size_t)bid * NV + q] = v[q];
        __threadfence();
        unsigned prev = atomicAdd(red.counter, 1u);
        am_last = (prev == nblocks - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (unsigned b = lane; b < nblocks; b += nthreads)
        for (int q = 0; q < NV; ++q) acc[q] += red.partials[(size_t)b * NV + q];
    __syncthreads();
    block_sum<NV>(acc, lds);
    if (lane == 0) {
        for (int q = 0; q < NV; ++q) tot[q] = acc[q];
        *red.counter = 0u;
    }
    return true;
}

}  // namespace tbfg
