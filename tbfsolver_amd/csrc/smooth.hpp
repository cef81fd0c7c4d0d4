// smooth.hpp -- device arithmetic shared by the Gauss-Seidel kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace tbfg {

// the smoother's division p = num / den (rbgs.f90:223-228): correctly rounded
// in FP64; in FP32 (the multigrid of mixed precision, a preconditioner) the
// hardware reciprocal (1 ulp) times num
template <class T>
__device__ __forceinline__ T smooth_div(T num, T den)
{
    return num / den;
}
template <>
__device__ __forceinline__ float smooth_div<float>(float num, float den)
{
    return num * __builtin_amdgcn_rcpf(den);
}

}  // namespace tbfg
