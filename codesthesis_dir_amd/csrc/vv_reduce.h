// vv_reduce.h -- workgroup reductions of the HIP kernels (device code only).
//
// Reductions over a field are done in two stages without atomics: every
// workgroup writes its partial result with a plain store, and one small
// consumer (a single-workgroup kernel, or the host) combines the partials.
// Maxima propagate NaN, so a non-finite field is never reported as finite.
#pragma once

#include <hip/hip_runtime.h>

namespace vv {

__device__ __forceinline__ double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Maximum over the workgroup; the result is valid in thread 0.  `sh` holds
// one double per wavefront of the workgroup (at most 16).
__device__ __forceinline__ double block_max(double m, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_down(m, off, 64));
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < nw; ++k) m = nanmax(m, sh[k]);
    return m;
}

// Sum over the workgroup (small integer counts held in doubles); the result
// is valid in thread 0.
__device__ __forceinline__ double block_sum(double s, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < nw; ++k) s += sh[k];
    return s;
}

}  // namespace vv
