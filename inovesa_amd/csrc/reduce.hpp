// Wavefront (64 lanes) and workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace vfps_impl {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, 64);
        v = v > u ? v : u;
    }
    return v;
}

// Sum over the workgroup; `part` holds one entry per wavefront (<= 16).
// Every thread receives the total.  Ends with a barrier, so the scratch
// can be reused right away.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* part)
{
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const unsigned nwaves = (blockDim.x + 63) / 64;
    v = wave_sum(v);
    if (lane == 0) {
        part[wave] = v;
    }
    __syncthreads();
    T t = (threadIdx.x < nwaves) ? part[threadIdx.x] : T(0);
    if (wave == 0) {
        t = wave_sum(t);
        if (lane == 0) {
            part[0] = t;
        }
    }
    __syncthreads();
    T r = part[0];
    __syncthreads();
    return r;
}

}  // namespace vfps_impl
