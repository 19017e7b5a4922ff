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

// One DPP move of a double: dpp_ctrl applied to both 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF,
                                               0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF,
                                               0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum of a double over the 64 lanes, every lane receives it.  Within each
// row of 16 lanes: neighbours (quad_perm 1,0,3,2), pairs (quad_perm
// 2,3,0,1), half-row mirror, row mirror; then the rows' sums read from
// lanes 0, 16, 32, 48 and added in that order.  Register moves only (no
// LDS permutes); the order is fixed, so every workgroup shape gives the
// same bits.
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    auto lane = [&](int l) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
        return __hiloint2double(hi, lo);
    };
    return ((lane(0) + lane(16)) + lane(32)) + lane(48);
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

// Sum of a double over a workgroup of NT threads (a multiple of 64, <= 16
// wavefronts); every thread receives it.  wave_sum_dpp within each
// wavefront, then every thread adds the wavefront sums in wavefront order.
// Ends with a barrier, so the scratch can be reused right away.
template <unsigned NT>
__device__ __forceinline__ double block_sum_dpp(double v, double* part)
{
    static_assert(NT % 64 == 0 && NT <= 1024, "whole wavefronts, at most 16");
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    constexpr unsigned nwaves = NT / 64;
    v = wave_sum_dpp(v);
    if (lane == 0) {
        part[wave] = v;
    }
    __syncthreads();
    double t = part[0];
    for (unsigned w = 1; w < nwaves; w++) {
        t += part[w];
    }
    __syncthreads();
    return t;
}

// Two sums over the workgroup in one pass; `part` holds two entries per
// wavefront (<= 8 wavefronts).  The per-wave and cross-wave orders are those
// of block_sum.  Ends with a barrier.
template <typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* part)
{
    const unsigned lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const unsigned nwaves = (blockDim.x + 63) / 64;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        part[wave] = a;
        part[8 + wave] = b;
    }
    __syncthreads();
    T ta = (threadIdx.x < nwaves) ? part[threadIdx.x] : T(0);
    T tb = (threadIdx.x < nwaves) ? part[8 + threadIdx.x] : T(0);
    if (wave == 0) {
        // only wavefront 0 read the entries, before lane 0 writes
        ta = wave_sum(ta);
        tb = wave_sum(tb);
        if (lane == 0) {
            part[0] = ta;
            part[8] = tb;
        }
    }
    __syncthreads();
    a = part[0];
    b = part[8];
    __syncthreads();
}

}  // namespace vfps_impl
