// Passive particle tracking (SourceMap::applyTo / applyToAll,
// src/SM/SourceMap.cpp:111-117) on the device: one thread per particle runs
// the applyTo of each map of a step, in the reference's float arithmetic.
// The stochastic Fokker-Planck tracking draws its normal noise from a
// counter-based generator (seed, particle, map, step) where the reference
// uses std::mt19937 seeded from std::random_device.
#include <cmath>

#include "common.hpp"
#include "maps.hpp"
#include "track.hpp"

namespace vfps_impl {

namespace {

// SplitMix64 finaliser: a well-mixed 64-bit hash of the counter
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// one standard normal deviate from the counter (Box-Muller on two 24-bit
// uniforms in (0, 1])
__device__ __forceinline__ float normal_from(uint64_t seed, uint64_t a,
                                             uint64_t b)
{
    const uint64_t h = mix64(seed ^ mix64(a ^ mix64(b)));
    const float u1 = ((float)(uint32_t)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
    const float u2 = (float)(uint32_t)((h >> 16) & 0xFFFFFFu) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}

// std::modf on a float, integer part cast to meshindex_t as the reference
// does
__device__ __forceinline__ float frac_idx(float v, uint32_t* i)
{
    float ip;
    const float f = vf_modf(v, &ip);
    *i = vf_f2u(ip);
    return f;
}

__device__ void apply_op(const TrackOp& o, float2& pos, uint64_t seed,
                         uint64_t pid, uint64_t count)
{
    switch (o.kind) {
    case TRACK_KICK_X: {
        // src/SM/KickMap.cpp:282-291
        uint32_t yi;
        const float yf = frac_idx(pos.y, &yi);
        if (yi + 1 < o.pd) {
            pos.x = pos.x - ((1 - yf) * o.off[yi] + yf * o.off[yi + 1]);
        }
        pos.x = fmaxf(1.0f, fminf(pos.x, (float)(o.kd - 1)));
        break;
    }
    case TRACK_KICK_Y: {
        // src/SM/KickMap.cpp:292-302
        uint32_t xi;
        const float xf = frac_idx(pos.x, &xi);
        if (xi + 1 < o.pd) {
            pos.y = pos.y - ((1 - xf) * o.off[xi] + xf * o.off[xi + 1]);
        }
        pos.y = fmaxf(1.0f, fminf(pos.y, (float)(o.kd - 1)));
        break;
    }
    case TRACK_ROTATION: {
        // src/SM/RotationMap.cpp:178-188
        const float dx = pos.x - o.zb0, dy = pos.y - o.zb1;
        float2 t;
        t.x = o.cs * dx + o.sn * dy + o.zb0;
        t.y = o.cs * dy - o.sn * dx + o.zb1;
        pos = t;
        break;
    }
    case TRACK_FP:
        switch (o.fptrack) {
        case 1: {
            // approximation1 (src/SM/FokkerPlanckMap.cpp:219-233)
            uint32_t yi = vf_f2u(floorf(pos.y));
            yi = yi < o.ysize ? yi : o.ysize - 1;
            float offset = 0.0f;
            for (uint32_t j = 0; j < o.ip; j++) {
                const uint32_t idx = o.fp_i[(size_t)j * o.ysize + yi];
                const float w = o.fp_w[(size_t)j * o.ysize + yi];
                const float dy = (float)yi - (float)idx;
                offset = offset + dy * w;
            }
            pos.y = fmaxf(1.0f, fminf(pos.y + offset, (float)(o.ysize - 1)));
            break;
        }
        case 2: {
            // approximation2 (src/SM/FokkerPlanckMap.cpp:234-262)
            int32_t xi = (int32_t)floorf(pos.x);
            xi = xi < (int32_t)o.nx - 1 ? xi : (int32_t)o.nx - 1;
            int32_t yi = (int32_t)floorf(pos.y);
            yi = yi < (int32_t)o.ysize - 1 ? yi : (int32_t)o.ysize - 1;
            if (xi < 0 || yi < 0) break;  // outside the mesh: no data
            const size_t offs = (size_t)xi * o.ysize;
            float offset = 0.0f, charge = 0.0f;
            for (uint32_t j = 0; j < o.ip; j++) {
                const uint32_t idx = o.fp_i[(size_t)j * o.ysize + yi];
                const float w = o.fp_w[(size_t)j * o.ysize + yi];
                const float dw = o.data_in[offs + idx] * w;
                charge = charge + dw;
                offset = offset + dw * (float)((int32_t)idx - yi);
            }
            offset = offset / charge;
            pos.y = fmaxf(1.0f, fminf(pos.y + offset, (float)(o.ysize - 1)));
            break;
        }
        case 3:
            // stochastic (src/SM/FokkerPlanckMap.cpp:263-265)
            pos.y = pos.y - (pos.y * o.e1 + o.sigma * normal_from(seed, pid,
                                                                  count));
            break;
        default:
            break;
        }
        break;
    default:
        break;
    }
}

__global__ void __launch_bounds__(256)
track_kernel(TrackProgram prog, float2* pos, uint32_t n,
             const uint32_t* counter, uint32_t call)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint64_t count = (uint64_t)(counter ? *counter : 0u) + call;
    float2 p = pos[i];
    for (int k = 0; k < prog.nops; k++) {
        apply_op(prog.op[k], p, prog.seed, i, count * TRACK_MAX_OPS + k);
    }
    pos[i] = p;
}

}  // namespace

bool track_op(const vfps_map* m, const float* kick_off, TrackOp* op)
{
    *op = TrackOp{};
    const vfps_mesh* ps = m->in;
    switch (m->kind) {
    case MAP_KICK_X:
    case MAP_KICK_Y: {
        const bool x = m->kind == MAP_KICK_X;
        op->kind = x ? TRACK_KICK_X : TRACK_KICK_Y;
        op->off = kick_off ? kick_off : m->offset;
        op->pd = ps->n;   // square meshes: _meshsize_pd = _meshsize_kd = n
        op->kd = ps->n;
        return true;
    }
    case MAP_ROTATION:
        op->kind = TRACK_ROTATION;
        op->cs = m->cos_dt;
        op->sn = m->sin_dt;
        op->zb0 = ps->ax0.zerobin;
        op->zb1 = ps->ax1.zerobin;
        return true;
    case MAP_FP:
        op->kind = TRACK_FP;
        op->fptrack = m->fptrack;
        op->ip = (uint32_t)m->ip;
        op->ysize = ps->n;
        op->nx = ps->n;
        op->fp_w = m->fp_w;
        op->fp_i = m->fp_i;
        op->data_in = ps->data;
        op->e1 = m->fp_e1;
        // std::normal_distribution<meshaxis_t>(0, std::sqrt(2*e1)/delta_p)
        op->sigma = std::sqrt(2 * m->fp_e1) / ps->ax1.delta;
        return true;
    default:
        return false;  // Identity
    }
}

int launch_track(const TrackProgram& prog, float2* pos, uint32_t n,
                 const uint32_t* counter, uint32_t call, hipStream_t st)
{
    if (n == 0 || prog.nops == 0) {
        return VFPS_OK;
    }
    hipLaunchKernelGGL(track_kernel, dim3(ceil_div(n, 256u)), dim3(256), 0,
                       st, prog, pos, n, counter, call);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

}  // namespace vfps_impl

using namespace vfps_impl;

extern "C" {

int vfps_map_apply_to(vfps_map* map, float* pos, uint32_t n, uint64_t seed)
{
    VFPS_CHECK_ARG(map && (pos || n == 0), "applyToAll: NULL argument");
    if (n == 0) return VFPS_OK;
    VFPS_TRY_BEGIN
    TrackProgram prog{};
    prog.seed = seed;
    prog.nops = track_op(map, nullptr, &prog.op[0]) ? 1 : 0;
    if (prog.nops == 0) {
        return VFPS_OK;
    }
    hipStream_t st = map->in->dev->stream;
    float2* d = nullptr;
    VFPS_HIP(hipMalloc(&d, sizeof(float2) * n));
    hipError_t e = hipMemcpyAsync(d, pos, sizeof(float2) * n,
                                  hipMemcpyHostToDevice, st);
    int rc = e == hipSuccess
                 ? launch_track(prog, d, n, nullptr, map->track_calls, st)
                 : fail(VFPS_ERR_HIP, hipGetErrorString(e));
    if (rc == VFPS_OK) {
        e = hipMemcpyAsync(pos, d, sizeof(float2) * n, hipMemcpyDeviceToHost,
                           st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    hipFree(d);
    map->track_calls++;
    return rc;
    VFPS_TRY_END
}

}  // extern "C"
