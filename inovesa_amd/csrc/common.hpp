// Internal definitions shared by the translation units of libvfps.
//
// All HIP sources are compiled with -ffp-contract=off: the map kernels
// reproduce the reference CPU arithmetic operation by operation, so a
// multiply followed by an add must stay two roundings.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "vfps_gpu.h"

namespace vfps_impl {

// ---- error plumbing ------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define VFPS_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            return ::vfps_impl::fail(VFPS_ERR_HIP,                            \
                                     std::string(#call) + ": " +              \
                                         hipGetErrorString(e_));              \
        }                                                                     \
    } while (0)

#define VFPS_CHECK_ARG(cond, msg)                                             \
    do {                                                                      \
        if (!(cond)) {                                                        \
            return ::vfps_impl::fail(VFPS_ERR_INVALID, msg);                  \
        }                                                                     \
    } while (0)

#define VFPS_TRY_BEGIN try {
#define VFPS_TRY_END                                                          \
    }                                                                         \
    catch (const std::bad_alloc&) {                                           \
        return ::vfps_impl::fail(VFPS_ERR_NOMEM, "out of host memory");       \
    }                                                                         \
    catch (const std::exception& ex) {                                        \
        return ::vfps_impl::fail(VFPS_ERR_INVALID, ex.what());                \
    }

// ---- host-side geometry: inc/PS/Ruler.hpp:31-51 ---------------------------
struct Ruler {
    uint32_t steps = 0;
    float min = 0, max = 0, delta = 0, zerobin = 0;
    std::vector<float> data;
    Ruler() = default;
    Ruler(uint32_t s, float mn, float mx);
};

// Simpson weights, src/PS/PhaseSpace.cpp:486-501
std::vector<float> simpson_weights(uint32_t n, float delta0);

}  // namespace vfps_impl

// ---- handles ---------------------------------------------------------------
struct vfps_dev {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int ncu = 0;
    size_t lds_max = 65536;
    std::string arch;
};

struct vfps_mesh {
    vfps_dev* dev = nullptr;
    uint32_t nb = 0, n = 0;
    vfps_impl::Ruler ax0, ax1;
    double qscale = 0, pscale = 0, charge = 0, current = 0;
    std::vector<float> ws_h;         // Simpson weights (host copy)
    std::vector<float> filling_set;  // set bunch populations
    float* data = nullptr;           // [nb][n][n]
    float* proj = nullptr;           // [2][nb][n]
    float* ws = nullptr;             // [n]
    float* fill_set_d = nullptr;     // [nb]
    double* filling = nullptr;       // [nb]
    double* integral = nullptr;      // [1]
    size_t cells() const { return (size_t)nb * n * n; }
};

enum MapKind {
    MAP_KICK_X = 0,
    MAP_KICK_Y = 1,
    MAP_ROTATION = 2,
    MAP_FP = 3,
    MAP_IDENTITY = 4
};

struct vfps_map {
    MapKind kind = MAP_IDENTITY;
    vfps_mesh* in = nullptr;
    vfps_mesh* out = nullptr;
    int interp = 4;
    // kick maps
    float* offset = nullptr;  // device, noff entries
    size_t noff = 0;
    vfps_field* field = nullptr;  // WakePotentialMap
    // rotation
    float cos_dt = 1, sin_dt = 0;
    int clamp = 0;
    uint32_t rotmapsize = 0;
    // Fokker-Planck
    int ip = 0;
    float* fp_w = nullptr;      // [ip][n]
    uint32_t* fp_i = nullptr;   // [ip][n]
    // FokkerPlanckMap particle tracking (FPTracking, src/SM/FokkerPlanckMap.cpp
    // :214-268): 0 none, 1 approximation1, 2 approximation2, 3 stochastic
    int fptrack = 0;
    float fp_e1 = 0;
    uint64_t track_seed = 0;   // stochastic tracking: generator seed
    uint32_t track_calls = 0;  // applyTo calls so far (generator counter)
    // DynamicRFKickMap: RF kick offsets recomputed before every apply from
    // the next (phase, amplitude) of a modulation queue built at
    // construction
    bool dynrf = false;
    bool rf_linear = false;
    float rf_angle = 0, rf_revolutionpart = 0, rf_V = 0, rf_f = 0, rf_V0 = 0;
    float rf_syncphase = 0;
    std::vector<float> modulation;  // [steps][2]: phase, amplitude
    uint32_t mod_next = 0;          // entries applied so far
    float* stage_h = nullptr;       // pinned host offsets of the next apply
    hipEvent_t stage_done = nullptr;  // the last upload from stage_h
};

struct vfps_impedance {
    vfps_dev* dev = nullptr;
    uint32_t n = 0;
    float fmax = 0;
    float2* z = nullptr;  // device complex<float> table
};

struct vfps_field {
    vfps_dev* dev = nullptr;
    vfps_mesh* ps = nullptr;
    vfps_impedance* z = nullptr;
    uint32_t nmax = 0, nb = 0, spacing = 0;
    std::vector<uint32_t> buckets;
    uint32_t* buckets_d = nullptr;
    float wakescaling = 0;       // already divided by nmax
    double f_rev = 0;
    float revolutionpart = 1;
    float* wake = nullptr;       // [nb][n]
    float* wake_padded = nullptr;  // [nmax] unscaled C2R output
    float* padded = nullptr;     // [nmax]
    double2* tw_m = nullptr;     // e^{-2 pi i t/M}, t < M (M = nmax/2)
    double2* tw_n = nullptr;     // e^{-2 pi i k/nmax}, k < M
    float* csr_renorm = nullptr; // [nmax]
    float* spectrum = nullptr;   // [nb][nmax]
    float* power = nullptr;      // [nb]
    float csr_fc = -1;           // fc the renorm table was built for
    float fdelta = 0;            // frequency ruler step (1/(dq (nmax-1)))
};

// ---- device helpers (bit-exact restatements of the reference arithmetic) --
namespace vfps_impl {

// std::modf: integer part and fraction carrying the sign of v
__host__ __device__ __forceinline__ float vf_modf(float v, float* ip)
{
    float t = truncf(v);
    *ip = t;
    return copysignf(v - t, v);
}

// static_cast<uint32_t>(float) of an integral value, as g++ on x86-64
// evaluates it: (-1, 2^32) converts; <= -1 wraps through int64 (so -1
// becomes 0xFFFFFFFF, out of range for every mesh).
__host__ __device__ __forceinline__ uint32_t vf_f2u(float v)
{
    if (v > -1.0f && v < 4294967296.0f) {
        return (uint32_t)v;
    }
    if (v > -9.2e18f && v < 9.2e18f) {
        return (uint32_t)(uint64_t)(int64_t)v;
    }
    return 0u;
}

// SourceMap::calcCoefficiants (src/SM/SourceMap.cpp:142-170)
template <int IT>
__host__ __device__ __forceinline__ void coefficients(float f, float* ic)
{
    const float one = 1.0f, two = 2.0f;
    if constexpr (IT == 1) {
        ic[0] = 1.0f;
    } else if constexpr (IT == 2) {
        ic[0] = one - f;
        ic[1] = f;
    } else if constexpr (IT == 3) {
        ic[0] = f * (f - one) / two;
        ic[1] = one - f * f;
        ic[2] = f * (f + one) / two;
    } else {
        const float m6 = (float)(-1. / 6.);
        const float p6 = (float)(1. / 6.);
        ic[0] = (f - one) * (f - two) * f * m6;
        ic[1] = (f + one) * (f - one) * (f - two) / two;
        ic[2] = (two - f) * f * (f + one) / two;
        ic[3] = f * (f + one) * (f - one) * p6;
    }
}

// One row of KickMap::updateSM (src/SM/KickMap.cpp:349-386): the tap
// (index, weight) list of a kick by `offset` on an axis of kd cells.
template <int IT>
__host__ __device__ __forceinline__ void kick_taps(float offset, uint32_t kd,
                                                   uint32_t* idx, float* w)
{
    float poffs = (float)(kd / 2) + offset;
    float qp_int;
    float xip = vf_modf(poffs, &qp_int);
    uint32_t jd = vf_f2u(qp_int);
    if (jd < kd) {
        float smc[4];
        coefficients<IT>(xip, smc);
#pragma unroll
        for (int j1 = 0; j1 < IT; j1++) {
            uint32_t j0 = jd + (uint32_t)j1 - (uint32_t)((IT - 1) / 2);
            if (j0 < kd) {
                idx[j1] = j0;
                w[j1] = smc[j1];
            } else {
                idx[j1] = kd / 2;
                w[j1] = 0.0f;
            }
        }
    } else {
#pragma unroll
        for (int j1 = 0; j1 < IT; j1++) {
            idx[j1] = kd / 2;
            w[j1] = 0.0f;
        }
    }
}

// std::max / std::min as libstdc++ defines them
__host__ __device__ __forceinline__ float std_max(float a, float b)
{
    return (a < b) ? b : a;
}
__host__ __device__ __forceinline__ float std_min(float a, float b)
{
    return (b < a) ? b : a;
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace vfps_impl
