// SourceMaps: KickMap (x, y), DriftMap, RFKickMap, RotationMap,
// FokkerPlanckMap, Identity, WakePotentialMap (inc/SM/*.hpp, src/SM/*.cpp).
//
// One kernel per SourceMap::apply().  The tap lists the reference keeps in
// its `hi` arrays are recomputed in registers from the offsets (kicks) or
// the geometry (rotation); the Fokker-Planck weights, one set per y, are a
// small table built at construction as in the reference.  Every tap is
// accumulated in the reference's order with separate multiply and add
// roundings, so outputs equal the reference CPU path bit for bit.
#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

#include "common.hpp"
#include "maps.hpp"

namespace vfps_impl {

// ---- kernels -------------------------------------------------------------

// KickMap::apply, Axis::y (src/SM/KickMap.cpp:244-267): one workgroup per
// (bunch, x) row; the row's taps come from offset[x].  For nb > 1 the
// reference indexes its tap table with min(b,last)*pd + x*ip + j; the
// kernel resolves that entry the same way.
template <int IT>
__global__ void __launch_bounds__(256)
kick_y_kernel(uint32_t nb, uint32_t n, const float* __restrict__ off,
              const float* __restrict__ in, float* __restrict__ out)
{
    const uint32_t row = blockIdx.x;
    const uint32_t b = row / n, x = row % n;
    const uint32_t last = nb - 1;
    const size_t offs2 = (size_t)(b < last ? b : last) * n;
    uint32_t idx[IT];
    float w[IT];
    if (offs2 == 0) {
        kick_taps<IT>(off[x], n, idx, w);
    } else {
#pragma unroll
        for (int j = 0; j < IT; j++) {
            size_t e = offs2 + (size_t)x * IT + j;
            uint32_t ii[IT];
            float ww[IT];
            kick_taps<IT>(off[e / IT], n, ii, ww);
            idx[j] = ii[e % IT];
            w[j] = ww[e % IT];
        }
    }
    const float* src = in + (size_t)row * n;
    float* dst = out + (size_t)row * n;
    const int32_t half = (int32_t)(n / 2);
    for (uint32_t y = threadIdx.x; y < n; y += blockDim.x) {
        // every tap's sample in flight at once (an out-of-range tap loads
        // sample 0 and is not added)
        uint32_t ys[IT];
        float s[IT];
#pragma unroll
        for (int j = 0; j < IT; j++) {
            ys[j] = (uint32_t)((int32_t)(y + idx[j]) - half);
            s[j] = src[ys[j] < n ? ys[j] : 0u];
        }
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < IT; j++) {
            if (ys[j] < n) {
                v = v + s[j] * w[j];
            }
        }
        dst[y] = v;
    }
}

// KickMap::apply, Axis::x (src/SM/KickMap.cpp:221-243): one thread per
// column y (taps from offset[y]), sliding down a block of rows.
template <int IT>
__global__ void __launch_bounds__(256)
kick_x_kernel(uint32_t n, uint32_t rows_per_block,
              const float* __restrict__ off, const float* __restrict__ in,
              float* __restrict__ out)
{
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = blockIdx.z;
    const uint32_t x0 = blockIdx.y * rows_per_block;
    if (y >= n) {
        return;
    }
    uint32_t idx[IT];
    float w[IT];
    kick_taps<IT>(off[y], n, idx, w);
    const float* src = in + (size_t)b * n * n;
    float* dst = out + (size_t)b * n * n;
    const int32_t half = (int32_t)(n / 2);
    const uint32_t x1 = min(x0 + rows_per_block, n);
    for (uint32_t x = x0; x < x1; x++) {
        // every tap's sample in flight at once (an out-of-range tap loads
        // row 0 and is not added)
        uint32_t xs[IT];
        float s[IT];
#pragma unroll
        for (int j = 0; j < IT; j++) {
            xs[j] = (uint32_t)((int32_t)(x + idx[j]) - half);
            s[j] = src[(size_t)(xs[j] < n ? xs[j] : 0u) * n + y];
        }
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < IT; j++) {
            if (xs[j] < n) {
                v = v + s[j] * w[j];
            }
        }
        dst[(size_t)x * n + y] = v;
    }
}

// FokkerPlanckMap::apply (src/SM/FokkerPlanckMap.cpp:194-212): the same
// per-y stencil on every row; tables w, ix are [IP][n].
template <int IP>
__global__ void __launch_bounds__(256)
fp_kernel(uint32_t n, const float* __restrict__ w,
          const uint32_t* __restrict__ ix, const float* __restrict__ in,
          float* __restrict__ out)
{
    const size_t row = blockIdx.x;
    const float* src = in + row * n;
    float* dst = out + row * n;
    for (uint32_t y = threadIdx.x; y < n; y += blockDim.x) {
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < IP; j++) {
            v = v + src[ix[(size_t)j * n + y]] * w[(size_t)j * n + y];
        }
        dst[y] = v;
    }
}

// RotationMap (src/SM/RotationMap.cpp:97-176, 190-262): taps generated on
// the fly (genHInfo), optional clamp to the inner 2x2 taps with the CPU
// path's initial bounds (FLT_MIN, FLT_MAX).
template <int IT, bool CLAMP>
__global__ void __launch_bounds__(256)
rotation_kernel(uint32_t n, RotGeom g, const float* __restrict__ in,
                float* __restrict__ out)
{
    const uint32_t y0 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t x0 = blockIdx.y;
    if (y0 >= n) {
        return;
    }
    out[(size_t)x0 * n + y0] = rotate_cell<IT, CLAMP>(n, g, x0, y0, in);
}

// ---- launchers -------------------------------------------------------------

template <int IT>
static void launch_kick_it(const vfps_map* m, hipStream_t st, const float* in,
                           float* out)
{
    const uint32_t n = m->in->n, nb = m->in->nb;
    if (m->kind == MAP_KICK_Y) {
        hipLaunchKernelGGL(kick_y_kernel<IT>, dim3(nb * n), dim3(256), 0, st,
                           nb, n, m->offset, in, out);
    } else {
        // rows per block: enough blocks to give every CU a few (a 512^2
        // mesh in blocks of 16 rows was 64 blocks: 12.4 us), at most 16
        // rows sharing one column's taps
        const uint64_t cols = ceil_div(n, 256);
        const uint64_t want = 1024;
        const uint32_t rpb = (uint32_t)std::min<uint64_t>(
            16, std::max<uint64_t>(1, cols * n * nb / want));
        hipLaunchKernelGGL(kick_x_kernel<IT>,
                           dim3(ceil_div(n, 256), ceil_div(n, rpb), nb),
                           dim3(256), 0, st, n, rpb, m->offset, in, out);
    }
}

int launch_map(const vfps_map* m, hipStream_t st, const float* in, float* out)
{
    const uint32_t n = m->in->n, nb = m->in->nb;
    switch (m->kind) {
    case MAP_KICK_X:
    case MAP_KICK_Y:
        switch (m->interp) {
        case 1: launch_kick_it<1>(m, st, in, out); break;
        case 2: launch_kick_it<2>(m, st, in, out); break;
        case 3: launch_kick_it<3>(m, st, in, out); break;
        default: launch_kick_it<4>(m, st, in, out); break;
        }
        break;
    case MAP_FP:
        switch (m->ip) {
        case 3:
            hipLaunchKernelGGL(fp_kernel<3>, dim3(nb * n), dim3(256), 0, st, n,
                               m->fp_w, m->fp_i, in, out);
            break;
        case 4:
            hipLaunchKernelGGL(fp_kernel<4>, dim3(nb * n), dim3(256), 0, st, n,
                               m->fp_w, m->fp_i, in, out);
            break;
        default:
            hipLaunchKernelGGL(fp_kernel<5>, dim3(nb * n), dim3(256), 0, st, n,
                               m->fp_w, m->fp_i, in, out);
            break;
        }
        break;
    case MAP_ROTATION: {
        RotGeom g = rot_geom(m);
        dim3 grid(ceil_div(n, 256), n);
        for (uint32_t b = 0; b < nb; b++) {
            const float* src = in + (size_t)b * n * n;
            float* dst = out + (size_t)b * n * n;
            if (m->clamp) {
                hipLaunchKernelGGL((rotation_kernel<4, true>), grid, dim3(256),
                                   0, st, n, g, src, dst);
            } else {
                switch (m->interp) {
                case 1:
                    hipLaunchKernelGGL((rotation_kernel<1, false>), grid,
                                       dim3(256), 0, st, n, g, src, dst);
                    break;
                case 2:
                    hipLaunchKernelGGL((rotation_kernel<2, false>), grid,
                                       dim3(256), 0, st, n, g, src, dst);
                    break;
                case 3:
                    hipLaunchKernelGGL((rotation_kernel<3, false>), grid,
                                       dim3(256), 0, st, n, g, src, dst);
                    break;
                default:
                    hipLaunchKernelGGL((rotation_kernel<4, false>), grid,
                                       dim3(256), 0, st, n, g, src, dst);
                    break;
                }
            }
        }
        break;
    }
    case MAP_IDENTITY:
        VFPS_HIP(hipMemcpyAsync(out, in, sizeof(float) * m->in->cells(),
                                hipMemcpyDeviceToDevice, st));
        break;
    }
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

RotGeom rot_geom(const vfps_map* m)
{
    RotGeom g;
    g.qmin = m->in->ax0.min;
    g.d0 = m->in->ax0.delta;
    g.zb0 = m->in->ax0.zerobin;
    g.pmin = m->in->ax1.min;
    g.d1 = m->in->ax1.delta;
    g.zb1 = m->in->ax1.zerobin;
    g.cs = m->cos_dt;
    g.sn = m->sin_dt;
    return g;
}

// ---- host-side tables (constructors) --------------------------------------

// DriftMap offsets (src/SM/DriftMap.cpp:17-28)
std::vector<float> drift_offsets(const Ruler& ax0, const Ruler& ax1,
                                 const float* slip, uint32_t nslip,
                                 float pscale, float E0)
{
    std::vector<float> off(ax1.steps);
    for (uint32_t y = 0; y < ax1.steps; y++) {
        float o = 0.0f;
        for (uint32_t i = 0; i < nslip; i++) {
            float t = slip[i] * ax1.data[y];
            float u = ax1.data[y] * pscale / E0;
            o = o + t * std::pow(u, (float)i);
        }
        off[y] = o / ax0.delta;
    }
    return off;
}

// _bl2phase of src/RFKickMap.cpp:30,53: scale("Meter")/c*f_RF*2pi in
// double, held as float
static float rf_bl2phase(double qscale, float f_RF)
{
    const double pi = 3.141592653589793238462643383279502884;
    return (float)((double)(float)qscale / 2.99792458e8 * (double)f_RF *
                   (2.0 * pi));
}

// RFKickMap::_calcKick(phase, ampl), linear (src/SM/RFKickMap.cpp:59-69)
std::vector<float> rfkick_offsets_linear_at(const Ruler& ax0, float angle,
                                            double qscale, float f_RF,
                                            float syncphase, float phase,
                                            float ampl)
{
    const float bl2phase = rf_bl2phase(qscale, f_RF);
    const float phaseoffs = syncphase - phase;
    const float xcenter = ax0.zerobin;
    std::vector<float> off(ax0.steps);
    for (uint32_t x = 0; x < ax0.steps; x++) {
        float o = std::tan(angle) * (xcenter - (float)x);
        o = o + std::tan(angle) * phaseoffs / bl2phase / ax0.delta;
        o = o * ampl;
        off[x] = o;
    }
    return off;
}

// RFKickMap::_calcKick(phase, ampl), sinusoidal (src/SM/RFKickMap.cpp:70-75)
std::vector<float> rfkick_offsets_at(const Ruler& ax0, const Ruler& ax1,
                                     float revolutionpart, float V_RF,
                                     float f_RF, float V0, double qscale,
                                     float pscale, float phase, float ampl)
{
    const float bl2phase = rf_bl2phase(qscale, f_RF);
    std::vector<float> off(ax0.steps);
    for (uint32_t x = 0; x < ax0.steps; x++) {
        float a = -ampl * V_RF;
        float s = std::sin(ax0.data[x] * bl2phase + phase);
        float o = revolutionpart * (a * s + V0);
        o = o / ax1.delta;
        o = o / pscale;
        off[x] = o;
    }
    return off;
}

// RFKickMap offsets at construction: _calcKick(_syncphase), linear
// (_syncphase 0) and sinusoidal (_syncphase = asin(V0/V_RF))
std::vector<float> rfkick_offsets_linear(const Ruler& ax0, float angle,
                                         double qscale, float f_RF)
{
    return rfkick_offsets_linear_at(ax0, angle, qscale, f_RF, 0.0f, 0.0f, 1.0f);
}

std::vector<float> rfkick_offsets(const Ruler& ax0, const Ruler& ax1,
                                  float revolutionpart, float V_RF,
                                  float f_RF, float V0, double qscale,
                                  float pscale)
{
    const float syncphase = std::asin(V0 / V_RF);
    return rfkick_offsets_at(ax0, ax1, revolutionpart, V_RF, f_RF, V0, qscale,
                             pscale, syncphase, 1.0f);
}

// DynamicRFKickMap::__calcModulation (src/SM/DynamicRFKickMap.cpp:80-95):
// per step, a normally distributed phase and amplitude noise and a
// sinusoidal phase modulation, drawn in the reference's order from the
// same generators (std::mt19937, std::normal_distribution<float>), seeded
// by the caller where the reference takes std::random_device.  [steps][2].
std::vector<float> dynrf_modulation(uint32_t steps, uint32_t seed,
                                    float syncphase, double revolutionpart,
                                    float phasespread, float amplspread,
                                    float modampl, double modtimeincrement)
{
    const double two_pi = 6.283185307179586476925286766559005768;
    const float phasenoise = (float)(phasespread / std::sqrt(revolutionpart));
    const float amplnoise = (float)(amplspread / std::sqrt(revolutionpart));
    const float modtimedelta = (float)(two_pi * modtimeincrement);
    std::mt19937 prng(seed);
    std::normal_distribution<float> dist(0.0f, 1.0f);
    std::vector<float> rv(2 * (size_t)steps);
    for (uint32_t i = 0; i < steps; i++) {
        const float pn = dist(prng) * phasenoise;
        const float an = dist(prng) * amplnoise;
        const float pm = modampl * std::sin(modtimedelta * (float)i);
        rv[2 * (size_t)i] = syncphase + pn + pm;
        rv[2 * (size_t)i + 1] = 1.0f + an;
    }
    return rv;
}

// FokkerPlanckMap weights (src/SM/FokkerPlanckMap.cpp:31-118); the
// five-point stencil (deriv 5) is this library's extension.  Returns ip,
// tables are [n][ip] (row-major like the reference's hi array).
int fp_weights(const Ruler& ax1, int fptype, float e1, int deriv,
               std::vector<uint32_t>& ix, std::vector<float>& w)
{
    const uint32_t n = ax1.steps;
    const float d = ax1.delta;
    const float e1_2d = e1 / (2.0f * d);
    const float e1_6d = e1 / (6.0f * d);
    const float e1_d2 = e1 / (d * d);
    const float ycenter = ax1.zerobin;
    const bool damp = (fptype != 0 && fptype != 2);
    const bool diff = (fptype != 0 && fptype != 1);
    int ip;
    auto set = [&](uint32_t j, int k, uint32_t i, float wt) {
        ix[(size_t)j * ip + k] = i;
        w[(size_t)j * ip + k] = wt;
    };
    switch (deriv) {
    case 3:
        ip = 3;
        ix.assign((size_t)n * ip, 0);
        w.assign((size_t)n * ip, 0.0f);
        for (uint32_t j = 1; j < n - 1; j++) {
            set(j, 0, j - 1, 0.0f);
            set(j, 1, j, 1.0f);
            set(j, 2, j + 1, 0.0f);
            float* h = &w[(size_t)j * ip];
            if (damp) {
                const float pos = ax1.data[j];
                h[0] = h[0] + (-e1_2d) * pos;
                h[1] = h[1] + e1;
                h[2] = h[2] + e1_2d * pos;
            }
            if (diff) {
                h[0] = h[0] + e1_d2;
                h[1] = h[1] + (-2.0f) * e1_d2;
                h[2] = h[2] + e1_d2;
            }
        }
        return ip;
    case 4: {
        ip = 4;
        ix.assign((size_t)n * ip, 0);
        w.assign((size_t)n * ip, 0.0f);
        for (uint32_t j = 2; (float)j < ycenter; j++) {
            const float pos = ax1.data[j];
            set(j, 0, j - 2, 0.0f);
            set(j, 1, j - 1, 0.0f);
            set(j, 2, j, 1.0f);
            set(j, 3, j + 1, 0.0f);
            float* h = &w[(size_t)j * ip];
            if (damp) {
                h[0] = h[0] + e1_6d * 1.0f * pos;
                h[1] = h[1] + e1_6d * (-6.0f) * pos;
                h[2] = h[2] + (e1 + e1_6d * 3.0f * pos);
                h[3] = h[3] + e1_6d * 2.0f * pos;
            }
            if (diff) {
                h[1] = h[1] + e1_d2;
                h[2] = h[2] + (-2.0f) * e1_d2;
                h[3] = h[3] + e1_d2;
            }
        }
        for (uint32_t j = vf_f2u(truncf(ycenter)); j < n - 2; j++) {
            const float pos = ax1.data[j];
            set(j, 0, j - 1, 0.0f);
            set(j, 1, j, 1.0f);
            set(j, 2, j + 1, 0.0f);
            set(j, 3, j + 2, 0.0f);
            float* h = &w[(size_t)j * ip];
            if (damp) {
                h[0] = h[0] + e1_6d * (-2.0f) * pos;
                h[1] = h[1] + (e1 + e1_6d * (-3.0f) * pos);
                h[2] = h[2] + e1_6d * 6.0f * pos;
                h[3] = h[3] + e1_6d * (-1.0f) * pos;
            }
            if (diff) {
                h[0] = h[0] + e1_d2;
                h[1] = h[1] + (-2.0f) * e1_d2;
                h[2] = h[2] + e1_d2;
            }
        }
        // rows n-2, n-1 (and 0, 1) stay {0, 0}
        for (int k = 0; k < ip; k++) {
            set(n - 2, k, 0, 0.0f);
            set(n - 1, k, 0, 0.0f);
        }
        return ip;
    }
    case 5: {
        ip = 5;
        ix.assign((size_t)n * ip, 0);
        w.assign((size_t)n * ip, 0.0f);
        const float e1_12d = e1 / (12.0f * d);
        const float e1_12d2 = e1 / (12.0f * d * d);
        const float dcoef[5] = {1.0f, -8.0f, 0.0f, 8.0f, -1.0f};
        const float ccoef[5] = {-1.0f, 16.0f, -30.0f, 16.0f, -1.0f};
        for (uint32_t j = 2; j < n - 2; j++) {
            const float pos = ax1.data[j];
            float* h = &w[(size_t)j * ip];
            for (int k = 0; k < 5; k++) {
                ix[(size_t)j * ip + k] = j + k - 2;
                h[k] = (k == 2) ? 1.0f : 0.0f;
            }
            if (damp) {
                h[2] = h[2] + e1;
                for (int k = 0; k < 5; k++) {
                    h[k] = h[k] + e1_12d * dcoef[k] * pos;
                }
            }
            if (diff) {
                for (int k = 0; k < 5; k++) {
                    h[k] = h[k] + e1_12d2 * ccoef[k];
                }
            }
        }
        return ip;
    }
    default:
        return -1;
    }
}

}  // namespace vfps_impl

using namespace vfps_impl;

static int check_pair(vfps_mesh* in, vfps_mesh* out)
{
    VFPS_CHECK_ARG(in && out, "SourceMap: NULL phase space");
    VFPS_CHECK_ARG(in->n == out->n && in->nb == out->nb && in->dev == out->dev,
                   "SourceMap: in and out phase spaces differ in size");
    return VFPS_OK;
}

static int new_kick(vfps_mesh* in, vfps_mesh* out, int interp, int axis,
                    vfps_map** map)
{
    int rc = check_pair(in, out);
    if (rc != VFPS_OK) return rc;
    VFPS_CHECK_ARG(map, "KickMap: NULL output");
    VFPS_CHECK_ARG(interp >= 1 && interp <= 4, "KickMap: interpolation type");
    VFPS_CHECK_ARG(axis == 0 || axis == 1, "KickMap: axis must be 0 or 1");
    auto* m = new vfps_map();
    m->kind = axis == 0 ? MAP_KICK_X : MAP_KICK_Y;
    m->in = in;
    m->out = out;
    m->interp = interp;
    m->noff = (size_t)in->n * in->nb;
    hipError_t e = hipMalloc(&m->offset, sizeof(float) * m->noff);
    if (e != hipSuccess) {
        delete m;
        return fail(VFPS_ERR_NOMEM, hipGetErrorString(e));
    }
    // _offset is zero-initialised (src/SM/KickMap.cpp:45-47)
    hipMemsetAsync(m->offset, 0, sizeof(float) * m->noff, in->dev->stream);
    *map = m;
    return VFPS_OK;
}

static int upload_offset(vfps_map* m, const std::vector<float>& off)
{
    VFPS_HIP(hipMemcpyAsync(m->offset, off.data(), sizeof(float) * off.size(),
                            hipMemcpyHostToDevice, m->in->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->in->dev->stream));
    return VFPS_OK;
}

extern "C" {

int vfps_map_create_kick(vfps_mesh* in, vfps_mesh* out, int interp, int clamp,
                         int axis, vfps_map** map)
{
    (void)clamp;  // ignored on the reference CPU path
    VFPS_TRY_BEGIN
    return new_kick(in, out, interp, axis, map);
    VFPS_TRY_END
}

int vfps_map_set_offset(vfps_map* map, const float* offset, size_t n)
{
    VFPS_CHECK_ARG(map && offset, "swapOffset: NULL argument");
    VFPS_CHECK_ARG(map->kind == MAP_KICK_X || map->kind == MAP_KICK_Y,
                   "swapOffset: not a KickMap");
    VFPS_CHECK_ARG(n == map->noff, "swapOffset: expected nx*nb offsets");
    VFPS_HIP(hipMemcpyAsync(map->offset, offset, sizeof(float) * n,
                            hipMemcpyHostToDevice, map->in->dev->stream));
    VFPS_HIP(hipStreamSynchronize(map->in->dev->stream));
    return VFPS_OK;
}

int vfps_map_get_offset(vfps_map* map, float* offset, size_t n)
{
    VFPS_CHECK_ARG(map && offset, "getForce: NULL argument");
    VFPS_CHECK_ARG(map->kind == MAP_KICK_X || map->kind == MAP_KICK_Y,
                   "getForce: not a KickMap");
    VFPS_CHECK_ARG(n == map->noff, "getForce: expected nx*nb offsets");
    VFPS_HIP(hipMemcpyAsync(offset, map->offset, sizeof(float) * n,
                            hipMemcpyDeviceToHost, map->in->dev->stream));
    VFPS_HIP(hipStreamSynchronize(map->in->dev->stream));
    return VFPS_OK;
}

int vfps_map_create_drift(vfps_mesh* in, vfps_mesh* out, const float* slip,
                          uint32_t nslip, float E0, int interp, int clamp,
                          vfps_map** map)
{
    (void)clamp;
    VFPS_CHECK_ARG(slip || nslip == 0, "DriftMap: NULL slip");
    VFPS_TRY_BEGIN
    int rc = new_kick(in, out, interp, 0, map);
    if (rc != VFPS_OK) return rc;
    auto off = drift_offsets(in->ax0, in->ax1, slip, nslip, (float)in->pscale,
                             E0);
    return upload_offset(*map, off);
    VFPS_TRY_END
}

int vfps_map_create_rfkick_linear(vfps_mesh* in, vfps_mesh* out, float angle,
                                  float f_RF, int interp, int clamp,
                                  vfps_map** map)
{
    (void)clamp;
    VFPS_TRY_BEGIN
    int rc = new_kick(in, out, interp, 1, map);
    if (rc != VFPS_OK) return rc;
    auto off = rfkick_offsets_linear(in->ax0, angle, in->qscale, f_RF);
    return upload_offset(*map, off);
    VFPS_TRY_END
}

int vfps_map_create_rfkick(vfps_mesh* in, vfps_mesh* out,
                           float revolutionpart, float V_RF, float f_RF,
                           float V0, int interp, int clamp, vfps_map** map)
{
    (void)clamp;
    VFPS_TRY_BEGIN
    int rc = new_kick(in, out, interp, 1, map);
    if (rc != VFPS_OK) return rc;
    auto off = rfkick_offsets(in->ax0, in->ax1, revolutionpart, V_RF, f_RF,
                              V0, in->qscale, (float)in->pscale);
    return upload_offset(*map, off);
    VFPS_TRY_END
}

int vfps_map_create_rotation(vfps_mesh* in, vfps_mesh* out, float angle,
                             int interp, int clamp, uint32_t rotmapsize,
                             vfps_map** map)
{
    int rc = check_pair(in, out);
    if (rc != VFPS_OK) return rc;
    VFPS_CHECK_ARG(map, "RotationMap: NULL output");
    VFPS_CHECK_ARG(interp >= 1 && interp <= 4, "RotationMap: interpolation");
    // src/SM/RotationMap.cpp:83-87
    if (clamp && !(interp == 4 && rotmapsize > 0)) {
        return fail(VFPS_ERR_INVALID,
                    "Clamping only supported with cubic interpolation and "
                    "rotmapsize > 0.");
    }
    VFPS_TRY_BEGIN
    auto* m = new vfps_map();
    m->kind = MAP_ROTATION;
    m->in = in;
    m->out = out;
    m->interp = interp;
    m->clamp = clamp ? 1 : 0;
    m->rotmapsize = rotmapsize;
    // _cos_dt(std::cos(-angle)), _sin_dt(std::sin(-angle)) on meshaxis_t
    m->cos_dt = std::cos(-angle);
    m->sin_dt = std::sin(-angle);
    *map = m;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_map_create_fokkerplanck(vfps_mesh* in, vfps_mesh* out, int fptype,
                                 int fptrack, float e1, int deriv,
                                 vfps_map** map)
{
    // particle tracking (applyTo) reads the mode and e1 from the map
    int rc = check_pair(in, out);
    if (rc != VFPS_OK) return rc;
    VFPS_CHECK_ARG(map, "FokkerPlanckMap: NULL output");
    VFPS_CHECK_ARG(fptype >= 0 && fptype <= 3, "FokkerPlanckMap: FPType");
    VFPS_TRY_BEGIN
    std::vector<uint32_t> ix;
    std::vector<float> w;
    int ip = fp_weights(in->ax1, fptype, e1, deriv, ix, w);
    if (ip < 0) {
        return fail(VFPS_ERR_INVALID,
                    "FokkerPlanckMap: unknown DerivationType");
    }
    const uint32_t n = in->n;
    std::vector<uint32_t> ixt((size_t)ip * n);
    std::vector<float> wt((size_t)ip * n);
    for (uint32_t y = 0; y < n; y++) {
        for (int j = 0; j < ip; j++) {
            ixt[(size_t)j * n + y] = ix[(size_t)y * ip + j];
            wt[(size_t)j * n + y] = w[(size_t)y * ip + j];
        }
    }
    VFPS_CHECK_ARG(fptrack >= 0 && fptrack <= 3, "FokkerPlanckMap: FPTracking");
    auto* m = new vfps_map();
    m->kind = MAP_FP;
    m->in = in;
    m->out = out;
    m->ip = ip;
    m->fptrack = fptrack;
    m->fp_e1 = e1;
    hipError_t e = hipMalloc(&m->fp_w, sizeof(float) * wt.size());
    if (e == hipSuccess) e = hipMalloc(&m->fp_i, sizeof(uint32_t) * ixt.size());
    if (e != hipSuccess) {
        hipFree(m->fp_w);
        delete m;
        return fail(VFPS_ERR_NOMEM, hipGetErrorString(e));
    }
    hipStream_t st = in->dev->stream;
    hipMemcpyAsync(m->fp_w, wt.data(), sizeof(float) * wt.size(),
                   hipMemcpyHostToDevice, st);
    hipMemcpyAsync(m->fp_i, ixt.data(), sizeof(uint32_t) * ixt.size(),
                   hipMemcpyHostToDevice, st);
    VFPS_HIP(hipStreamSynchronize(st));
    *map = m;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_map_create_identity(vfps_mesh* in, vfps_mesh* out, vfps_map** map)
{
    int rc = check_pair(in, out);
    if (rc != VFPS_OK) return rc;
    VFPS_CHECK_ARG(map, "Identity: NULL output");
    VFPS_TRY_BEGIN
    auto* m = new vfps_map();
    m->kind = MAP_IDENTITY;
    m->in = in;
    m->out = out;
    *map = m;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_map_create_wakepotential(vfps_mesh* in, vfps_mesh* out,
                                  vfps_field* field, int interp, int clamp,
                                  vfps_map** map)
{
    (void)clamp;
    VFPS_CHECK_ARG(field, "WakePotentialMap: NULL field");
    VFPS_TRY_BEGIN
    int rc = new_kick(in, out, interp, 1, map);
    if (rc != VFPS_OK) return rc;
    (*map)->field = field;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_map_update(vfps_map* map)
{
    VFPS_CHECK_ARG(map && map->field, "update: not a WakePotentialMap");
    return launch_wake_potential(map->field, map->offset);
}

// DynamicRFKickMap::apply (src/SM/DynamicRFKickMap.cpp:103-111): the RF
// offsets of the next queue entry (computed on the host in the reference's
// float arithmetic, staged through pinned memory), then the kick
static int dynrf_stage(vfps_map* m)
{
    VFPS_CHECK_ARG((size_t)m->mod_next * 2 < m->modulation.size(),
                   "DynamicRFKickMap::apply: more steps than the modulation "
                   "queue holds");
    const float phase = m->modulation[2 * (size_t)m->mod_next];
    const float ampl = m->modulation[2 * (size_t)m->mod_next + 1];
    const vfps_mesh* ps = m->in;
    const std::vector<float> off =
        m->rf_linear
            ? rfkick_offsets_linear_at(ps->ax0, m->rf_angle, ps->qscale,
                                       m->rf_f, m->rf_syncphase, phase, ampl)
            : rfkick_offsets_at(ps->ax0, ps->ax1, m->rf_revolutionpart,
                                m->rf_V, m->rf_f, m->rf_V0, ps->qscale,
                                (float)ps->pscale, phase, ampl);
    hipStream_t st = ps->dev->stream;
    // the previous upload from the staging buffer has finished reading it
    VFPS_HIP(hipEventSynchronize(m->stage_done));
    std::memcpy(m->stage_h, off.data(), sizeof(float) * off.size());
    VFPS_HIP(hipMemcpyAsync(m->offset, m->stage_h, sizeof(float) * off.size(),
                            hipMemcpyHostToDevice, st));
    VFPS_HIP(hipEventRecord(m->stage_done, st));
    m->mod_next++;
    return VFPS_OK;
}

int vfps_map_apply(vfps_map* map)
{
    VFPS_CHECK_ARG(map, "apply: NULL map");
    if (map->dynrf) {
        VFPS_TRY_BEGIN
        int rc = dynrf_stage(map);
        if (rc != VFPS_OK) return rc;
        VFPS_TRY_END
    }
    return launch_map(map, map->in->dev->stream, map->in->data,
                      map->out->data);
}

static int dynrf_create(vfps_mesh* in, vfps_mesh* out, bool linear,
                        float angle, double revolutionpart, float V_RF,
                        float f_RF, float V0, float phasespread,
                        float amplspread, float modampl,
                        double modtimeincrement, uint32_t steps, uint32_t seed,
                        int interp, int clamp, vfps_map** map)
{
    VFPS_CHECK_ARG(in && out && map && steps > 0 && revolutionpart > 0,
                   "DynamicRFKickMap: bad argument");
    VFPS_CHECK_ARG(linear || V_RF != 0, "DynamicRFKickMap: V_RF is zero");
    int rc = vfps_map_create_kick(in, out, interp, clamp, 1, map);
    if (rc != VFPS_OK) return rc;
    vfps_map* m = *map;
    VFPS_TRY_BEGIN
    m->dynrf = true;
    m->rf_linear = linear;
    m->rf_angle = angle;
    m->rf_revolutionpart = (float)revolutionpart;
    m->rf_V = V_RF;
    m->rf_f = f_RF;
    m->rf_V0 = V0;
    m->rf_syncphase = linear ? 0.0f : std::asin(V0 / V_RF);
    m->modulation = dynrf_modulation(steps, seed, m->rf_syncphase,
                                     revolutionpart, phasespread,
                                     amplspread, modampl, modtimeincrement);
    if (hipHostMalloc(&m->stage_h, sizeof(float) * in->n,
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&m->stage_done, hipEventDisableTiming) !=
            hipSuccess) {
        vfps_map_destroy(m);
        *map = nullptr;
        return fail(VFPS_ERR_NOMEM, "DynamicRFKickMap: staging buffer");
    }
    VFPS_HIP(hipEventRecord(m->stage_done, in->dev->stream));
    // the static RF kick of the constructor (RFKickMap::_calcKick(_syncphase))
    const auto off =
        linear ? rfkick_offsets_linear(in->ax0, angle, in->qscale, f_RF)
               : rfkick_offsets(in->ax0, in->ax1, (float)revolutionpart, V_RF,
                                f_RF, V0, in->qscale, (float)in->pscale);
    return upload_offset(m, off);
    VFPS_TRY_END
}

int vfps_map_create_dynrfkick(vfps_mesh* in, vfps_mesh* out,
                              double revolutionpart, float V_RF, float f_RF,
                              float V0, float phasespread, float amplspread,
                              float modampl, double modtimeincrement,
                              uint32_t steps, uint32_t seed, int interp,
                              int clamp, vfps_map** map)
{
    return dynrf_create(in, out, false, 0.0f, revolutionpart, V_RF, f_RF, V0,
                        phasespread, amplspread, modampl, modtimeincrement,
                        steps, seed, interp, clamp, map);
}

int vfps_map_create_dynrfkick_linear(vfps_mesh* in, vfps_mesh* out,
                                     float angle, double revolutionpart,
                                     float f_RF, float phasespread,
                                     float amplspread, float modampl,
                                     double modtimeincrement, uint32_t steps,
                                     uint32_t seed, int interp, int clamp,
                                     vfps_map** map)
{
    return dynrf_create(in, out, true, angle, revolutionpart, 0.0f, f_RF, 0.0f,
                        phasespread, amplspread, modampl, modtimeincrement,
                        steps, seed, interp, clamp, map);
}

int vfps_map_get_modulation(const vfps_map* map, uint32_t first,
                            uint32_t count, float* out, uint32_t* applied)
{
    VFPS_CHECK_ARG(map && map->dynrf, "getPastModulation: not a "
                                      "DynamicRFKickMap");
    VFPS_CHECK_ARG(!count || out, "getPastModulation: NULL output");
    VFPS_CHECK_ARG(((size_t)first + count) * 2 <= map->modulation.size(),
                   "getPastModulation: beyond the queue");
    if (count) {
        std::memcpy(out, map->modulation.data() + 2 * (size_t)first,
                    sizeof(float) * 2 * count);
    }
    if (applied) *applied = map->mod_next;
    return VFPS_OK;
}

int vfps_dynrf_modulation(uint32_t steps, uint32_t seed, float syncphase,
                          double revolutionpart, float phasespread,
                          float amplspread, float modampl,
                          double modtimeincrement, float* out)
{
    VFPS_CHECK_ARG(out && revolutionpart > 0, "vfps_dynrf_modulation: bad "
                                              "argument");
    VFPS_TRY_BEGIN
    const auto q = dynrf_modulation(steps, seed, syncphase, revolutionpart,
                                    phasespread, amplspread, modampl,
                                    modtimeincrement);
    std::memcpy(out, q.data(), sizeof(float) * q.size());
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_map_destroy(vfps_map* map)
{
    if (!map) return VFPS_OK;
    hipStreamSynchronize(map->in->dev->stream);
    hipFree(map->offset);
    hipFree(map->fp_w);
    hipFree(map->fp_i);
    if (map->stage_done) {
        hipEventSynchronize(map->stage_done);
        hipEventDestroy(map->stage_done);
    }
    if (map->stage_h) hipHostFree(map->stage_h);
    delete map;
    return VFPS_OK;
}

}  // extern "C"
