// Map internals shared by the per-map kernels and the fused step kernels.
#pragma once

#include <vector>

#include "common.hpp"

namespace vfps_impl {

// Geometry of a RotationMap: axes (min, delta, zerobin) and cos/sin of
// the backward rotation angle.
struct RotGeom {
    float qmin, d0, zb0;
    float pmin, d1, zb1;
    float cs, sn;
};

RotGeom rot_geom(const vfps_map* m);

// One output cell of RotationMap::apply with the taps of genHInfo
// (src/SM/RotationMap.cpp:190-262) and, if CLAMP, the CPU path's clamp to
// the inner 2x2 taps (src/SM/RotationMap.cpp:159-172).  src(ok, i0, j0)
// returns the source sample of tap (i0, j0); for taps out of range (ok ==
// false) it must return in[0], the sample the reference reads there.
// rows(i0) returns the IT consecutive source rows from i0 on, `stride`
// samples apart, if they are all at hand, else null: a stencil wholly
// inside the mesh and those rows reads its taps without per-tap checks,
// with the same products in the same order.
template <int IT, bool CLAMP, typename Src, typename Rows>
__device__ __forceinline__ float rotate_cell_src(uint32_t n, const RotGeom& g,
                                                 uint32_t x0, uint32_t y0,
                                                 Src src, Rows rows,
                                                 uint32_t stride)
{
    const float at0 = g.qmin + ((float)x0 * g.d0);
    const float at1 = g.pmin + ((float)y0 * g.d1);
    const float x1r = (g.cs * at0 - g.sn * at1) / g.d0 + g.zb0;
    const float y1r = (g.sn * at0 + g.cs * at1) / g.d1 + g.zb1;
    float xi, yi;
    const float xf = vf_modf(x1r, &xi);
    const float yf = vf_modf(y1r, &yi);
    const uint32_t x1 = vf_f2u(xi), y1 = vf_f2u(yi);
    float v = 0.0f;
    float s11 = 0.0f, s12 = 0.0f, s21 = 0.0f, s22 = 0.0f;
    constexpr uint32_t LO = (uint32_t)((IT - 1) / 2), HI = (uint32_t)(IT / 2);
    const float* r0 = nullptr;
    // (written without x1 + HI: a source position in [-2, -1) casts to
    // 0xFFFFFFFF, as the reference's cast does, and the sum would wrap)
    if (n > HI && x1 >= LO && x1 < n - HI && y1 >= LO && y1 < n - HI) {
        r0 = rows(x1 - LO);
    }
    if (r0) {
        float icq[4], icp[4];
        coefficients<IT>(xf, icq);
        coefficients<IT>(yf, icp);
        const float* p = r0 + (y1 - LO);
#pragma unroll
        for (int i1 = 0; i1 < IT; i1++) {
#pragma unroll
            for (int j1 = 0; j1 < IT; j1++) {
                const float w = icq[i1] * icp[j1];
                const float s = p[(size_t)i1 * stride + (uint32_t)j1];
                v = v + s * w;
                if constexpr (CLAMP) {
                    if (i1 == 1 && j1 == 1) s11 = s;
                    if (i1 == 1 && j1 == 2) s12 = s;
                    if (i1 == 2 && j1 == 1) s21 = s;
                    if (i1 == 2 && j1 == 2) s22 = s;
                }
            }
        }
    } else if (x1 < n && y1 < n) {
        float icq[4], icp[4];
        coefficients<IT>(xf, icq);
        coefficients<IT>(yf, icp);
#pragma unroll
        for (int i1 = 0; i1 < IT; i1++) {
            const uint32_t i0 = x1 + (uint32_t)i1 - (uint32_t)((IT - 1) / 2);
#pragma unroll
            for (int j1 = 0; j1 < IT; j1++) {
                const uint32_t j0 = y1 + (uint32_t)j1 - (uint32_t)((IT - 1) / 2);
                const bool ok = i0 < n && j0 < n;
                const float w = ok ? icq[i1] * icp[j1] : 0.0f;
                const float s = src(ok, i0, j0);
                v = v + s * w;
                if constexpr (CLAMP) {
                    if (i1 == 1 && j1 == 1) s11 = s;
                    if (i1 == 1 && j1 == 2) s12 = s;
                    if (i1 == 2 && j1 == 1) s21 = s;
                    if (i1 == 2 && j1 == 2) s22 = s;
                }
            }
        }
    } else {
        const float s = src(false, 0u, 0u);
#pragma unroll
        for (int j = 0; j < IT * IT; j++) {
            v = v + s * 0.0f;
        }
        s11 = s12 = s21 = s22 = s;
    }
    if constexpr (CLAMP) {
        float ceil_ = __FLT_MIN__;  // numeric_limits<float>::min()
        float flor = __FLT_MAX__;
        ceil_ = std_max(ceil_, s11);
        flor = std_min(flor, s11);
        ceil_ = std_max(ceil_, s12);
        flor = std_min(flor, s12);
        ceil_ = std_max(ceil_, s21);
        flor = std_min(flor, s21);
        ceil_ = std_max(ceil_, s22);
        flor = std_min(flor, s22);
        v = std_max(std_min(ceil_, v), flor);
    }
    return v;
}

template <int IT, bool CLAMP>
__device__ __forceinline__ float rotate_cell(uint32_t n, const RotGeom& g,
                                             uint32_t x0, uint32_t y0,
                                             const float* __restrict__ in)
{
    return rotate_cell_src<IT, CLAMP>(
        n, g, x0, y0,
        [&](bool ok, uint32_t i0, uint32_t j0) {
            return in[ok ? i0 * n + j0 : 0u];
        },
        [&](uint32_t i0) { return in + (size_t)i0 * n; }, n);
}

// host restatement of the source position of a rotated cell (the same
// float operations as rotate_cell_src); returns x1 as the reference's
// uint32 cast computes it
inline uint32_t rotation_source_row(const RotGeom& g, uint32_t x0, uint32_t y0,
                                    uint32_t* y1out)
{
    const float at0 = g.qmin + ((float)x0 * g.d0);
    const float at1 = g.pmin + ((float)y0 * g.d1);
    const float x1r = (g.cs * at0 - g.sn * at1) / g.d0 + g.zb0;
    const float y1r = (g.sn * at0 + g.cs * at1) / g.d1 + g.zb1;
    float xi, yi;
    vf_modf(x1r, &xi);
    vf_modf(y1r, &yi);
    *y1out = vf_f2u(yi);
    return vf_f2u(xi);
}

// apply() of any map from `in` to `out` on stream st
int launch_map(const vfps_map* m, hipStream_t st, const float* in, float* out);

// WakePotentialMap::update: wake potential of the field into dst (nb*n)
int launch_wake_potential(vfps_field* f, float* dst);

// updateXProjection / integrate on a mesh (runtime.hip)
int launch_project_x(vfps_mesh* m, const float* src, float* proj);
int launch_integrate(vfps_mesh* m);

// constructor tables
std::vector<float> drift_offsets(const Ruler& ax0, const Ruler& ax1,
                                 const float* slip, uint32_t nslip,
                                 float pscale, float E0);
std::vector<float> rfkick_offsets_linear(const Ruler& ax0, float angle,
                                         double qscale, float f_RF);
std::vector<float> rfkick_offsets(const Ruler& ax0, const Ruler& ax1,
                                  float revolutionpart, float V_RF,
                                  float f_RF, float V0, double qscale,
                                  float pscale);
int fp_weights(const Ruler& ax1, int fptype, float e1, int deriv,
               std::vector<uint32_t>& ix, std::vector<float>& w);

}  // namespace vfps_impl
