// ElectricField kernel parameters, shared with the fused step.
#pragma once

#include <vector>

#include "common.hpp"

namespace vfps_impl {

struct FieldParams {
    uint32_t nb, n, M, spacing;
    const uint32_t* buckets;
    float scaling;            // wake scaling already divided by nmax
    const float* proj;        // [nb][n] x projection
    const float* proj_alt;    // x projection read at odd parity (fused step)
    const float* ws;          // Simpson weights (for the FP64 integral)
    const float2* Z;          // wake impedance [2M]
    const double2* tw_m;      // e^{-2 pi i t/M}
    const double2* tw_n;      // e^{-2 pi i k/2M}
    float* wake;              // [nb][n]
    float* wake2;             // optional copy (a kick map's offsets)
    float* padded;            // optional [2M] padded profile
    float* wake_padded;       // optional [2M] unscaled C2R output
                              // (getPaddedWakePotential)
    double* filling;          // optional [nb+1] FP64 integral (+ total)
    const float2* Zc;         // CSR impedance [2M]
    const float* renorm;      // CSR renormalisation [2M]
    float fdelta;             // frequency step of the CSR sum
    float* spectrum;          // optional [nb][2M]
    float* power;             // optional [nb]
    uint32_t* counter;        // optional step counter (history slot)
    uint32_t counter_pingpong;  // 1: read counter[parity], write [parity^1]
    uint32_t parity;          // set by the fused step's field kernel
    float* csr_hist;          // optional per-step CSR power history
    uint32_t hist_cap;        // history capacity (steps)
    unsigned long long* tstamp;  // optional phase timestamps (block 0)
};

FieldParams field_params(vfps_field* f);
size_t field_lds_bytes(uint32_t M);
int field_set_lds_limit(size_t bytes);
int launch_field(const FieldParams& p, bool wake, bool csr, uint32_t blocks,
                 hipStream_t st);
void make_twiddles(uint32_t M, std::vector<double2>& tw_m,
                   std::vector<double2>& tw_n);
std::vector<float> csr_renorm(uint32_t nmax, float delta0, double qscale,
                              float fc, float* fdelta);

}  // namespace vfps_impl
