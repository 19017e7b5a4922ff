// Fused per-step kernels shared by the single-mesh step and the batch.
#pragma once

#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "maps.hpp"

namespace vfps_impl {

// One instance (one mesh) as the fused map kernel sees it.
struct InstDesc {
    uint32_t n;
    uint32_t hlo, hhi;            // halo rows needed by the drift/rotation
    float* mesh0;                 // ping-pong state buffers
    float* mesh1;
    float* proj;                  // [n] x projection (written every step)
    const float* ws;              // [n] Simpson weights
    const float* wake_off;        // [n] wake offsets (nullptr: no wake map)
    const float* rf_off;          // [n] RF kick offsets (scheme 0)
    const float* drift_off;       // [n] drift offsets (scheme 0)
    const float* fp_w;            // [ip][n] Fokker-Planck weights
    const uint32_t* fp_i;         // [ip][n] Fokker-Planck indices
    RotGeom rot;                  // scheme 1
    float* prof_hist;             // optional per-step profile history
    uint32_t hist_cap;
    const uint32_t* counter;      // steps counted by the field kernel
    uint32_t* err;                // set if a tap falls outside the halo
    uint32_t seq_proj;            // 1: x projection in the reference's order
    unsigned long long* tstamp;   // optional phase timestamps of workgroup 0
};

// kernel variant key: scheme (0 kick-drift, 1 rotation), interpolation,
// FP taps (0 = identity), wake on/off, clamp
struct MapsVariant {
    int scheme, it, ipfp;
    bool wake, clamp;
};

// Launch the fused maps kernel over `ninst` instances of the same n with
// `tiles` row tiles each (tile height T rows) and `threads` per block.
int launch_fused_maps(const MapsVariant& v, const InstDesc* descs_d,
                      uint32_t ninst, uint32_t n, uint32_t T,
                      uint32_t threads, uint32_t hmax, uint32_t parity,
                      hipStream_t st);
size_t fused_lds_bytes(uint32_t n, uint32_t T, uint32_t hmax);
int fused_set_lds_limit(const MapsVariant& v, size_t bytes);

// Field kernel over `ninst` instances (one workgroup each).
int launch_fields(bool wake, bool csr, const FieldParams* params_d,
                  uint32_t ninst, uint32_t Mmax, hipStream_t st);

// exact halo of a kick along x (drift) with the given offsets
void kick_halo(const std::vector<float>& off, uint32_t n, int it,
               uint32_t* hlo, uint32_t* hhi);
// exact halo of a rotation
void rotation_halo(const RotGeom& g, uint32_t n, int it, uint32_t* hlo,
                   uint32_t* hhi);

}  // namespace vfps_impl
