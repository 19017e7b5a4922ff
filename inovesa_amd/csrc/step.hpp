// Fused per-step kernels shared by the single-mesh step and the batch.
#pragma once

#include <map>
#include <utility>
#include <vector>

#include "common.hpp"
#include "field.hpp"
#include "maps.hpp"
#include "track.hpp"

namespace vfps_impl {

// One instance (one mesh) as the step kernel sees it.
struct InstDesc {
    uint32_t n;
    uint32_t hlo, hhi;            // halo rows needed by the drift/rotation
    uint32_t tiles;               // row tiles of this instance
    uint32_t T;                   // output rows per tile
    uint32_t stride;              // LDS buffer rows (>= T + hlo + hhi)
    float* mesh[2];               // ping-pong state buffers
    float* proj[2];               // x projection of mesh[k]
    const float* ws;              // [n] Simpson weights
    const float* wake_off;        // FFT wake: [n] offsets of the field kernel
    const double* wake_g;         // direct wake: g[d + n - 1], |d| < n
    float* wake_out;              // direct wake: [n] offsets of the step
    float wake_scale;             // wake scaling divided by nmax
    int32_t use_wake, use_csr;
    const double* csr_h;          // direct CSR: h[d + n - 1], |d| < n
    double* csr_part;             // [2][n] CSR shares of the tiles, by parity
    uint32_t* csr_valid;          // [2] step (1-based) whose shares csr_part
                                  // holds, by parity (0xFFFFFFFF: none)
    float* csr_power;             // [1] CSR power of the last step
    double* filling;              // [2] FP64 charge integral (+ total)
    const float* rf_off;          // [n] RF kick offsets (scheme 0)
    const float* drift_off;       // [n] drift offsets (scheme 0)
    const float* fp_w;            // [ip][n] Fokker-Planck weights
    const uint32_t* fp_i;         // [ip][n] Fokker-Planck indices
    RotGeom rot;                  // scheme 1
    float* prof_hist;             // optional per-step profile history
    float* csr_hist;              // optional per-step CSR power history
    uint32_t hist_cap;
    uint32_t* counter;            // [2] completed steps, read at parity p,
                                  // written at p^1 by the workgroup that
                                  // closes the step (direct form) or by
                                  // the field kernel
    uint32_t* err;                // set if a tap falls outside the halo
    uint32_t K;                   // strip form: tiles per strip (0: none)
    uint32_t ringw;               // strip form: rows of the kicked-row ring
    uint32_t seq_proj;            // 1: x projection in the reference's order
    uint32_t norm_every;          // > 0: renormalise the charge before the
                                  // maps of every step whose count is a
                                  // multiple (src/main.cpp:953-955)
    float fill_set;               // the set bunch population
    unsigned long long* tstamp;   // optional timestamps (phases of tile 0,
                                  // entry/exit of every workgroup)
};

// kernel variant key: scheme (0 kick-drift, 1 rotation), interpolation,
// FP taps (0 = identity), direct wake (else the field kernel's offsets),
// clamp
struct StepVariant {
    int scheme, it, ipfp;
    bool direct, clamp;
};

// form: StepForm (step_kernel.hpp); maxc: columns per thread
const void* step_kernel_for(const StepVariant& v, int form, uint32_t maxc);
// one_mesh: the one-mesh form's layout (kernel windows in LDS), else a
// batch's
size_t step_lds_bytes(const StepVariant& v, uint32_t n, uint32_t T,
                      uint32_t stride, bool one_mesh);
int step_set_lds_limit(const StepVariant& v, size_t bytes);
// Launch the step kernel: `nblocks` workgroups of `threads` with
// `lds_bytes` of dynamic LDS (at least step_lds_bytes of every instance);
// blockmap_d (optional) gives (instance << 16 | tile) per logical
// workgroup (STEP_NO_TILE: padding), else instance 0, tile = workgroup.
// maxn: the largest mesh size of the launch (sets the columns per thread).
// One mesh: desc_h (its descriptor, passed to the kernel by value), no
// block map, one workgroup per tile (direct wake: and one more that closes
// the step).  A batch: blockmap_d of nblocks
// entries (a multiple of 8); grid (a multiple of 8, at most nblocks; 0:
// nblocks) workgroups walk it, several tiles each when grid < nblocks.
int launch_step_kernel(const StepVariant& v, const InstDesc* descs_d,
                       const uint32_t* blockmap_d, uint32_t nblocks,
                       size_t lds_bytes, uint32_t threads, uint32_t maxn,
                       uint32_t parity, hipStream_t st,
                       const InstDesc* desc_h = nullptr, uint32_t grid = 0);
// workgroups of this launch shape that stay resident on one CU
int step_occupancy(const StepVariant& v, uint32_t threads, uint32_t maxn,
                   size_t lds_bytes, int* per_cu);
// Strip form (strip_kernel.hpp; kick-drift maps, direct wake, batches): a
// workgroup advances K tiles of one instance.
const void* strip_kernel_for(const StepVariant& v, uint32_t maxc);
size_t strip_lds_bytes(uint32_t n, uint32_t T, uint32_t K, uint32_t halo);
uint32_t strip_ring(uint32_t T, uint32_t halo);
int strip_set_lds_limit(const StepVariant& v, size_t bytes);
// nblocks (a multiple of 8) workgroups, one per block-map entry
// (instance << 16 | strip)
int launch_strip_kernel(const StepVariant& v, const InstDesc* descs_d,
                        const uint32_t* blockmap_d, uint32_t nblocks,
                        size_t lds_bytes, uint32_t threads, uint32_t maxn,
                        uint32_t parity, hipStream_t st);
// FP64 Toeplitz kernels of the direct wake (g) and CSR power (h) forms
void toeplitz_tables(uint32_t n, uint32_t M, const float2* Z,
                     const float2* Zc, const float* renorm, float fdelta,
                     std::vector<double>* g, std::vector<double>* h);

// Field kernel over `ninst` instances (one workgroup each).  params_h:
// the host copy of the parameters of a single instance, passed by value,
// or null to read them from params_d.
int launch_fields(bool wake, bool csr, const FieldParams* params_d,
                  uint32_t ninst, uint32_t Mmax, uint32_t parity,
                  hipStream_t st, const FieldParams* params_h = nullptr);
// close a run: with close_csr, the last step's CSR power from the tiles'
// shares (direct form; the next step's closing workgroup does this
// otherwise); with err_host, every instance's error flag into that pinned
// host array, so the host reads it after the stream's last wait
int launch_step_finalize(const InstDesc* descs_d, uint32_t ninst,
                         uint32_t parity, hipStream_t st,
                         uint32_t* err_host = nullptr, bool close_csr = true);

// exact halo of a kick along x (drift) with the given offsets
void kick_halo(const std::vector<float>& off, uint32_t n, int it,
               uint32_t* hlo, uint32_t* hhi);
// exact halo of a rotation
void rotation_halo(const RotGeom& g, uint32_t n, int it, uint32_t* hlo,
                   uint32_t* hhi);

}  // namespace vfps_impl

// A whole step on one mesh (the object behind vfps_step_*).
struct vfps_step {
    vfps_dev* dev = nullptr;
    vfps_step_desc desc{};
    vfps_mesh *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
    vfps_map *wm = nullptr, *rfm = nullptr, *drm = nullptr, *rot = nullptr,
             *fpm = nullptr;
    vfps_field* wfield = nullptr;
    vfps_impedance* zwake = nullptr;
    vfps_impedance* zcsr = nullptr;
    float* csr_renorm = nullptr;
    double2* csr_tw = nullptr;
    float csr_fdelta = 0;
    double* filling = nullptr;  // [2]
    uint32_t* counter = nullptr;    // [2] (ping-pong in the fused step)
    uint32_t* err = nullptr;
    uint32_t* err_h = nullptr;  // pinned host copy, read after each run
    float* csr_hist = nullptr;
    float* prof_hist = nullptr;
    uint32_t hist_cap = 0;
    // direct wake / CSR forms
    double* wake_g = nullptr;   // [2n-1]
    double* csr_h = nullptr;    // [2n-1]
    double* csr_part = nullptr; // [2][n]
    uint32_t* csr_valid = nullptr;  // [2]
    float* csr_power = nullptr;
    vfps_impl::InstDesc* inst_d = nullptr;
    vfps_impl::InstDesc inst_h{};     // host copy, passed by value at launch
    vfps_impl::FieldParams* fparams_d = nullptr;
    vfps_impl::FieldParams fparams_h{};
    bool desc_dirty = true;           // inst_h/fparams_h not yet on the device
    bool halo_valid = false;          // halo rows of the drift/rotation
    uint32_t halo_lo = 0, halo_hi = 0;
    vfps_impl::StepVariant variant{};
    uint32_t T = 1, threads = 256, stride = 1, hmax = 0, M = 0;
    bool persist = false;  // in a batch: persistent launch (grid < tiles)
    uint32_t strip = 0;    // in a batch: tiles per strip (strip form; 0: a
                           // workgroup per tile)
    bool in_batch = false; // a batch member (the batch forms' LDS layout)
    bool fused = true;
    bool direct = true;
    // passive particle tracking through every step's maps
    float2* particles = nullptr;
    uint32_t npart = 0;
    vfps_impl::TrackProgram track{};
    uint32_t parity = 0;  // fused path: buffer holding the state (0: t1)
    // captured runs of k steps from parity p, key k << 1 | p
    std::map<uint32_t, std::pair<hipGraph_t, hipGraphExec_t>> graphs;

    uint32_t tiles() const { return vfps_impl::ceil_div(desc.n, T); }
    uint32_t strips() const { return vfps_impl::ceil_div(desc.n, T * strip); }
};

namespace vfps_impl {
// reset: forget pending CSR shares afterwards (step_reset_pending); a
// caller that restores the whole state itself passes false
int step_retile(vfps_step* s, uint32_t T, uint32_t threads, bool reset = true);
// LDS bytes of the step's tile in the form it runs in
size_t step_lds_of(const vfps_step* s);
}  // namespace vfps_impl

namespace vfps_impl {
// forget pending CSR shares and make both counter entries the current
// count (after a retile, a new mesh or a new history), in stream order
int step_reset_pending(vfps_step* s);
}  // namespace vfps_impl
