// The time step of src/main.cpp:948-1053 on the device.
//
// Fused path, two kernels per step:
//   fields_kernel: x projection -> FFT wake potential (+ CSR power, FP64
//                  charge integral, step counter), one workgroup per mesh;
//   maps_kernel:   wake kick -> RF kick -> drift -> Fokker-Planck (scheme 0)
//                  or wake kick -> rotation -> Fokker-Planck (scheme 1) on a
//                  tile of T output rows held in LDS with the halo rows the
//                  drift/rotation reads, then the x projection of the tile.
// The mesh makes one trip HBM -> LDS -> HBM per step (8 bytes per cell plus
// halo re-reads).  Each map inside keeps the reference's arithmetic, so the
// fused step equals the per-map path bit for bit except for the summation
// order of the x projection.
//
// Per-map path (fused = 0): one kernel per SourceMap::apply, as the
// reference calls them.  Both paths are captured into a hipGraph of an
// even number of steps (the fused path ping-pongs two buffers).
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <memory>

#include "common.hpp"
#include "field_body.hpp"
#include "reduce.hpp"
#include "step.hpp"

namespace vfps_impl {

// ---- fused maps kernel -------------------------------------------------------

template <int IT>
__device__ __forceinline__ void row_kick(const float* __restrict__ src,
                                         float* __restrict__ dst, uint32_t n,
                                         float off, uint32_t lane)
{
    uint32_t idx[IT];
    float w[IT];
    kick_taps<IT>(off, n, idx, w);
    const int32_t half = (int32_t)(n / 2);
    for (uint32_t y = lane; y < n; y += 64) {
        float v = 0.0f;
#pragma unroll
        for (int j = 0; j < IT; j++) {
            const uint32_t ys = (uint32_t)((int32_t)(y + idx[j]) - half);
            if (ys < n) {
                v = v + src[ys] * w[j];
            }
        }
        dst[y] = v;
    }
}

// most mesh columns one thread handles (n <= 1024 with >= 256 threads)
constexpr int MAPS_MAX_COLS = 4;
// float4 loads one thread keeps in flight while filling the tile
constexpr int MAPS_LOAD_BATCH = 8;

#define VFPS_STAMP(k)                                                         \
    do {                                                                      \
        if (d.tstamp && blockIdx.x == 0) {                                    \
            __syncthreads();                                                  \
            if (threadIdx.x == 0) d.tstamp[k] = wall_clock64();               \
        }                                                                     \
    } while (0)

template <int IT, int IPFP, bool WAKE, int SCHEME, bool CLAMP>
__global__ void __launch_bounds__(512)
maps_kernel(const InstDesc* __restrict__ descs, uint32_t tiles_per_inst,
            uint32_t T, uint32_t hmax, uint32_t parity)
{
    extern __shared__ __align__(16) float lds[];
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs,
    // so give each XCD a contiguous run of tiles (neighbouring tiles share
    // halo rows, which then hit in that XCD's L2).
    const uint32_t nblocks = gridDim.x;
    uint32_t logical = blockIdx.x;
    if ((nblocks & 7u) == 0u) {
        logical = (blockIdx.x & 7u) * (nblocks >> 3) + (blockIdx.x >> 3);
    }
    const InstDesc& d = descs[logical / tiles_per_inst];
    const uint32_t tile = logical % tiles_per_inst;
    const uint32_t n = d.n;
    const int x0 = (int)(tile * T);
    if (x0 >= (int)n) {
        return;
    }
    VFPS_STAMP(0);
    const int rows_out = min((int)T, (int)n - x0);
    const int hlo = (int)d.hlo;
    const int rbase = x0 - hlo;
    const int R = rows_out + hlo + (int)d.hhi;
    const int stride = (int)(T + 2 * hmax);  // rows per LDS buffer
    const uint32_t tstride = n + 4;           // padded row of the output tile
    float* bufA = lds;
    float* bufB = bufA + (size_t)stride * n;
    float* tile_out = bufB + (size_t)stride * n;  // [T][n+4] products
    float* offw = tile_out + (size_t)T * tstride;  // [stride] wake offsets
    float* offr = offw + stride;                   // [stride] RF offsets
    const float* in = parity ? d.mesh1 : d.mesh0;
    float* out = parity ? d.mesh0 : d.mesh1;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6, nw = bs >> 6;
    const int32_t half = (int32_t)(n / 2);

    // phase 1: every global read the block needs, issued together
    //  - per-row kick offsets of the tile's rows
    if ((int)tid < R) {
        const int r = rbase + (int)tid;
        if (r >= 0 && r < (int)n) {
            if (WAKE) offw[tid] = d.wake_off[r];
            if (SCHEME == 0) offr[tid] = d.rf_off[r];
        }
    }
    //  - per-column tables (drift offset, Fokker-Planck stencil, Simpson
    //    weight) of this thread's columns, kept in registers
    const int ncols = (int)((n + bs - 1) / bs);
    float col_drift[MAPS_MAX_COLS], col_ws[MAPS_MAX_COLS];
    uint32_t col_fi[MAPS_MAX_COLS][IPFP > 0 ? IPFP : 1];
    float col_fw[MAPS_MAX_COLS][IPFP > 0 ? IPFP : 1];
#pragma unroll
    for (int c = 0; c < MAPS_MAX_COLS; c++) {
        const uint32_t y = tid + (uint32_t)c * bs;
        if (c < ncols && y < n) {
            if (SCHEME == 0) col_drift[c] = d.drift_off[y];
            col_ws[c] = d.ws[y];
#pragma unroll
            for (int k = 0; k < IPFP; k++) {
                col_fi[c][k] = d.fp_i[(size_t)k * n + y];
                col_fw[c][k] = d.fp_w[(size_t)k * n + y];
            }
        }
    }
    //  - rows [rbase, rbase+R) of the input, MAPS_LOAD_BATCH float4 loads
    //    in flight per thread
    if ((n & 3u) == 0u) {
        const uint32_t n4 = n >> 2;
        const float4* in4 = reinterpret_cast<const float4*>(in);
        float4* a4 = reinterpret_cast<float4*>(bufA);
        const uint32_t total = (uint32_t)R * n4;
        for (uint32_t base = 0; base < total; base += bs * MAPS_LOAD_BATCH) {
            float4 v[MAPS_LOAD_BATCH];
#pragma unroll
            for (int b = 0; b < MAPS_LOAD_BATCH; b++) {
                const uint32_t i = base + tid + (uint32_t)b * bs;
                const int r = rbase + (int)(i / n4);
                if (i < total && r >= 0 && r < (int)n) {
                    v[b] = in4[(size_t)r * n4 + (i % n4)];
                }
            }
#pragma unroll
            for (int b = 0; b < MAPS_LOAD_BATCH; b++) {
                const uint32_t i = base + tid + (uint32_t)b * bs;
                const int r = rbase + (int)(i / n4);
                if (i < total && r >= 0 && r < (int)n) {
                    a4[i] = v[b];
                }
            }
        }
    } else {
        for (uint32_t i = tid; i < (uint32_t)R * n; i += bs) {
            const int r = rbase + (int)(i / n);
            if (r >= 0 && r < (int)n) {
                bufA[i] = in[(size_t)r * n + (i % n)];
            }
        }
    }
    __syncthreads();
    VFPS_STAMP(1);
    float* cur = bufA;
    float* oth = bufB;
    // phase 2: wake kick along y (WakePotentialMap::apply), row by row
    if (WAKE) {
        for (int row = (int)wave; row < R; row += (int)nw) {
            const int r = rbase + row;
            if (r < 0 || r >= (int)n) {
                continue;
            }
            row_kick<IT>(cur + (size_t)row * n, oth + (size_t)row * n, n,
                         offw[row], lane);
        }
        __syncthreads();
        float* t = cur;
        cur = oth;
        oth = t;
    }
    VFPS_STAMP(2);
    if constexpr (SCHEME == 0) {
        // phase 3: RF kick along y (RFKickMap::apply)
        for (int row = (int)wave; row < R; row += (int)nw) {
            const int r = rbase + row;
            if (r < 0 || r >= (int)n) {
                continue;
            }
            row_kick<IT>(cur + (size_t)row * n, oth + (size_t)row * n, n,
                         offr[row], lane);
        }
        __syncthreads();
        float* t = cur;
        cur = oth;
        oth = t;
        VFPS_STAMP(3);
        // phase 4: drift along x (DriftMap::apply), one column per thread
#pragma unroll
        for (int c = 0; c < MAPS_MAX_COLS; c++) {
            const uint32_t y = tid + (uint32_t)c * bs;
            if (c >= ncols || y >= n) {
                continue;
            }
            uint32_t idx[IT];
            float w[IT];
            kick_taps<IT>(col_drift[c], n, idx, w);
            for (int orow = 0; orow < rows_out; orow++) {
                const uint32_t x = (uint32_t)(x0 + orow);
                float v = 0.0f;
#pragma unroll
                for (int j = 0; j < IT; j++) {
                    const uint32_t xs = (uint32_t)((int32_t)(x + idx[j]) - half);
                    if (xs < n) {
                        int lr = (int)xs - rbase;
                        if (lr < 0 || lr >= R) {
                            atomicOr(d.err, 1u);
                            lr = lr < 0 ? 0 : R - 1;
                        }
                        v = v + cur[(size_t)lr * n + y] * w[j];
                    }
                }
                oth[(size_t)orow * n + y] = v;
            }
        }
    } else {
        // phase 3: rotation (RotationMap::apply).  Taps out of range read
        // sample 0 of the (wake-kicked) input, as the reference does.
        float s00;
        if (rbase <= 0) {
            s00 = cur[(size_t)(-rbase) * n];
        } else if (WAKE) {
            uint32_t idx[IT];
            float w[IT];
            kick_taps<IT>(d.wake_off[0], n, idx, w);
            float v = 0.0f;
#pragma unroll
            for (int j = 0; j < IT; j++) {
                const uint32_t ys = (uint32_t)((int32_t)idx[j] - half);
                if (ys < n) {
                    v = v + in[ys] * w[j];
                }
            }
            s00 = v;
        } else {
            s00 = in[0];
        }
        VFPS_STAMP(3);
        const RotGeom g = d.rot;
        const float* src = cur;
        for (uint32_t i = tid; i < (uint32_t)rows_out * n; i += bs) {
            const uint32_t orow = i / n, y0 = i % n;
            const float v = rotate_cell_src<IT, CLAMP>(
                n, g, (uint32_t)x0 + orow, y0,
                [&](bool ok, uint32_t i0, uint32_t j0) -> float {
                    if (!ok) {
                        return s00;
                    }
                    int lr = (int)i0 - rbase;
                    if (lr < 0 || lr >= R) {
                        atomicOr(d.err, 2u);
                        lr = lr < 0 ? 0 : R - 1;
                    }
                    return src[(size_t)lr * n + j0];
                });
            oth[i] = v;
        }
    }
    __syncthreads();
    VFPS_STAMP(4);
    // phase 5: Fokker-Planck along y (FokkerPlanckMap::apply), the output
    // store and the products mesh*weight of the x projection, one column
    // per thread with the column's stencil in registers
#pragma unroll
    for (int c = 0; c < MAPS_MAX_COLS; c++) {
        const uint32_t y = tid + (uint32_t)c * bs;
        if (c >= ncols || y >= n) {
            continue;
        }
        for (int orow = 0; orow < rows_out; orow++) {
            const float* src = oth + (size_t)orow * n;
            float v;
            if constexpr (IPFP == 0) {
                v = src[y];
            } else {
                v = 0.0f;
#pragma unroll
                for (int k = 0; k < IPFP; k++) {
                    v = v + src[col_fi[c][k]] * col_fw[c][k];
                }
            }
            out[(size_t)(x0 + orow) * n + y] = v;
            tile_out[(size_t)orow * tstride + y] = v * col_ws[c];
        }
    }
    uint32_t slot = 0xFFFFFFFFu;
    if (d.prof_hist) {
        const uint32_t cnt = *d.counter;
        if (cnt >= 1 && cnt - 1 < d.hist_cap) {
            slot = cnt - 1;
        }
    }
    __syncthreads();
    VFPS_STAMP(5);
    // phase 6: the x projection (PhaseSpace::updateXProjection)
    if (d.seq_proj) {
        // the reference's std::inner_product order: y = 0..n-1, one thread
        // per row, loads batched ahead of the dependent adds
        if ((int)tid < rows_out) {
            const float* r = tile_out + (size_t)tid * tstride;
            float acc = 0.0f;
            uint32_t y = 0;
            if ((n & 3u) == 0u) {
                const float4* r4 = reinterpret_cast<const float4*>(r);
                for (; y + 16 <= n; y += 16) {
                    float4 q[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        q[k] = r4[(y >> 2) + k];
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        acc = acc + q[k].x;
                        acc = acc + q[k].y;
                        acc = acc + q[k].z;
                        acc = acc + q[k].w;
                    }
                }
            }
            for (; y < n; y++) {
                acc = acc + r[y];
            }
            const uint32_t x = (uint32_t)x0 + tid;
            d.proj[x] = acc;
            if (slot != 0xFFFFFFFFu) {
                d.prof_hist[(size_t)slot * n + x] = acc;
            }
        }
    } else {
        for (int orow = (int)wave; orow < rows_out; orow += (int)nw) {
            const float* r = tile_out + (size_t)orow * tstride;
            float acc = 0.0f;
            for (uint32_t y = lane; y < n; y += 64) {
                acc = acc + r[y];
            }
            acc = wave_sum(acc);
            if (lane == 0) {
                const uint32_t x = (uint32_t)(x0 + orow);
                d.proj[x] = acc;
                if (slot != 0xFFFFFFFFu) {
                    d.prof_hist[(size_t)slot * n + x] = acc;
                }
            }
        }
    }
    VFPS_STAMP(6);
}

template <int IT, int IPFP, bool WAKE, int SCHEME, bool CLAMP>
static const void* maps_fn()
{
    return reinterpret_cast<const void*>(&maps_kernel<IT, IPFP, WAKE, SCHEME, CLAMP>);
}

// resolve a variant to its kernel (cubic and linear interpolation)
static const void* maps_kernel_for(const MapsVariant& v)
{
#define VFPS_SEL(IT, IP)                                                      \
    if (v.scheme == 0) {                                                      \
        return v.wake ? maps_fn<IT, IP, true, 0, false>()                     \
                      : maps_fn<IT, IP, false, 0, false>();                   \
    }                                                                         \
    if (v.clamp) {                                                            \
        return v.wake ? maps_fn<IT, IP, true, 1, true>()                      \
                      : maps_fn<IT, IP, false, 1, true>();                    \
    }                                                                         \
    return v.wake ? maps_fn<IT, IP, true, 1, false>()                         \
                  : maps_fn<IT, IP, false, 1, false>();
    if (v.it == 4) {
        switch (v.ipfp) {
        case 0: { VFPS_SEL(4, 0) }
        case 3: { VFPS_SEL(4, 3) }
        case 4: { VFPS_SEL(4, 4) }
        case 5: { VFPS_SEL(4, 5) }
        }
    } else if (v.it == 2 && !v.clamp) {
        switch (v.ipfp) {
        case 0: { VFPS_SEL(2, 0) }
        case 3: { VFPS_SEL(2, 3) }
        case 4: { VFPS_SEL(2, 4) }
        case 5: { VFPS_SEL(2, 5) }
        }
    }
#undef VFPS_SEL
    return nullptr;
}

size_t fused_lds_bytes(uint32_t n, uint32_t T, uint32_t hmax)
{
    const size_t stride = T + 2 * hmax;
    return sizeof(float) * (2 * stride * n + (size_t)T * (n + 4) + 2 * stride);
}

int fused_set_lds_limit(const MapsVariant& v, size_t bytes)
{
    const void* fn = maps_kernel_for(v);
    if (!fn) {
        return fail(VFPS_ERR_UNSUPPORTED, "fused step: no kernel for this "
                                          "interpolation/stencil combination");
    }
    if (bytes > 65536) {
        VFPS_HIP(hipFuncSetAttribute(fn,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes));
    }
    return VFPS_OK;
}

int launch_fused_maps(const MapsVariant& v, const InstDesc* descs_d,
                      uint32_t ninst, uint32_t n, uint32_t T, uint32_t threads,
                      uint32_t hmax, uint32_t parity, hipStream_t st)
{
    const void* fn = maps_kernel_for(v);
    if (!fn) {
        return fail(VFPS_ERR_UNSUPPORTED, "fused step: no kernel for this "
                                          "interpolation/stencil combination");
    }
    uint32_t tiles = ceil_div(n, T);
    void* args[] = {(void*)&descs_d, (void*)&tiles, (void*)&T, (void*)&hmax,
                    (void*)&parity};
    VFPS_HIP(hipLaunchKernel(fn, dim3(tiles * ninst), dim3(threads), args,
                             fused_lds_bytes(n, T, hmax), st));
    return VFPS_OK;
}

// ---- fields kernel (one workgroup per instance) ------------------------------

template <bool WAKE, bool CSR>
__global__ void __launch_bounds__(1024)
fields_kernel(const FieldParams* __restrict__ params)
{
    extern __shared__ double2 flds[];
    const FieldParams p = params[blockIdx.x];
    field_body<WAKE, CSR>(p, 0, flds);
}

int launch_fields(bool wake, bool csr, const FieldParams* params_d,
                  uint32_t ninst, uint32_t Mmax, hipStream_t st)
{
    const size_t lds = field_lds_bytes(Mmax);
    const void* fn;
    if (wake && csr) {
        fn = reinterpret_cast<const void*>(&fields_kernel<true, true>);
    } else if (wake) {
        fn = reinterpret_cast<const void*>(&fields_kernel<true, false>);
    } else if (csr) {
        fn = reinterpret_cast<const void*>(&fields_kernel<false, true>);
    } else {
        fn = reinterpret_cast<const void*>(&fields_kernel<false, false>);
    }
    if (lds > 65536) {
        VFPS_HIP(hipFuncSetAttribute(fn,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    }
    void* args[] = {(void*)&params_d};
    VFPS_HIP(hipLaunchKernel(fn, dim3(ninst), dim3(FFT_THREADS), args, lds, st));
    return VFPS_OK;
}

// ---- halos ----------------------------------------------------------------------

void kick_halo(const std::vector<float>& off, uint32_t n, int it,
               uint32_t* hlo, uint32_t* hhi)
{
    int lo = 0, hi = 0;
    const int half = (int)(n / 2);
    for (uint32_t y = 0; y < n && y < off.size(); y++) {
        uint32_t idx[4];
        float w[4];
        switch (it) {
        case 1: kick_taps<1>(off[y], n, idx, w); break;
        case 2: kick_taps<2>(off[y], n, idx, w); break;
        case 3: kick_taps<3>(off[y], n, idx, w); break;
        default: kick_taps<4>(off[y], n, idx, w); break;
        }
        for (int j = 0; j < it; j++) {
            const int rel = (int)idx[j] - half;
            lo = std::max(lo, -rel);
            hi = std::max(hi, rel);
        }
    }
    *hlo = (uint32_t)lo;
    *hhi = (uint32_t)hi;
}

void rotation_halo(const RotGeom& g, uint32_t n, int it, uint32_t* hlo,
                   uint32_t* hhi)
{
    int lo = 0, hi = 0;
    for (uint32_t x0 = 0; x0 < n; x0++) {
        for (uint32_t y0 = 0; y0 < n; y0++) {
            uint32_t y1;
            const uint32_t x1 = rotation_source_row(g, x0, y0, &y1);
            if (!(x1 < n && y1 < n)) {
                continue;
            }
            for (int i1 = 0; i1 < it; i1++) {
                const uint32_t i0 = x1 + (uint32_t)i1 - (uint32_t)((it - 1) / 2);
                if (i0 < n) {
                    lo = std::max(lo, (int)x0 - (int)i0);
                    hi = std::max(hi, (int)i0 - (int)x0);
                }
            }
        }
    }
    *hlo = (uint32_t)lo;
    *hhi = (uint32_t)hi;
}

// ---- record the profile on the per-map path --------------------------------------
__global__ void record_kernel(uint32_t n, const float* proj, float* hist,
                              const uint32_t* counter, uint32_t cap)
{
    const uint32_t c = *counter;
    if (c == 0 || c - 1 >= cap) {
        return;
    }
    for (uint32_t x = threadIdx.x; x < n; x += blockDim.x) {
        hist[(size_t)(c - 1) * n + x] = proj[x];
    }
}

}  // namespace vfps_impl

using namespace vfps_impl;

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
    uint32_t* counter = nullptr;
    uint32_t* err = nullptr;
    float* csr_hist = nullptr;
    float* prof_hist = nullptr;
    uint32_t hist_cap = 0;
    InstDesc* inst_d = nullptr;
    FieldParams* fparams_d = nullptr;
    MapsVariant variant{};
    uint32_t T = 8, threads = 256, hmax = 0, M = 0;
    bool fused = true;
    uint32_t parity = 0;  // fused path: buffer holding the state (0: t1)
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    uint32_t graph_steps = 0;
};

static int step_launch_one(vfps_step* s, uint32_t parity)
{
    hipStream_t st = s->dev->stream;
    const bool wake = s->desc.use_wake != 0;
    const bool csr = s->desc.use_csr != 0;
    if (s->fused) {
        int rc = launch_fields(wake, csr, s->fparams_d, 1, s->M, st);
        if (rc != VFPS_OK) return rc;
        return launch_fused_maps(s->variant, s->inst_d, 1, s->desc.n, s->T,
                                 s->threads, s->hmax, parity, st);
    }
    // per-map path, the loop body of src/main.cpp:948-1053
    int rc = launch_fields(wake, csr, s->fparams_d, 1, s->M, st);
    if (rc != VFPS_OK) return rc;
    vfps_mesh *t1 = s->t1, *t2 = s->t2, *t3 = s->t3;
    if (wake) {
        rc = launch_map(s->wm, st, t1->data, t2->data);
    } else {
        VFPS_HIP(hipMemcpyAsync(t2->data, t1->data, sizeof(float) * t1->cells(),
                                hipMemcpyDeviceToDevice, st));
    }
    if (rc != VFPS_OK) return rc;
    if (s->desc.scheme == 0) {
        if ((rc = launch_map(s->rfm, st, t2->data, t1->data)) != VFPS_OK) return rc;
        if ((rc = launch_map(s->drm, st, t1->data, t3->data)) != VFPS_OK) return rc;
    } else {
        if ((rc = launch_map(s->rot, st, t2->data, t3->data)) != VFPS_OK) return rc;
    }
    if (s->fpm) {
        rc = launch_map(s->fpm, st, t3->data, t1->data);
    } else {
        VFPS_HIP(hipMemcpyAsync(t1->data, t3->data, sizeof(float) * t1->cells(),
                                hipMemcpyDeviceToDevice, st));
    }
    if (rc != VFPS_OK) return rc;
    rc = launch_project_x(t1, t1->data, t1->proj);
    if (rc != VFPS_OK) return rc;
    if (s->prof_hist) {
        hipLaunchKernelGGL(record_kernel, dim3(1), dim3(256), 0, st, t1->n,
                           t1->proj, s->prof_hist, s->counter, s->hist_cap);
        VFPS_HIP(hipGetLastError());
    }
    return VFPS_OK;
}

static void step_free(vfps_step* s)
{
    if (!s) return;
    if (s->dev) hipStreamSynchronize(s->dev->stream);
    if (s->graph) hipGraphExecDestroy(s->graph);
    if (s->graph_def) hipGraphDestroy(s->graph_def);
    vfps_map_destroy(s->wm);
    vfps_map_destroy(s->rfm);
    vfps_map_destroy(s->drm);
    vfps_map_destroy(s->rot);
    vfps_map_destroy(s->fpm);
    vfps_field_destroy(s->wfield);
    vfps_mesh_destroy(s->t1);
    vfps_mesh_destroy(s->t2);
    vfps_mesh_destroy(s->t3);
    hipFree(s->csr_renorm);
    hipFree(s->csr_tw);
    hipFree(s->filling);
    hipFree(s->counter);
    hipFree(s->err);
    hipFree(s->csr_hist);
    hipFree(s->prof_hist);
    hipFree(s->inst_d);
    hipFree(s->fparams_d);
    delete s;
}

// rebuild the device descriptors (history pointers change with record())
static int step_upload_desc(vfps_step* s)
{
    vfps_mesh* t1 = s->t1;
    InstDesc d{};
    d.n = s->desc.n;
    d.mesh0 = t1->data;
    d.mesh1 = s->t2->data;
    d.proj = t1->proj;
    d.ws = t1->ws;
    d.wake_off = s->desc.use_wake ? s->wfield->wake : nullptr;
    d.rf_off = s->rfm ? s->rfm->offset : nullptr;
    d.drift_off = s->drm ? s->drm->offset : nullptr;
    d.fp_w = s->fpm ? s->fpm->fp_w : nullptr;
    d.fp_i = s->fpm ? s->fpm->fp_i : nullptr;
    if (s->rot) d.rot = rot_geom(s->rot);
    d.prof_hist = s->prof_hist;
    d.hist_cap = s->hist_cap;
    d.counter = s->counter;
    d.err = s->err;
    d.seq_proj = s->desc.fast_projection ? 0u : 1u;
    uint32_t hlo = 0, hhi = 0;
    if (s->desc.scheme == 0) {
        std::vector<float> off(t1->n);
        VFPS_HIP(hipMemcpy(off.data(), s->drm->offset, sizeof(float) * t1->n,
                           hipMemcpyDeviceToHost));
        kick_halo(off, t1->n, s->desc.interp, &hlo, &hhi);
    } else {
        rotation_halo(d.rot, t1->n, s->desc.interp, &hlo, &hhi);
    }
    d.hlo = hlo;
    d.hhi = hhi;
    s->hmax = std::max(hlo, hhi);
    VFPS_HIP(hipMemcpy(s->inst_d, &d, sizeof(InstDesc), hipMemcpyHostToDevice));

    FieldParams p{};
    p.nb = 1;
    p.n = t1->n;
    p.M = s->M;
    p.spacing = s->desc.spacing_bins;
    p.proj = t1->proj;
    p.ws = t1->ws;
    p.filling = s->filling;
    p.counter = s->counter;
    p.csr_hist = s->csr_hist;
    p.hist_cap = s->hist_cap;
    if (s->wfield) {
        p.buckets = s->wfield->buckets_d;
        p.scaling = s->wfield->wakescaling;
        p.Z = s->wfield->z->z;
        p.tw_m = s->wfield->tw_m;
        p.tw_n = s->wfield->tw_n;
        p.wake = s->wfield->wake;
        p.wake2 = s->fused ? nullptr : s->wm->offset;
    }
    if (s->desc.use_csr) {
        p.Zc = s->zcsr->z;
        p.renorm = s->csr_renorm;
        p.fdelta = s->csr_fdelta;
        if (!s->wfield) {
            // CSR only: twiddles from a field-less table
            p.tw_m = s->csr_tw;
            p.tw_n = s->csr_tw + s->M;
        }
    }
    VFPS_HIP(hipMemcpy(s->fparams_d, &p, sizeof(FieldParams),
                       hipMemcpyHostToDevice));
    return VFPS_OK;
}

static int step_drop_graph(vfps_step* s)
{
    if (s->graph) {
        VFPS_HIP(hipGraphExecDestroy(s->graph));
        s->graph = nullptr;
    }
    if (s->graph_def) {
        VFPS_HIP(hipGraphDestroy(s->graph_def));
        s->graph_def = nullptr;
    }
    return VFPS_OK;
}

static int step_capture(vfps_step* s, uint32_t steps)
{
    int rc = step_drop_graph(s);
    if (rc != VFPS_OK) return rc;
    hipStream_t st = s->dev->stream;
    VFPS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (uint32_t k = 0; k < steps; k++) {
        rc = step_launch_one(s, k & 1u);
        if (rc != VFPS_OK) {
            hipGraph_t g;
            hipStreamEndCapture(st, &g);
            if (g) hipGraphDestroy(g);
            return rc;
        }
    }
    VFPS_HIP(hipStreamEndCapture(st, &s->graph_def));
    VFPS_HIP(hipGraphInstantiate(&s->graph, s->graph_def, nullptr, nullptr, 0));
    s->graph_steps = steps;
    return VFPS_OK;
}

extern "C" {

int vfps_step_create(vfps_dev* dev, const vfps_step_desc* desc,
                     const float* init, vfps_impedance* zwake,
                     vfps_impedance* zcsr, vfps_step** out)
{
    VFPS_CHECK_ARG(dev && desc && init && out, "vfps_step_create: NULL argument");
    VFPS_CHECK_ARG(desc->scheme <= 1, "vfps_step_create: scheme must be 0 or 1");
    VFPS_CHECK_ARG(!desc->use_wake || zwake, "vfps_step_create: wake impedance");
    VFPS_CHECK_ARG(!desc->use_csr || zcsr, "vfps_step_create: CSR impedance");
    VFPS_TRY_BEGIN
    auto* s = new vfps_step();
    std::unique_ptr<vfps_step, void (*)(vfps_step*)> guard(s, step_free);
    s->dev = dev;
    s->desc = *desc;
    s->fused = desc->fused != 0;
    s->zwake = zwake;
    s->zcsr = zcsr;
    const uint32_t n = desc->n;
    const float one = 1.0f;
    int rc;
    // three grids (src/main.cpp:470-535)
    rc = vfps_mesh_create(dev, 1, n, desc->qmin, desc->qmax, desc->qscale,
                          desc->pmin, desc->pmax, desc->pscale, 1.0, 1.0,
                          &one, 1.0, init, &s->t1);
    if (rc != VFPS_OK) return rc;
    if ((rc = vfps_mesh_clone(s->t1, &s->t2)) != VFPS_OK) return rc;
    if ((rc = vfps_mesh_clone(s->t1, &s->t3)) != VFPS_OK) return rc;
    const int it = (int)desc->interp;
    if (desc->use_wake) {
        const uint32_t bucket0 = 0;
        rc = vfps_field_create(s->t1, zwake, &bucket0, 1, desc->spacing_bins,
                               1.0, 1.0f, desc->wakescaling, &s->wfield);
        if (rc != VFPS_OK) return rc;
        rc = vfps_map_create_wakepotential(s->t1, s->t2, s->wfield, it, 0,
                                           &s->wm);
        if (rc != VFPS_OK) return rc;
        s->M = s->wfield->nmax / 2;
    } else if (desc->use_csr) {
        VFPS_CHECK_ARG(zcsr->n >= n && zcsr->n >= 4 && zcsr->n <= 8192 &&
                           (zcsr->n & (zcsr->n - 1)) == 0,
                       "vfps_step_create: CSR impedance length must be a "
                       "power of two >= the mesh size");
        s->M = zcsr->n / 2;
    } else {
        uint32_t m = 2;
        while (m < n) m <<= 1;
        s->M = m;
    }
    if (desc->scheme == 0) {
        if (desc->rf_linear) {
            rc = vfps_map_create_rfkick_linear(s->t2, s->t1, desc->rf_angle,
                                               desc->f_RF, it, 0, &s->rfm);
        } else {
            rc = vfps_map_create_rfkick(s->t2, s->t1, desc->revolutionpart,
                                        desc->V_RF, desc->f_RF, desc->V0, it,
                                        0, &s->rfm);
        }
        if (rc != VFPS_OK) return rc;
        rc = vfps_map_create_drift(s->t1, s->t3, desc->slip, desc->nslip,
                                   desc->E0, it, 0, &s->drm);
        if (rc != VFPS_OK) return rc;
    } else {
        rc = vfps_map_create_rotation(s->t2, s->t3, desc->angle, it,
                                      desc->rot_clamp, desc->rot_clamp ? n * n : 0,
                                      &s->rot);
        if (rc != VFPS_OK) return rc;
    }
    if (desc->use_fp) {
        rc = vfps_map_create_fokkerplanck(s->t3, s->t1, (int)desc->fptype, 0,
                                          desc->e1, (int)desc->deriv, &s->fpm);
        if (rc != VFPS_OK) return rc;
    }
    if (desc->use_csr) {
        VFPS_CHECK_ARG(zcsr->n == 2 * s->M,
                       "vfps_step_create: CSR and wake FFT lengths differ");
        const uint32_t nmax = zcsr->n;
        auto renorm = csr_renorm(nmax, s->t1->ax0.delta, desc->qscale, desc->fc,
                                 &s->csr_fdelta);
        std::vector<double2> twm, twn;
        make_twiddles(s->M, twm, twn);
        // twiddles of their own, used when there is no wake field
        VFPS_HIP(hipMalloc(&s->csr_renorm, sizeof(float) * nmax));
        VFPS_HIP(hipMemcpy(s->csr_renorm, renorm.data(), sizeof(float) * nmax,
                           hipMemcpyHostToDevice));
        VFPS_HIP(hipMalloc(&s->csr_tw, sizeof(double2) * 2 * s->M));
        VFPS_HIP(hipMemcpy(s->csr_tw, twm.data(), sizeof(double2) * s->M,
                           hipMemcpyHostToDevice));
        VFPS_HIP(hipMemcpy(s->csr_tw + s->M, twn.data(),
                           sizeof(double2) * s->M, hipMemcpyHostToDevice));
    }
    VFPS_HIP(hipMalloc(&s->filling, sizeof(double) * 2));
    VFPS_HIP(hipMalloc(&s->counter, sizeof(uint32_t)));
    VFPS_HIP(hipMalloc(&s->err, sizeof(uint32_t)));
    VFPS_HIP(hipMemset(s->counter, 0, sizeof(uint32_t)));
    VFPS_HIP(hipMemset(s->err, 0, sizeof(uint32_t)));
    VFPS_HIP(hipMalloc(&s->inst_d, sizeof(InstDesc)));
    VFPS_HIP(hipMalloc(&s->fparams_d, sizeof(FieldParams)));
    VFPS_HIP(hipSetDevice(dev->ordinal));
    VFPS_CHECK_ARG(field_lds_bytes(s->M) <= dev->lds_max,
                   "vfps_step_create: FFT length exceeds the LDS of one CU");
    // fused kernel variant and tile
    s->variant.scheme = (int)desc->scheme;
    s->variant.it = it;
    s->variant.ipfp = s->fpm ? s->fpm->ip : 0;
    s->variant.wake = desc->use_wake != 0;
    s->variant.clamp = desc->rot_clamp != 0;
    if ((rc = step_upload_desc(s)) != VFPS_OK) return rc;
    if (s->fused) {
        if (!maps_kernel_for(s->variant)) {
            return fail(VFPS_ERR_UNSUPPORTED,
                        "vfps_step_create: fused step needs linear or cubic "
                        "interpolation (use fused = 0)");
        }
        uint32_t T = desc->tile_rows ? desc->tile_rows : (n >= 512 ? 4u : 2u);
        while (T > 1 && fused_lds_bytes(n, T, s->hmax) > dev->lds_max) {
            T >>= 1;
        }
        VFPS_CHECK_ARG(fused_lds_bytes(n, T, s->hmax) <= dev->lds_max,
                       "vfps_step_create: mesh row too long for the LDS of one "
                       "CU (use fused = 0)");
        s->T = T;
        s->threads = 256;
        if ((rc = fused_set_lds_limit(s->variant,
                                      fused_lds_bytes(n, T, s->hmax))) != VFPS_OK)
            return rc;
    }
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    *out = guard.release();
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_step_destroy(vfps_step* s)
{
    step_free(s);
    return VFPS_OK;
}

int vfps_step_run(vfps_step* s, uint32_t nsteps)
{
    VFPS_CHECK_ARG(s, "vfps_step_run: NULL step");
    hipStream_t st = s->dev->stream;
    uint32_t left = nsteps;
    int rc;
    // bring the fused ping-pong to parity 0 before replaying the graph
    if (s->fused && (s->parity & 1u) && left > 0) {
        if ((rc = step_launch_one(s, 1)) != VFPS_OK) return rc;
        s->parity = 0;
        left--;
    }
    const uint32_t G = 64;
    if (left >= G) {
        if (!s->graph || s->graph_steps != G) {
            if ((rc = step_capture(s, G)) != VFPS_OK) return rc;
        }
        while (left >= G) {
            VFPS_HIP(hipGraphLaunch(s->graph, st));
            left -= G;
        }
    }
    while (left > 0) {
        if ((rc = step_launch_one(s, s->fused ? s->parity : 0)) != VFPS_OK)
            return rc;
        if (s->fused) s->parity ^= 1u;
        left--;
    }
    VFPS_HIP(hipStreamSynchronize(st));
    uint32_t err = 0;
    VFPS_HIP(hipMemcpy(&err, s->err, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (err) {
        return fail(VFPS_ERR_INVALID, "fused step: a tap fell outside the "
                                      "halo rows (internal error)");
    }
    return VFPS_OK;
}

int vfps_step_record(vfps_step* s, uint32_t capacity, int want_profiles)
{
    VFPS_CHECK_ARG(s, "vfps_step_record: NULL step");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    hipFree(s->csr_hist);
    hipFree(s->prof_hist);
    s->csr_hist = nullptr;
    s->prof_hist = nullptr;
    s->hist_cap = capacity;
    if (capacity) {
        VFPS_HIP(hipMalloc(&s->csr_hist, sizeof(float) * capacity));
        VFPS_HIP(hipMemset(s->csr_hist, 0, sizeof(float) * capacity));
        if (want_profiles) {
            VFPS_HIP(hipMalloc(&s->prof_hist,
                               sizeof(float) * (size_t)capacity * s->desc.n));
        }
    }
    VFPS_HIP(hipMemset(s->counter, 0, sizeof(uint32_t)));
    int rc = step_drop_graph(s);
    if (rc != VFPS_OK) return rc;
    return step_upload_desc(s);
}

int vfps_step_get_history(vfps_step* s, uint32_t* count, float* csr_power,
                          float* profiles)
{
    VFPS_CHECK_ARG(s, "vfps_step_get_history: NULL step");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    uint32_t c = 0;
    VFPS_HIP(hipMemcpy(&c, s->counter, sizeof(uint32_t), hipMemcpyDeviceToHost));
    c = std::min(c, s->hist_cap);
    if (count) *count = c;
    if (csr_power && s->csr_hist && c) {
        VFPS_HIP(hipMemcpy(csr_power, s->csr_hist, sizeof(float) * c,
                           hipMemcpyDeviceToHost));
    }
    if (profiles && s->prof_hist && c) {
        VFPS_HIP(hipMemcpy(profiles, s->prof_hist,
                           sizeof(float) * (size_t)c * s->desc.n,
                           hipMemcpyDeviceToHost));
    }
    return VFPS_OK;
}

int vfps_step_get_mesh(vfps_step* s, float* mesh, float* proj, float* wake)
{
    VFPS_CHECK_ARG(s, "vfps_step_get_mesh: NULL step");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    const size_t n = s->desc.n;
    const float* cur = (s->fused && (s->parity & 1u)) ? s->t2->data : s->t1->data;
    if (mesh) {
        VFPS_HIP(hipMemcpy(mesh, cur, sizeof(float) * n * n, hipMemcpyDeviceToHost));
    }
    if (proj) {
        VFPS_HIP(hipMemcpy(proj, s->t1->proj, sizeof(float) * n,
                           hipMemcpyDeviceToHost));
    }
    if (wake) {
        if (s->wfield) {
            VFPS_HIP(hipMemcpy(wake, s->wfield->wake, sizeof(float) * n,
                               hipMemcpyDeviceToHost));
        } else {
            std::fill(wake, wake + n, 0.0f);
        }
    }
    return VFPS_OK;
}

int vfps_step_set_mesh(vfps_step* s, const float* mesh)
{
    VFPS_CHECK_ARG(s && mesh, "vfps_step_set_mesh: NULL argument");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    VFPS_HIP(hipMemcpy(s->t1->data, mesh, sizeof(float) * s->t1->cells(),
                       hipMemcpyHostToDevice));
    s->parity = 0;
    int rc = launch_project_x(s->t1, s->t1->data, s->t1->proj);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    return VFPS_OK;
}

int vfps_step_phases(vfps_step* s, double* ns_out, uint32_t nout)
{
    VFPS_CHECK_ARG(s && ns_out && nout >= 16, "vfps_step_phases: bad argument");
    VFPS_CHECK_ARG(s->fused, "vfps_step_phases: needs the fused step");
    hipStream_t st = s->dev->stream;
    VFPS_HIP(hipStreamSynchronize(st));
    unsigned long long* buf = nullptr;
    VFPS_HIP(hipMalloc(&buf, sizeof(unsigned long long) * 16));
    VFPS_HIP(hipMemset(buf, 0, sizeof(unsigned long long) * 16));
    // point both descriptors at the stamp buffer for one step
    unsigned long long* mstamp = buf;
    unsigned long long* fstamp = buf + 8;
    VFPS_HIP(hipMemcpy(reinterpret_cast<char*>(s->inst_d) + offsetof(InstDesc, tstamp),
                       &mstamp, sizeof(mstamp), hipMemcpyHostToDevice));
    VFPS_HIP(hipMemcpy(reinterpret_cast<char*>(s->fparams_d) + offsetof(FieldParams, tstamp),
                       &fstamp, sizeof(fstamp), hipMemcpyHostToDevice));
    int rc = step_launch_one(s, s->parity);
    s->parity ^= 1u;
    hipError_t e = hipStreamSynchronize(st);
    unsigned long long h[16] = {0};
    hipMemcpy(h, buf, sizeof(h), hipMemcpyDeviceToHost);
    unsigned long long* nul = nullptr;
    hipMemcpy(reinterpret_cast<char*>(s->inst_d) + offsetof(InstDesc, tstamp),
              &nul, sizeof(nul), hipMemcpyHostToDevice);
    hipMemcpy(reinterpret_cast<char*>(s->fparams_d) + offsetof(FieldParams, tstamp),
              &nul, sizeof(nul), hipMemcpyHostToDevice);
    hipFree(buf);
    if (rc != VFPS_OK) return rc;
    if (e != hipSuccess) return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    int khz = 0;
    VFPS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                                   s->dev->ordinal));
    const double ns_per_tick = khz > 0 ? 1e6 / khz : 10.0;
    // [0..6]: maps kernel phase ends relative to its start;
    // [8..13]: field kernel phase ends relative to its start; 15: tick ns
    for (uint32_t i = 0; i < 16; i++) ns_out[i] = 0;
    for (int i = 1; i < 7; i++) {
        if (h[i] && h[0]) ns_out[i] = (double)(h[i] - h[0]) * ns_per_tick;
    }
    for (int i = 9; i < 14; i++) {
        if (h[i] && h[8]) ns_out[i] = (double)(h[i] - h[8]) * ns_per_tick;
    }
    ns_out[15] = ns_per_tick;
    return VFPS_OK;
}

int vfps_step_timing(vfps_step* s, uint32_t nsteps, float* ms_wake,
                     float* ms_maps, float* ms_step)
{
    VFPS_CHECK_ARG(s && nsteps > 0, "vfps_step_timing: bad argument");
    hipStream_t st = s->dev->stream;
    hipEvent_t e0, e1, e2;
    VFPS_HIP(hipEventCreate(&e0));
    VFPS_HIP(hipEventCreate(&e1));
    VFPS_HIP(hipEventCreate(&e2));
    const bool wake = s->desc.use_wake != 0;
    const bool csr = s->desc.use_csr != 0;
    double tw = 0, tm = 0;
    int rc = VFPS_OK;
    for (uint32_t k = 0; k < nsteps && rc == VFPS_OK; k++) {
        const uint32_t par = s->fused ? s->parity : 0;
        hipEventRecord(e0, st);
        rc = launch_fields(wake, csr, s->fparams_d, 1, s->M, st);
        hipEventRecord(e1, st);
        if (rc == VFPS_OK) {
            if (s->fused) {
                rc = launch_fused_maps(s->variant, s->inst_d, 1, s->desc.n,
                                       s->T, s->threads, s->hmax, par, st);
            } else {
                rc = VFPS_ERR_UNSUPPORTED;
                set_error("vfps_step_timing: per-kernel timing needs the "
                          "fused step");
            }
        }
        hipEventRecord(e2, st);
        hipEventSynchronize(e2);
        float a = 0, b = 0;
        hipEventElapsedTime(&a, e0, e1);
        hipEventElapsedTime(&b, e1, e2);
        tw += a;
        tm += b;
        if (s->fused) s->parity ^= 1u;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (rc != VFPS_OK) return rc;
    if (ms_wake) *ms_wake = (float)(tw / nsteps);
    if (ms_maps) *ms_maps = (float)(tm / nsteps);
    if (ms_step) *ms_step = (float)((tw + tm) / nsteps);
    return VFPS_OK;
}

}  // extern "C"
