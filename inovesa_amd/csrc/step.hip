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
#include <map>
#include <memory>
#include <mutex>

#include "common.hpp"
#include "field_body.hpp"
#include "reduce.hpp"
#include "step.hpp"
#include "step_kernel.hpp"
#include "step_select.hpp"
#include "strip_kernel.hpp"

namespace vfps_impl {

// ---- fused step kernel ---------------------------------------------------------

template <int IT, bool DIRECT>
static const void* step_fn_fp(const StepVariant& v, int form, uint32_t maxc)
{
    switch (v.ipfp) {
    case 0: return step_kernels<IT, DIRECT, 0>(v, form, maxc);
    case 3: return step_kernels<IT, DIRECT, 3>(v, form, maxc);
    case 4: return step_kernels<IT, DIRECT, 4>(v, form, maxc);
    case 5: return step_kernels<IT, DIRECT, 5>(v, form, maxc);
    }
    return nullptr;
}

// resolve a variant to its kernel (cubic and linear interpolation; the
// clamp needs cubic interpolation); form: StepForm; maxc: most columns per
// thread (1, 2, else up to STEP_MAX_COLS)
const void* step_kernel_for(const StepVariant& v, int form, uint32_t maxc)
{
    if (maxc > (uint32_t)STEP_MAX_COLS) {
        return nullptr;
    }
    if (v.it == 4) {
        return v.direct ? step_fn_fp<4, true>(v, form, maxc)
                        : step_fn_fp<4, false>(v, form, maxc);
    }
    if (v.it == 2 && !v.clamp) {
        return v.direct ? step_fn_fp<2, true>(v, form, maxc)
                        : step_fn_fp<2, false>(v, form, maxc);
    }
    return nullptr;
}

size_t step_lds_bytes(const StepVariant& v, uint32_t n, uint32_t T,
                      uint32_t stride, bool one_mesh)
{
    return sizeof(float) *
           step_lds(n, T, stride, v.direct, (uint32_t)v.it, v.scheme, one_mesh)
               .words;
}

size_t step_lds_of(const vfps_step* s)
{
    if (s->strip) {
        return strip_lds_bytes(s->desc.n, s->T, s->strip, s->stride - s->T);
    }
    return step_lds_bytes(s->variant, s->desc.n, s->T, s->stride, !s->in_batch);
}

int step_set_lds_limit(const StepVariant& v, size_t bytes)
{
    // the largest limit already set per kernel (raising it is a runtime
    // call; retiling the instances of a batch asks for it many times)
    static std::mutex mu;
    static std::map<const void*, size_t> set;
    std::lock_guard<std::mutex> lock(mu);
    for (int form : {STEP_SINGLE, STEP_PER_TILE, STEP_PERSISTENT}) {
        for (uint32_t maxc : {1u, 2u, (uint32_t)STEP_MAX_COLS}) {
            const void* fn = step_kernel_for(v, form, maxc);
            if (!fn && form == STEP_PERSISTENT && !v.direct) {
                continue;  // the resident-grid form is direct-wake only
            }
            if (!fn) {
                return fail(VFPS_ERR_UNSUPPORTED, "fused step: no kernel for "
                                                  "this interpolation/stencil "
                                                  "combination");
            }
            if (bytes > 65536 && bytes > set[fn]) {
                VFPS_HIP(hipFuncSetAttribute(
                    fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
                set[fn] = bytes;
            }
        }
    }
    return VFPS_OK;
}

int launch_step_kernel(const StepVariant& v, const InstDesc* descs_d,
                       const uint32_t* blockmap_d, uint32_t nblocks,
                       size_t lds_bytes, uint32_t threads, uint32_t maxn,
                       uint32_t parity, hipStream_t st, const InstDesc* desc_h,
                       uint32_t grid)
{
    const bool single = desc_h != nullptr;
    if (single == (blockmap_d != nullptr)) {
        return fail(VFPS_ERR_INVALID, "fused step: one mesh takes a by-value "
                                      "descriptor, a batch a block map");
    }
    if (grid == 0) {
        grid = nblocks;
    }
    if (!single && (nblocks % 8u != 0u || grid % 8u != 0u || grid == 0 ||
                    grid > nblocks)) {
        return fail(VFPS_ERR_INVALID, "fused step: a batch's block map and "
                                      "grid must be multiples of 8, the grid "
                                      "at most the map");
    }
    if (single && grid != nblocks) {
        return fail(VFPS_ERR_INVALID, "fused step: one mesh runs one "
                                      "workgroup per tile");
    }
    if (single && v.direct) {
        grid = nblocks + 1u;  // and one that closes the step
    }
    if (threads == 0 || threads > STEP_MAX_THREADS || threads % 64u != 0u) {
        return fail(VFPS_ERR_INVALID, "fused step: workgroup size");
    }
    const int form = single ? STEP_SINGLE
                            : (grid < nblocks ? STEP_PERSISTENT : STEP_PER_TILE);
    const void* fn = step_kernel_for(v, form, ceil_div(maxn, threads));
    if (!fn) {
        return fail(VFPS_ERR_UNSUPPORTED, "fused step: no kernel for this "
                                          "interpolation/stencil combination");
    }
    InstDesc d0{};
    if (single) {
        d0 = *desc_h;
    }
    void* args[] = {(void*)&descs_d, (void*)&blockmap_d, (void*)&nblocks,
                    (void*)&parity, (void*)&d0};
    VFPS_HIP(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds_bytes,
                             st));
    return VFPS_OK;
}

int step_occupancy(const StepVariant& v, uint32_t threads, uint32_t maxn,
                   size_t lds_bytes, int* per_cu)
{
    const void* fn = step_kernel_for(v, STEP_PERSISTENT, ceil_div(maxn, threads));
    if (!fn) {
        return fail(VFPS_ERR_UNSUPPORTED, "fused step: no kernel for this "
                                          "interpolation/stencil combination");
    }
    int occ = 0;
    VFPS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, (int)threads,
                                                          lds_bytes));
    *per_cu = occ;
    return VFPS_OK;
}

// Toeplitz kernels of the direct wake and CSR forms, in FP64 from the
// float tables.  With N = nmax = 2M and theta = 2 pi k d / N:
//   g[d] = Re Z_0 + 2 sum_{k=1}^{M-1} (Re Z_k cos theta - Im Z_k sin theta)
// is C2R(Z) with the Nyquist bin zero and Im Z_0 dropped, so that
//   C2R(Z * R2C(pad(p)))[x] = sum_x' g[x - x'] p[x'];
//   h[d] = sum_{k=0}^{M} c_k cos theta,  c_k = fdelta renorm_k Re Zc_k,
// gives the CSR power sum_k c_k |F_k|^2 = sum_{x,x'} p[x] h[x - x'] p[x'].
// Tables hold d = -(n-1)..(n-1) at index d + n - 1.
void toeplitz_tables(uint32_t n, uint32_t M, const float2* Z,
                     const float2* Zc, const float* renorm, float fdelta,
                     std::vector<double>* g, std::vector<double>* h)
{
    const uint64_t N = 2ull * M;
    std::vector<double> cs(N), sn(N);
    for (uint64_t j = 0; j < N; j++) {
        const double a = 2.0 * M_PI * (double)j / (double)N;
        cs[j] = std::cos(a);
        sn[j] = std::sin(a);
    }
    const int64_t nd = 2 * (int64_t)n - 1;
    if (g) {
        g->assign((size_t)nd, 0.0);
        for (int64_t i = 0; i < nd; i++) {
            const int64_t dd = i - (int64_t)(n - 1);
            const uint64_t step = (uint64_t)((dd % (int64_t)N + (int64_t)N) % (int64_t)N);
            double acc = 0.0;
            uint64_t j = 0;
            for (uint32_t k = 1; k < M; k++) {
                j += step;
                if (j >= N) j -= N;
                acc += (double)Z[k].x * cs[j] - (double)Z[k].y * sn[j];
            }
            (*g)[(size_t)i] = (double)Z[0].x + 2.0 * acc;
        }
    }
    if (h) {
        std::vector<double> c(M + 1);
        for (uint32_t k = 0; k <= M; k++) {
            c[k] = (double)fdelta * (double)renorm[k] * (double)Zc[k].x;
        }
        h->assign((size_t)nd, 0.0);
        for (int64_t i = 0; i < nd; i++) {
            const int64_t dd = i - (int64_t)(n - 1);
            const uint64_t step = (uint64_t)((dd % (int64_t)N + (int64_t)N) % (int64_t)N);
            double acc = c[0];
            uint64_t j = 0;
            for (uint32_t k = 1; k <= M; k++) {
                j += step;
                if (j >= N) j -= N;
                acc += c[k] * cs[j];
            }
            (*h)[(size_t)i] = acc;
        }
    }
}

// ---- fields kernel (one workgroup per instance) ------------------------------

template <bool WAKE, bool CSR>
__global__ void __launch_bounds__(1024)
fields_kernel(const FieldParams* __restrict__ params, uint32_t parity)
{
    extern __shared__ double2 flds[];
    FieldParams p = params[blockIdx.x];
    if (parity && p.proj_alt) {
        p.proj = p.proj_alt;
    }
    p.parity = parity;
    field_body<WAKE, CSR>(p, 0, flds);
}

// One instance with its parameters by value: the kernel starts from its
// arguments instead of a load of the parameter block through a pointer.
template <bool WAKE, bool CSR>
__global__ void __launch_bounds__(1024)
field_one_kernel(FieldParams p, uint32_t parity)
{
    extern __shared__ double2 flds[];
    if (parity && p.proj_alt) {
        p.proj = p.proj_alt;
    }
    p.parity = parity;
    field_body<WAKE, CSR>(p, 0, flds);
}

template <bool WAKE, bool CSR>
static const void* fields_fn(bool one)
{
    return one ? reinterpret_cast<const void*>(&field_one_kernel<WAKE, CSR>)
               : reinterpret_cast<const void*>(&fields_kernel<WAKE, CSR>);
}

int launch_fields(bool wake, bool csr, const FieldParams* params_d,
                  uint32_t ninst, uint32_t Mmax, uint32_t parity,
                  hipStream_t st, const FieldParams* params_h)
{
    const size_t lds = field_lds_bytes(Mmax);
    const bool one = ninst == 1 && params_h;
    const void* fn;
    if (wake && csr) {
        fn = fields_fn<true, true>(one);
    } else if (wake) {
        fn = fields_fn<true, false>(one);
    } else if (csr) {
        fn = fields_fn<false, true>(one);
    } else {
        fn = fields_fn<false, false>(one);
    }
    if (lds > 65536) {
        VFPS_HIP(hipFuncSetAttribute(fn,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    }
    if (one) {
        FieldParams p = *params_h;
        void* args[] = {(void*)&p, (void*)&parity};
        VFPS_HIP(hipLaunchKernel(fn, dim3(1), dim3(FFT_THREADS), args, lds, st));
        return VFPS_OK;
    }
    void* args[] = {(void*)&params_d, (void*)&parity};
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

// the fused step's kernels for one step at the given parity
static int step_launch_fused(vfps_step* s, uint32_t parity, hipStream_t st)
{
    if (!s->direct) {
        int rc = launch_fields(s->desc.use_wake != 0, s->desc.use_csr != 0,
                               s->fparams_d, 1, s->M, parity, st,
                               &s->fparams_h);
        if (rc != VFPS_OK) return rc;
    }
    return launch_step_kernel(s->variant, s->inst_d, nullptr, s->tiles(),
                              step_lds_bytes(s->variant, s->desc.n, s->T,
                                                   s->stride, true),
                              s->threads, s->desc.n, parity, st, &s->inst_h);
}

static int step_launch_maps(vfps_step* s, uint32_t parity);

// one step, then the tracked particles through its maps (applyToAll after
// each apply, src/main.cpp:1034-1041); the step counter the noise of the
// stochastic tracking draws on is the one the step has just written
static int step_launch_one(vfps_step* s, uint32_t parity)
{
    int rc = step_launch_maps(s, parity);
    if (rc != VFPS_OK || s->npart == 0) return rc;
    const uint32_t* cnt = s->counter + (s->fused ? (parity ^ 1u) : 0u);
    return launch_track(s->track, s->particles, s->npart, cnt, 0,
                        s->dev->stream);
}

static int step_launch_maps(vfps_step* s, uint32_t parity)
{
    hipStream_t st = s->dev->stream;
    const bool wake = s->desc.use_wake != 0;
    const bool csr = s->desc.use_csr != 0;
    if (s->fused) {
        return step_launch_fused(s, parity, st);
    }
    // per-map path, the loop body of src/main.cpp:948-1053
    int rc = launch_fields(wake, csr, s->fparams_d, 1, s->M, 0, st,
                           &s->fparams_h);
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
    for (auto& kv : s->graphs) {
        hipGraphExecDestroy(kv.second.second);
        hipGraphDestroy(kv.second.first);
    }
    s->graphs.clear();
    hipFree(s->particles);
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
    if (s->err_h) hipHostFree(s->err_h);
    hipFree(s->csr_hist);
    hipFree(s->prof_hist);
    hipFree(s->wake_g);
    hipFree(s->csr_h);
    hipFree(s->csr_part);
    hipFree(s->csr_valid);
    hipFree(s->csr_power);
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
    d.tiles = s->tiles();
    d.mesh[0] = t1->data;
    d.mesh[1] = s->t2->data;
    d.proj[0] = t1->proj;
    d.proj[1] = s->t2->proj;
    d.ws = t1->ws;
    d.use_wake = s->desc.use_wake ? 1 : 0;
    d.use_csr = s->desc.use_csr ? 1 : 0;
    d.wake_off = s->desc.use_wake ? s->wfield->wake : nullptr;
    d.wake_out = s->desc.use_wake ? s->wfield->wake : nullptr;
    d.wake_g = s->wake_g;
    d.wake_scale = s->wfield ? s->wfield->wakescaling : 0.0f;
    d.csr_h = s->csr_h;
    d.csr_part = s->csr_part;
    d.csr_valid = s->csr_valid;
    d.csr_power = s->csr_power;
    d.filling = s->filling;
    d.rf_off = s->rfm ? s->rfm->offset : nullptr;
    d.drift_off = s->drm ? s->drm->offset : nullptr;
    d.fp_w = s->fpm ? s->fpm->fp_w : nullptr;
    d.fp_i = s->fpm ? s->fpm->fp_i : nullptr;
    if (s->rot) d.rot = rot_geom(s->rot);
    d.prof_hist = s->prof_hist;
    d.csr_hist = s->csr_hist;
    d.hist_cap = s->hist_cap;
    d.counter = s->counter;
    d.err = s->err;
    d.seq_proj = s->desc.fast_projection ? 0u : 1u;
    d.norm_every = s->desc.renormalize;
    d.fill_set = t1->filling_set.empty() ? 1.0f : t1->filling_set[0];
    // the halo depends on the maps only: computed once
    if (!s->halo_valid) {
        if (s->desc.scheme == 0) {
            std::vector<float> off(t1->n);
            VFPS_HIP(hipMemcpy(off.data(), s->drm->offset, sizeof(float) * t1->n,
                               hipMemcpyDeviceToHost));
            kick_halo(off, t1->n, s->desc.interp, &s->halo_lo, &s->halo_hi);
        } else {
            rotation_halo(d.rot, t1->n, s->desc.interp, &s->halo_lo, &s->halo_hi);
        }
        s->halo_valid = true;
    }
    const uint32_t hlo = s->halo_lo, hhi = s->halo_hi;
    d.hlo = hlo;
    d.hhi = hhi;
    s->hmax = hlo + hhi;
    s->stride = s->T + hlo + hhi;
    // the fused kernel gives each buffer row (tile plus halo) one thread
    VFPS_CHECK_ARG(!s->fused || s->stride <= s->threads,
                   "fused step: tile rows plus halo rows exceed the workgroup "
                   "size");
    d.T = s->T;
    d.stride = s->stride;
    d.K = s->strip;
    d.ringw = s->strip ? strip_ring(s->T, hlo + hhi) : 0u;
    s->inst_h = d;

    FieldParams p{};
    p.nb = 1;
    p.n = t1->n;
    p.M = s->M;
    p.spacing = s->desc.spacing_bins;
    p.proj = t1->proj;
    p.proj_alt = s->fused ? s->t2->proj : nullptr;
    p.ws = t1->ws;
    p.filling = s->filling;
    p.counter = s->counter;
    p.counter_pingpong = s->fused ? 1u : 0u;
    p.csr_hist = s->csr_hist;
    p.hist_cap = s->hist_cap;
    p.power = s->csr_power;
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
    s->fparams_h = p;
    // uploaded before the next launch that reads them (step_flush_desc), so
    // retiling the instances of a batch costs no device copies
    s->desc_dirty = true;
    return VFPS_OK;
}

// the device copies of the descriptors, if the host copies changed
static int step_flush_desc(vfps_step* s)
{
    if (s->desc_dirty) {
        VFPS_HIP(hipMemcpy(s->inst_d, &s->inst_h, sizeof(InstDesc),
                           hipMemcpyHostToDevice));
        VFPS_HIP(hipMemcpy(s->fparams_d, &s->fparams_h, sizeof(FieldParams),
                           hipMemcpyHostToDevice));
        s->desc_dirty = false;
    }
    return VFPS_OK;
}

static int step_drop_graph(vfps_step* s)
{
    for (auto& kv : s->graphs) {
        VFPS_HIP(hipGraphExecDestroy(kv.second.second));
        VFPS_HIP(hipGraphDestroy(kv.second.first));
    }
    s->graphs.clear();
    return VFPS_OK;
}

// longest captured run; a run of nsteps replays it nsteps / STEP_GRAPH
// times, then one captured run of the remaining steps
constexpr uint32_t STEP_GRAPH = 64;

// the captured run of `steps` steps from parity p (captured on first use)
// (fin, fused steps: the run ends with this graph, which then also closes
// it: the last step's CSR power and the error flag into its pinned host
// copy)
static int step_graph(vfps_step* s, uint32_t steps, uint32_t p, bool fin,
                      hipGraphExec_t* out)
{
    fin = fin && s->fused;
    const uint32_t key = steps << 2 | (fin ? 2u : 0u) | (p & 1u);
    auto it = s->graphs.find(key);
    if (it != s->graphs.end()) {
        *out = it->second.second;
        return VFPS_OK;
    }
    // remainders differ from run to run: keep the cache small
    if (s->graphs.size() >= 16) {
        int rc = step_drop_graph(s);
        if (rc != VFPS_OK) return rc;
    }
    hipStream_t st = s->dev->stream;
    VFPS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = VFPS_OK;
    for (uint32_t k = 0; k < steps && rc == VFPS_OK; k++) {
        rc = step_launch_one(s, s->fused ? ((p + k) & 1u) : 0u);
    }
    if (rc == VFPS_OK && fin) {
        rc = launch_step_finalize(s->inst_d, 1, (p + steps) & 1u, st, s->err_h,
                                  s->direct);
    }
    if (rc != VFPS_OK) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(st, &g);
        if (g) hipGraphDestroy(g);
        return rc;
    }
    hipGraph_t def = nullptr;
    hipGraphExec_t exec = nullptr;
    VFPS_HIP(hipStreamEndCapture(st, &def));
    hipError_t e = hipGraphInstantiate(&exec, def, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(def);
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    s->graphs[key] = {def, exec};
    // uploaded now, so that the first replay does not pay for it
    VFPS_HIP(hipGraphUpload(exec, st));
    *out = exec;
    return VFPS_OK;
}

// Toeplitz tables and scratch of the direct wake / CSR forms
static int step_setup_direct(vfps_step* s)
{
    const uint32_t n = s->desc.n;
    const uint32_t M = s->M;
    const uint32_t nmax = 2 * M;
    std::vector<float2> z, zc;
    std::vector<float> rn;
    if (s->desc.use_wake) {
        z.resize(nmax);
        VFPS_HIP(hipMemcpy(z.data(), s->zwake->z, sizeof(float2) * nmax,
                           hipMemcpyDeviceToHost));
    }
    if (s->desc.use_csr) {
        zc.resize(nmax);
        rn.resize(nmax);
        VFPS_HIP(hipMemcpy(zc.data(), s->zcsr->z, sizeof(float2) * nmax,
                           hipMemcpyDeviceToHost));
        VFPS_HIP(hipMemcpy(rn.data(), s->csr_renorm, sizeof(float) * nmax,
                           hipMemcpyDeviceToHost));
    }
    std::vector<double> g, h;
    toeplitz_tables(n, M, z.empty() ? nullptr : z.data(),
                    zc.empty() ? nullptr : zc.data(),
                    rn.empty() ? nullptr : rn.data(), s->csr_fdelta,
                    s->desc.use_wake ? &g : nullptr,
                    s->desc.use_csr ? &h : nullptr);
    const size_t nd = 2 * (size_t)n - 1;
    if (!g.empty()) {
        VFPS_HIP(hipMalloc(&s->wake_g, sizeof(double) * nd));
        VFPS_HIP(hipMemcpy(s->wake_g, g.data(), sizeof(double) * nd,
                           hipMemcpyHostToDevice));
    }
    if (!h.empty()) {
        VFPS_HIP(hipMalloc(&s->csr_h, sizeof(double) * nd));
        VFPS_HIP(hipMemcpy(s->csr_h, h.data(), sizeof(double) * nd,
                           hipMemcpyHostToDevice));
    }
    VFPS_HIP(hipMalloc(&s->csr_part, sizeof(double) * 2 * n));
    VFPS_HIP(hipMemset(s->csr_part, 0, sizeof(double) * 2 * n));
    VFPS_HIP(hipMalloc(&s->csr_valid, sizeof(uint32_t) * 2));
    VFPS_HIP(hipMemset(s->csr_valid, 0xFF, sizeof(uint32_t) * 2));
    return VFPS_OK;
}

namespace vfps_impl {

// change the tile height and workgroup size of a fused step (the results
// do not depend on them)
int step_retile(vfps_step* s, uint32_t T, uint32_t threads, bool reset)
{
    VFPS_CHECK_ARG(s && s->fused, "retile: needs a fused step");
    VFPS_CHECK_ARG(threads % 64 == 0 && threads >= 64 &&
                       threads <= STEP_MAX_THREADS,
                   "retile: threads must be a multiple of 64 up to 512");
    VFPS_CHECK_ARG(T >= 1 && T <= threads && T <= s->desc.n,
                   "retile: tile rows must be in [1, min(threads, n)]");
    VFPS_CHECK_ARG(ceil_div(s->desc.n, threads) <= (uint32_t)STEP_MAX_COLS,
                   "retile: too few threads for the mesh size");
    if (s->strip) {
        // the strip form: kick-drift maps with the direct wake, in a batch;
        // one wavefront sums the x projection of a strip, a row per lane
        VFPS_CHECK_ARG(s->in_batch && strip_kernel_for(s->variant, 1) != nullptr &&
                           !s->desc.fast_projection,
                       "retile: strips need a batch instance with the "
                       "kick-drift maps, the direct wake and the reference "
                       "order of the x projection");
        VFPS_CHECK_ARG(T <= STRIP_MAX_T && T * s->strip <= STRIP_MAX_ROWS &&
                           s->desc.n % 4 == 0 &&
                           (strip_ring(T, s->hmax) + T) * s->desc.n >=
                               16 * T * s->strip,
                       "retile: strips of at most 64 rows in tiles of at most "
                       "8 rows");
    }
    const uint32_t oldT = s->T, oldthr = s->threads;
    s->T = T;
    s->threads = threads;
    int rc = step_upload_desc(s);
    const size_t lds = step_lds_of(s);
    if (rc == VFPS_OK && lds > s->dev->lds_max) {
        rc = fail(VFPS_ERR_INVALID, "retile: tile exceeds the LDS of one CU");
    }
    if (rc == VFPS_OK) {
        rc = s->strip ? strip_set_lds_limit(s->variant, lds)
                      : step_set_lds_limit(s->variant, lds);
    }
    if (rc != VFPS_OK) {
        s->T = oldT;
        s->threads = oldthr;
        step_upload_desc(s);
        return rc;
    }
    if (reset && (rc = step_reset_pending(s)) != VFPS_OK) return rc;
    return step_drop_graph(s);
}

// The CSR power of the last step of a direct-form run (the shares at
// parity^1), one workgroup per instance; the counter is at parity.
__global__ void __launch_bounds__(256)
step_finalize_kernel(const InstDesc* __restrict__ descs, uint32_t parity,
                     uint32_t* err_host, uint32_t close_csr)
{
    __shared__ double red[16];
    const InstDesc& d = descs[blockIdx.x];
    if (err_host && threadIdx.x == 0) {
        err_host[blockIdx.x] = *d.err;
    }
    if (!close_csr || !d.use_csr || !d.counter) {
        return;
    }
    const uint32_t cnt = d.counter[parity];
    if (cnt < 1 || d.csr_valid[parity ^ 1u] != cnt) {
        return;
    }
    double a = 0.0;
    for (uint32_t k = threadIdx.x; k < d.tiles; k += blockDim.x) {
        a += d.csr_part[(size_t)(parity ^ 1u) * d.n + k];
    }
    const double p = block_sum(a, red);
    if (threadIdx.x == 0) {
        *d.csr_power = (float)p;
        if (d.csr_hist && cnt - 1u < d.hist_cap) {
            d.csr_hist[cnt - 1u] = (float)p;
        }
    }
}

int step_reset_pending(vfps_step* s)
{
    // in stream order, after whatever step is still in flight: both
    // counter entries hold the current count, and no CSR share is pending
    hipStream_t st = s->dev->stream;
    const uint32_t p = s->parity & 1u;
    VFPS_HIP(hipMemcpyAsync(s->counter + (p ^ 1u), s->counter + p,
                            sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    if (s->csr_valid) {
        VFPS_HIP(hipMemsetAsync(s->csr_valid, 0xFF, sizeof(uint32_t) * 2, st));
    }
    return VFPS_OK;
}

int launch_step_finalize(const InstDesc* descs_d, uint32_t ninst,
                         uint32_t parity, hipStream_t st, uint32_t* err_host,
                         bool close_csr)
{
    hipLaunchKernelGGL(step_finalize_kernel, dim3(ninst), dim3(256), 0, st,
                       descs_d, parity, err_host, close_csr ? 1u : 0u);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

}  // namespace vfps_impl

extern "C" {

int vfps_step_create(vfps_dev* dev, const vfps_step_desc* desc,
                     const float* init, vfps_impedance* zwake,
                     vfps_impedance* zcsr, vfps_step** out)
{
    VFPS_CHECK_ARG(dev && desc && init && out, "vfps_step_create: NULL argument");
    VFPS_CHECK_ARG(desc->scheme <= 1, "vfps_step_create: scheme must be 0 or 1");
    VFPS_CHECK_ARG(!desc->use_wake || zwake, "vfps_step_create: wake impedance");
    VFPS_CHECK_ARG(!desc->use_csr || zcsr, "vfps_step_create: CSR impedance");
    VFPS_CHECK_ARG(desc->wake_method <= 2,
                   "vfps_step_create: wake_method must be 0, 1 or 2");
    VFPS_CHECK_ARG(desc->threads == 0 ||
                       (desc->threads % 64 == 0 && desc->threads <= STEP_MAX_THREADS),
                   "vfps_step_create: threads must be a multiple of 64 up to 512");
    VFPS_TRY_BEGIN
    auto* s = new vfps_step();
    std::unique_ptr<vfps_step, void (*)(vfps_step*)> guard(s, step_free);
    s->dev = dev;
    s->desc = *desc;
    s->fused = desc->fused != 0;
    s->direct = s->fused && desc->wake_method != 1;
    VFPS_CHECK_ARG(desc->renormalize == 0 || s->direct,
                   "vfps_step_create: charge renormalisation needs the fused "
                   "step with the direct wake");
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
                           zcsr->n % 2 == 0,
                       "vfps_step_create: CSR impedance length must be even, "
                       "at least the mesh size and at most 8192");
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
    VFPS_HIP(hipMalloc(&s->counter, sizeof(uint32_t) * 2));
    VFPS_HIP(hipMalloc(&s->err, sizeof(uint32_t)));
    VFPS_HIP(hipHostMalloc(&s->err_h, sizeof(uint32_t), hipHostMallocDefault));
    *s->err_h = 0;
    VFPS_HIP(hipMalloc(&s->csr_power, sizeof(float)));
    VFPS_HIP(hipMemset(s->filling, 0, sizeof(double) * 2));
    VFPS_HIP(hipMemset(s->counter, 0, sizeof(uint32_t) * 2));
    VFPS_HIP(hipMemset(s->err, 0, sizeof(uint32_t)));
    VFPS_HIP(hipMemset(s->csr_power, 0, sizeof(float)));
    VFPS_HIP(hipMalloc(&s->inst_d, sizeof(InstDesc)));
    VFPS_HIP(hipMalloc(&s->fparams_d, sizeof(FieldParams)));
    VFPS_HIP(hipSetDevice(dev->ordinal));
    VFPS_CHECK_ARG(field_lds_bytes(s->M) <= dev->lds_max,
                   "vfps_step_create: FFT length exceeds the LDS of one CU");
    if (s->direct && (rc = step_setup_direct(s)) != VFPS_OK) return rc;
    // fused kernel variant, tile and workgroup
    s->variant.scheme = (int)desc->scheme;
    s->variant.it = it;
    s->variant.ipfp = s->fpm ? s->fpm->ip : 0;
    s->variant.direct = s->direct;
    s->variant.clamp = desc->rot_clamp != 0;
    if (s->fused) {
        if (!step_kernel_for(s->variant, STEP_SINGLE, 1)) {
            return fail(VFPS_ERR_UNSUPPORTED,
                        "vfps_step_create: fused step needs linear or cubic "
                        "interpolation (use fused = 0)");
        }
        // 512-thread workgroups from 256^2 on measured fastest on one mesh
        // (256^2: 6.4 against 7.4 us/step with 256 threads)
        uint32_t thr = desc->threads ? desc->threads : (n >= 256 ? 512u : 256u);
        if (ceil_div(n, thr) > (uint32_t)STEP_MAX_COLS) {
            return fail(VFPS_ERR_INVALID, "vfps_step_create: too few threads "
                                          "for the mesh size");
        }
        s->threads = thr;
        // two-row tiles from 384^2 on measured fastest on one mesh
        s->T = desc->tile_rows ? desc->tile_rows : (n >= 384 ? 2u : 1u);
        VFPS_CHECK_ARG(s->T <= thr, "vfps_step_create: tile_rows > threads");
    }
    if ((rc = step_upload_desc(s)) != VFPS_OK) return rc;
    if (s->fused) {
        while (s->T > 1 &&
               step_lds_bytes(s->variant, n, s->T, s->stride, true) > dev->lds_max) {
            s->T >>= 1;
            if ((rc = step_upload_desc(s)) != VFPS_OK) return rc;
        }
        const size_t lds = step_lds_bytes(s->variant, n, s->T, s->stride, true);
        VFPS_CHECK_ARG(lds <= dev->lds_max,
                       "vfps_step_create: mesh row too long for the LDS of one "
                       "CU (use fused = 0)");
        if ((rc = step_set_lds_limit(s->variant, lds)) != VFPS_OK) return rc;
    }
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    *out = guard.release();
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_toeplitz_kernels(uint32_t nmax, uint32_t n, const vfps_cplx* z,
                          const vfps_cplx* zc, const float* renorm,
                          float fdelta, double* g, double* h)
{
    VFPS_CHECK_ARG(nmax >= 4 && nmax % 2 == 0 && n >= 1 && n <= nmax / 2,
                   "vfps_toeplitz_kernels: need an even nmax >= 2 n");
    VFPS_CHECK_ARG(!g || z, "vfps_toeplitz_kernels: g needs z");
    VFPS_CHECK_ARG(!h || (zc && renorm), "vfps_toeplitz_kernels: h needs zc "
                                         "and renorm");
    VFPS_TRY_BEGIN
    std::vector<double> gv, hv;
    toeplitz_tables(n, nmax / 2, reinterpret_cast<const float2*>(z),
                    reinterpret_cast<const float2*>(zc), renorm, fdelta,
                    g ? &gv : nullptr, h ? &hv : nullptr);
    if (g) std::copy(gv.begin(), gv.end(), g);
    if (h) std::copy(hv.begin(), hv.end(), h);
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
    int rc = step_flush_desc(s);
    if (rc != VFPS_OK) return rc;
    bool closed = false;
    if (!s->desc.no_graph) {
        // captured runs: full ones of STEP_GRAPH steps (even, so the parity
        // stays), then one of the remaining steps; the last one also closes
        // the run (fused steps)
        const uint32_t p = s->fused ? (s->parity & 1u) : 0u;
        while (left > 0) {
            const uint32_t k = left >= STEP_GRAPH ? STEP_GRAPH : left;
            hipGraphExec_t g = nullptr;
            if ((rc = step_graph(s, k, p, k == left, &g)) != VFPS_OK) return rc;
            VFPS_HIP(hipGraphLaunch(g, st));
            left -= k;
            if (s->fused && (k & 1u)) s->parity ^= 1u;
        }
        closed = s->fused && nsteps > 0;
    }
    while (left > 0) {
        if ((rc = step_launch_one(s, s->fused ? s->parity : 0)) != VFPS_OK)
            return rc;
        if (s->fused) s->parity ^= 1u;
        left--;
    }
    if (!closed && s->fused && nsteps > 0) {
        // the same close, launched from here
        if ((rc = launch_step_finalize(s->inst_d, 1, s->parity, st, s->err_h,
                                       s->direct)) != VFPS_OK)
            return rc;
        closed = true;
    }
    if (!closed) {
        // the error flag comes back in the same stream: one wait for both
        VFPS_HIP(hipMemcpyAsync(s->err_h, s->err, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, st));
    }
    VFPS_HIP(hipStreamSynchronize(st));
    const uint32_t err = *s->err_h;
    if (err) {
        return fail(VFPS_ERR_INVALID, "fused step: a tap fell outside the "
                                      "halo rows (internal error)");
    }
    return VFPS_OK;
}

int vfps_step_prepare(vfps_step* s, uint32_t nsteps)
{
    VFPS_CHECK_ARG(s, "vfps_step_prepare: NULL step");
    int rc = step_flush_desc(s);
    if (rc != VFPS_OK) return rc;
    if (s->desc.no_graph || nsteps == 0) {
        return VFPS_OK;
    }
    // the graphs vfps_step_run(s, nsteps) replays: full ones that do not end
    // the run, and the one that ends it
    const uint32_t p = s->fused ? (s->parity & 1u) : 0u;
    const uint32_t full = nsteps / STEP_GRAPH, rem = nsteps % STEP_GRAPH;
    hipGraphExec_t g = nullptr;
    if ((full > 1 || (full == 1 && rem)) &&
        (rc = step_graph(s, STEP_GRAPH, p, false, &g)) != VFPS_OK) {
        return rc;
    }
    if ((rc = step_graph(s, rem ? rem : STEP_GRAPH, p, true, &g)) != VFPS_OK) {
        return rc;
    }
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
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
    VFPS_HIP(hipMemset(s->counter, 0, sizeof(uint32_t) * 2));
    if (s->csr_valid) {
        VFPS_HIP(hipMemset(s->csr_valid, 0xFF, sizeof(uint32_t) * 2));
    }
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
    const uint32_t at = s->fused ? (s->parity & 1u) : 0u;
    VFPS_HIP(hipMemcpy(&c, s->counter + at, sizeof(uint32_t),
                       hipMemcpyDeviceToHost));
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
    const bool odd = s->fused && (s->parity & 1u);
    const vfps_mesh* cur = odd ? s->t2 : s->t1;
    if (mesh) {
        VFPS_HIP(hipMemcpy(mesh, cur->data, sizeof(float) * n * n,
                           hipMemcpyDeviceToHost));
    }
    if (proj) {
        VFPS_HIP(hipMemcpy(proj, cur->proj, sizeof(float) * n,
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

int vfps_step_set_particles(vfps_step* s, const float* pos, uint32_t n,
                            int fptrack, uint64_t seed)
{
    VFPS_CHECK_ARG(s && (pos || n == 0), "vfps_step_set_particles: NULL "
                                         "argument");
    VFPS_CHECK_ARG(fptrack >= 0 && fptrack <= 3,
                   "vfps_step_set_particles: FPTracking");
    VFPS_CHECK_ARG(!(s->fused && fptrack == 2),
                   "vfps_step_set_particles: FPTracking approximation2 reads "
                   "the mesh between drift and Fokker-Planck, which only the "
                   "per-map step (fused = 0) keeps");
    VFPS_TRY_BEGIN
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    hipFree(s->particles);
    s->particles = nullptr;
    s->npart = 0;
    TrackProgram prog{};
    prog.seed = seed;
    auto add = [&](const vfps_map* m, const float* off) {
        if (m && track_op(m, off, &prog.op[prog.nops])) {
            prog.nops++;
        }
    };
    if (s->desc.use_wake) add(s->wm, s->wfield->wake);
    if (s->desc.scheme == 0) {
        add(s->rfm, nullptr);
        add(s->drm, nullptr);
    } else {
        add(s->rot, nullptr);
    }
    if (s->fpm) {
        add(s->fpm, nullptr);
        prog.op[prog.nops - 1].fptrack = fptrack;
    }
    s->track = prog;
    if (n) {
        VFPS_HIP(hipMalloc(&s->particles, sizeof(float2) * n));
        VFPS_HIP(hipMemcpy(s->particles, pos, sizeof(float2) * n,
                           hipMemcpyHostToDevice));
        s->npart = n;
    }
    // the captured steps gain (or lose) the tracking launch
    return step_drop_graph(s);
    VFPS_TRY_END
}

int vfps_step_get_particles(vfps_step* s, float* pos, uint32_t n)
{
    VFPS_CHECK_ARG(s && pos && n == s->npart,
                   "vfps_step_get_particles: bad argument");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    if (n) {
        VFPS_HIP(hipMemcpy(pos, s->particles, sizeof(float2) * n,
                           hipMemcpyDeviceToHost));
    }
    return VFPS_OK;
}

int vfps_step_moments(vfps_step* s, float* xproj, float* yproj,
                      float* population, float* position, float* length,
                      float* average, float* spread)
{
    VFPS_CHECK_ARG(s, "vfps_step_moments: NULL step");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    vfps_mesh* cur = (s->fused && (s->parity & 1u)) ? s->t2 : s->t1;
    double fill = 0.0;
    // src/main.cpp:961-967: integrate(); variance(0); updateYProjection();
    // variance(1)
    int rc = vfps_mesh_integrate(cur, &fill, nullptr);
    if (rc != VFPS_OK) return rc;
    float mean = 0.0f, var = 0.0f, rms = 0.0f;
    if ((rc = vfps_mesh_variance(cur, 0, &mean, &var, &rms)) != VFPS_OK) {
        return rc;
    }
    if (position) *position = mean;
    if (length) *length = rms;
    if ((rc = vfps_mesh_project_y(cur)) != VFPS_OK) return rc;
    if ((rc = vfps_mesh_variance(cur, 1, &mean, &var, &rms)) != VFPS_OK) {
        return rc;
    }
    if (average) *average = mean;
    if (spread) *spread = rms;
    if (population) *population = (float)fill;
    if (xproj && (rc = vfps_mesh_get_projection(cur, 0, xproj)) != VFPS_OK) {
        return rc;
    }
    if (yproj && (rc = vfps_mesh_get_projection(cur, 1, yproj)) != VFPS_OK) {
        return rc;
    }
    return VFPS_OK;
}

int vfps_step_set_mesh(vfps_step* s, const float* mesh)
{
    VFPS_CHECK_ARG(s && mesh, "vfps_step_set_mesh: NULL argument");
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    VFPS_HIP(hipMemcpy(s->t1->data, mesh, sizeof(float) * s->t1->cells(),
                       hipMemcpyHostToDevice));
    if (s->fused) {
        int rc0 = step_reset_pending(s);
        if (rc0 != VFPS_OK) return rc0;
    }
    s->parity = 0;
    int rc = launch_project_x(s->t1, s->t1->data, s->t1->proj);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(s->dev->stream));
    return VFPS_OK;
}

int vfps_step_config(vfps_step* s, uint32_t* out, uint32_t nout)
{
    VFPS_CHECK_ARG(s && out && nout >= 6, "vfps_step_config: bad argument");
    out[0] = s->fused ? 1u : 0u;
    out[1] = s->direct ? 2u : 1u;
    out[2] = s->T;
    out[3] = s->threads;
    out[4] = s->stride;
    out[5] = s->fused ? (uint32_t)step_lds_of(s) : 0u;
    if (nout >= 7) {
        out[6] = s->persist ? 1u : 0u;
    }
    if (nout >= 8) {
        out[7] = s->strip;
    }
    return VFPS_OK;
}

int vfps_step_phases(vfps_step* s, double* ns_out, uint32_t nout)
{
    VFPS_CHECK_ARG(s && ns_out && nout >= 16, "vfps_step_phases: bad argument");
    VFPS_CHECK_ARG(s->fused, "vfps_step_phases: needs the fused step");
    hipStream_t st = s->dev->stream;
    VFPS_HIP(hipStreamSynchronize(st));
    if (int rc0 = step_flush_desc(s); rc0 != VFPS_OK) return rc0;
    const uint32_t nblk = s->tiles();
    const size_t nstamp = 16 + 2 * (size_t)nblk;
    unsigned long long* buf = nullptr;
    VFPS_HIP(hipMalloc(&buf, sizeof(unsigned long long) * nstamp));
    VFPS_HIP(hipMemset(buf, 0, sizeof(unsigned long long) * nstamp));
    // point both descriptors at the stamp buffer for one step
    unsigned long long* mstamp = buf;
    unsigned long long* fstamp = buf + 8;
    VFPS_HIP(hipMemcpy(reinterpret_cast<char*>(s->inst_d) + offsetof(InstDesc, tstamp),
                       &mstamp, sizeof(mstamp), hipMemcpyHostToDevice));
    s->inst_h.tstamp = mstamp;
    VFPS_HIP(hipMemcpy(reinterpret_cast<char*>(s->fparams_d) + offsetof(FieldParams, tstamp),
                       &fstamp, sizeof(fstamp), hipMemcpyHostToDevice));
    s->fparams_h.tstamp = fstamp;
    int rc = step_launch_one(s, s->parity);
    s->parity ^= 1u;
    hipError_t e = hipStreamSynchronize(st);
    std::vector<unsigned long long> h(nstamp, 0);
    hipMemcpy(h.data(), buf, sizeof(unsigned long long) * nstamp,
              hipMemcpyDeviceToHost);
    unsigned long long* nul = nullptr;
    hipMemcpy(reinterpret_cast<char*>(s->inst_d) + offsetof(InstDesc, tstamp),
              &nul, sizeof(nul), hipMemcpyHostToDevice);
    s->inst_h.tstamp = nullptr;
    hipMemcpy(reinterpret_cast<char*>(s->fparams_d) + offsetof(FieldParams, tstamp),
              &nul, sizeof(nul), hipMemcpyHostToDevice);
    s->fparams_h.tstamp = nullptr;
    hipFree(buf);
    if (rc != VFPS_OK) return rc;
    if (e != hipSuccess) return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    int khz = 0;
    VFPS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                                   s->dev->ordinal));
    const double ns_per_tick = khz > 0 ? 1e6 / khz : 10.0;
    // [1..7]: step kernel phase ends (tile 0) relative to its start;
    // [9..13]: field kernel phase ends relative to its start; 15: tick ns;
    // [16..21] (nout >= 24): step workgroups: first entry after the field
    // kernel's last stamp, spread of entries, shortest, median and longest
    // workgroup, last exit after first entry
    for (uint32_t i = 0; i < nout; i++) ns_out[i] = 0;
    for (int i = 1; i < 8; i++) {
        if (h[i] && h[0]) ns_out[i] = (double)(h[i] - h[0]) * ns_per_tick;
    }
    for (int i = 9; i < 14; i++) {
        if (h[i] && h[8]) ns_out[i] = (double)(h[i] - h[8]) * ns_per_tick;
    }
    ns_out[15] = ns_per_tick;
    if (nout >= 24) {
        unsigned long long e0 = ~0ull, e1 = 0, x1 = 0;
        std::vector<double> dur;
        for (uint32_t b = 0; b < nblk; b++) {
            const unsigned long long a = h[16 + 2 * b], z = h[17 + 2 * b];
            if (!a || !z) continue;
            e0 = std::min(e0, a);
            e1 = std::max(e1, a);
            x1 = std::max(x1, z);
            dur.push_back((double)(z - a) * ns_per_tick);
        }
        if (!dur.empty()) {
            std::sort(dur.begin(), dur.end());
            const unsigned long long fend = h[13];
            ns_out[16] = fend ? ((double)e0 - (double)fend) * ns_per_tick : 0;
            ns_out[17] = (double)(e1 - e0) * ns_per_tick;
            ns_out[18] = dur.front();
            ns_out[19] = dur[dur.size() / 2];
            ns_out[20] = dur.back();
            ns_out[21] = (double)(x1 - e0) * ns_per_tick;
            // [24 + 2 t], [25 + 2 t] (nout >= 24 + 2 tiles): entry and exit
            // of tile t after the first entry
            if (nout >= 24 + 2 * (size_t)nblk) {
                for (uint32_t b = 0; b < nblk; b++) {
                    const unsigned long long a = h[16 + 2 * b], z = h[17 + 2 * b];
                    if (!a || !z) continue;
                    ns_out[24 + 2 * b] = (double)(a - e0) * ns_per_tick;
                    ns_out[25 + 2 * b] = (double)(z - e0) * ns_per_tick;
                }
            }
        }
    }
    return VFPS_OK;
}

int vfps_step_timing(vfps_step* s, uint32_t nsteps, float* ms_wake,
                     float* ms_maps, float* ms_step)
{
    VFPS_CHECK_ARG(s && nsteps > 0, "vfps_step_timing: bad argument");
    VFPS_CHECK_ARG(s->fused, "vfps_step_timing: per-kernel timing needs the "
                             "fused step");
    if (int rc0 = step_flush_desc(s); rc0 != VFPS_OK) return rc0;
    hipStream_t st = s->dev->stream;
    hipEvent_t e0, e1, e2;
    VFPS_HIP(hipEventCreate(&e0));
    VFPS_HIP(hipEventCreate(&e1));
    VFPS_HIP(hipEventCreate(&e2));
    double tw = 0, tm = 0;
    int rc = VFPS_OK;
    if (s->direct) {
        // one kernel per step: nsteps launches back to back between two
        // events, so the figure is the kernel's own average duration
        hipEventRecord(e0, st);
        for (uint32_t k = 0; k < nsteps && rc == VFPS_OK; k++) {
            rc = launch_step_kernel(s->variant, s->inst_d, nullptr, s->tiles(),
                                    step_lds_bytes(s->variant, s->desc.n, s->T,
                                                   s->stride, true),
                                    s->threads, s->desc.n, s->parity, st,
                                    &s->inst_h);
            s->parity ^= 1u;
        }
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        tm = ms;
    } else {
        // field kernel and step kernel of each step between three events
        for (uint32_t k = 0; k < nsteps && rc == VFPS_OK; k++) {
            const uint32_t par = s->parity;
            hipEventRecord(e0, st);
            rc = launch_fields(s->desc.use_wake != 0, s->desc.use_csr != 0,
                               s->fparams_d, 1, s->M, par, st, &s->fparams_h);
            hipEventRecord(e1, st);
            if (rc == VFPS_OK) {
                rc = launch_step_kernel(s->variant, s->inst_d, nullptr,
                                        s->tiles(),
                                        step_lds_bytes(s->variant, s->desc.n, s->T,
                                                   s->stride, true),
                                        s->threads, s->desc.n, par, st,
                                        &s->inst_h);
            }
            hipEventRecord(e2, st);
            hipEventSynchronize(e2);
            float a = 0, b = 0;
            hipEventElapsedTime(&a, e0, e1);
            hipEventElapsedTime(&b, e1, e2);
            tw += a;
            tm += b;
            s->parity ^= 1u;
        }
    }
    if (s->direct && rc == VFPS_OK) {
        rc = launch_step_finalize(s->inst_d, 1, s->parity, st);
        hipStreamSynchronize(st);
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
