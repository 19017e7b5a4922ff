// The fused time step of src/main.cpp:948-1053: one kernel per step.
//
// A workgroup owns a tile of T output rows of one mesh (one instance) and
// holds, in LDS, the input rows the tile depends on (the tile plus the halo
// rows the drift or rotation reads).  It runs, in order:
//   A  every global read: the input rows, per-row kick offsets, per-column
//      drift/Fokker-Planck/Simpson tables and, for the direct wake, the
//      previous x projection and the window of the wake kernel the tile's
//      rows need;
//   B  (direct wake) the wake offsets of its rows as the Toeplitz product of
//      the previous profile with the kernel g (FP64), and its share of the
//      CSR power as the quadratic form p^T H p;
//   C  wake kick along y (WakePotentialMap::apply);
//   D  RF kick along y (scheme 0, RFKickMap::apply);
//   E  drift along x (scheme 0, DriftMap::apply) or rotation (scheme 1);
//   F  Fokker-Planck along y, the store of the output rows, and the
//      products mesh*weight of the x projection;
//   G  the x projection of its rows (PhaseSpace::updateXProjection), summed
//      in std::inner_product order;
//   H  (direct wake) every tile stores its CSR share; tile 0 (a one-mesh
//      launch: a workgroup of its own) counts the step, computes the FP64
//      charge integral and sums the previous step's shares in a fixed order
//      (the launch boundary orders them).
// The mesh makes one trip HBM -> LDS -> HBM per step.  The state buffers
// and the x projections ping-pong (`parity`), so a step reads only what
// the previous launch wrote.
//
// Every map keeps the reference's float arithmetic and operation order, so
// with the same wake offsets the output equals the per-map kernels bit for
// bit.  The direct wake differs from the FFT of the reference only by
// rounding: g[d] = Re Z_0 + 2 sum_{k=1}^{M-1} Re(Z_k e^{2 pi i k d/N}) is
// the exact impulse response of "zero-pad, R2C, times Z, C2R" of
// src/PS/ElectricField.cpp:215-241 (Nyquist bin zero, Im W_0 ignored by
// the C2R), and the product is accumulated in FP64 and rounded to float
// where FFTW's output is float.
#pragma once

#include "common.hpp"
#include "maps.hpp"
#include "reduce.hpp"
#include "step.hpp"

namespace vfps_impl {

// LDS layout of the step kernel (offsets in 4-byte words; the FP64 arrays
// come first so they stay 8-byte aligned)
struct StepLds {
    uint32_t gwin, hwin;          // FP64 kernel windows (word offsets)
    uint32_t g0win;               // FP64 kernel window of row 0 (g[0..n))
    uint32_t pprev;               // previous x projection
    uint32_t row0;                // input row 0 (scheme 1 out-of-range taps)
    uint32_t bufA, bufB;          // row buffers, `stride` rows each
    uint32_t tile_out;            // [T][n + 4] projection products, or
                                  // STEP_TILE_IN_BUF: in the row buffer the
                                  // drift/rotation read (dead by then)
    uint32_t offw, offr;          // per-row kick offsets
    uint32_t tapw, tapr;          // per-row taps: IT shifts, then IT weights
    uint32_t csr_rows;            // FP64 CSR share of each output row
    uint32_t fout;                // [T][n] output rows (one-mesh form),
                                  // stored to global memory during the
                                  // projection
    uint32_t words;
};

__host__ __device__ inline uint32_t step_round4(uint32_t v) { return (v + 3u) & ~3u; }

// The direct form's dot products (one per wake row, CSR row and, for the
// rotation, row 0) are each summed by one wavefront in a fixed order that
// depends on neither the tile height nor the workgroup size: lane l takes
// x' = l, l + 64, ... in increasing order, then wave_sum_dpp over the lanes.
// LDS words of one row's kick taps
constexpr uint32_t STEP_TAP_WORDS = 8;


constexpr uint32_t STEP_TILE_IN_BUF = 0xFFFFFFFFu;

// scheme 1 (rotation) with the direct wake also keeps input row 0 and the
// kernel window of its wake offset.  windows: the FP64 kernel windows live
// in LDS (the one-mesh form, where they load with the mesh); batches read
// them from global memory (tables in L2 shared by every tile) and spend
// the LDS on occupancy instead.
__host__ __device__ inline StepLds step_lds(uint32_t n, uint32_t T, uint32_t stride,
                                            bool direct, uint32_t it,
                                            int scheme, bool windows)
{
    StepLds L{};
    uint32_t w = 0;
    if (direct) {
        if (windows) {
            L.gwin = w;
            w += 2 * (n + stride);
            L.hwin = w;
            w += 2 * (n + T);
        }
        L.csr_rows = w;
        w += 2 * T;
        if (scheme == 1 && windows) {
            L.g0win = w;
            w += 2 * n;
        }
        w = step_round4(w);
        L.pprev = w;
        w += step_round4(n);
        if (scheme == 1) {
            L.row0 = w;
            w += step_round4(n);
        }
    }
    L.bufA = w;
    w += step_round4(stride * n);
    L.bufB = w;
    w += step_round4(stride * n);
    if ((size_t)stride * n >= (size_t)T * (n + 4)) {
        L.tile_out = STEP_TILE_IN_BUF;
    } else {
        L.tile_out = w;
        w += T * (n + 4);
    }
    L.offw = w;
    w += step_round4(stride);
    L.offr = w;
    w += step_round4(stride);
    L.tapw = w;
    w += STEP_TAP_WORDS * stride;
    L.tapr = w;
    w += STEP_TAP_WORDS * stride;
    if (windows) {
        w = step_round4(w);
        L.fout = w;
        w += step_round4(T * n);
    }
    L.words = w;
    return L;
}

// global-memory views of pointers read from a descriptor (without them the
// compiler sees generic pointers and emits flat accesses)
typedef float vec4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vec4f gfloat4;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) uint32_t guint;

// most mesh columns one thread handles
constexpr int STEP_MAX_COLS = 4;
// float4 loads one thread keeps in flight while filling the tile
constexpr int STEP_LOAD_BATCH = 8;
// float4 loads of the input rows per thread issued before the wake products
// (one-mesh launches)
constexpr int STEP_EARLY_LOADS = 2;
// largest workgroup
constexpr uint32_t STEP_MAX_THREADS = 512;
// block-map entry of a padding workgroup
constexpr uint32_t STEP_NO_TILE = 0xFFFFFFFFu;

#define VFPS_STAMP(k)                                                         \
    do {                                                                      \
        if (d.tstamp && tile == 0 && inst == 0) {                             \
            __syncthreads();                                                  \
            if (threadIdx.x == 0) d.tstamp[k] = wall_clock64();               \
        }                                                                     \
    } while (0)

// Taps of one kick row (KickMap::updateSM, src/SM/KickMap.cpp:349-386):
// STEP_TAP_WORDS words per row, the IT weights (zero-padded to 4) then the
// shift of tap 0 relative to the output sample.  The taps of a row are
// consecutive, index jd-(IT-1)/2 .. jd+IT/2; a tap the reference sends to
// index n/2 (out of range) has weight zero, so giving it the consecutive
// index instead adds the same +0 to the sum.

template <int IT>
__device__ __forceinline__ void store_row_taps(float off, uint32_t n, float* taps)
{
    uint32_t idx[IT];
    float w[IT];
    kick_taps<IT>(off, n, idx, w);
    const float poffs = (float)(n / 2) + off;
    float qi;
    vf_modf(poffs, &qi);
    const uint32_t jd = vf_f2u(qi);
    const int32_t base = (int32_t)jd - (IT - 1) / 2 - (int32_t)(n / 2);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        taps[j] = j < IT ? w[j] : 0.0f;
    }
    taps[4] = __int_as_float(base);
}

// The NV consecutive LDS words s[lo .. lo + NV) (lo >= 0, s 16-byte
// aligned), handed to use(v) as v[0 .. NV), for a wavefront whose lanes all
// have lo % 4 == sft (sft in an SGPR): 16-byte loads from s[lo - sft], and
// use is inlined once per shift, reading the loaded registers in place.  A
// kick's lanes start their windows 4 or 8 words apart, so NV one-word reads
// hit a few banks many times over; the 16-byte reads of 8 neighbouring
// lanes spread over the banks instead.  Reads up to s[lo + 4 NQ - 1], at
// most 6 words past the window.
template <int NV, typename F>
__device__ __forceinline__ void lds_window16(const float* s, int32_t lo,
                                             uint32_t sft, F&& use)
{
    constexpr int NQ = (NV + 6) / 4;  // sft + NV <= 4 NQ
    const vec4f* q = reinterpret_cast<const vec4f*>(s + (lo - (int32_t)sft));
    float w[4 * NQ];
#pragma unroll
    for (int j = 0; j < NQ; j++) {
        const vec4f t = q[j];
        w[4 * j] = t.x;
        w[4 * j + 1] = t.y;
        w[4 * j + 2] = t.z;
        w[4 * j + 3] = t.w;
    }
    switch (sft) {
    case 0:
        use(w);
        break;
    case 1:
        use(w + 1);
        break;
    case 2:
        use(w + 2);
        break;
    default:
        use(w + 3);
        break;
    }
}

// KickMap::apply along y (src/SM/KickMap.cpp:244-267) on local rows
// [r0, r1) of a row buffer: each thread makes four consecutive outputs of
// a row from IT+3 consecutive inputs, taps in order (n % 4 == 0).
template <int IT>
__device__ __forceinline__ void kick_rows(const float* __restrict__ src,
                                          float* __restrict__ dst, uint32_t n,
                                          int r0, int r1,
                                          const float* __restrict__ taps)
{
    if (r1 <= r0) {
        return;
    }
    const uint32_t bs = blockDim.x;
    const uint32_t ng = n >> 2;
    const uint32_t drow = bs / ng, dg = bs % ng;
    uint32_t row = (uint32_t)r0 + threadIdx.x / ng;
    uint32_t g = threadIdx.x % ng;
    while (row < (uint32_t)r1) {
        const float* tp = taps + row * STEP_TAP_WORDS;
        const vec4f w4 = *reinterpret_cast<const vec4f*>(tp);
        const int32_t base = __float_as_int(tp[4]);
        const float* s = src + (size_t)row * n;
        const int32_t y0 = (int32_t)(4 * g);
        float v[IT + 3];
#pragma unroll
        for (int k = 0; k < IT + 3; k++) {
            const uint32_t ys = (uint32_t)(y0 + base + k);
            v[k] = ys < n ? s[ys] : 0.0f;
        }
        // an input out of range is skipped by the reference; here it is a
        // zero, whose product adds +0 or -0 to the sum: the same result
        vec4f o;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < IT; j++) {
                acc = acc + v[e + j] * w4[j];
            }
            o[e] = acc;
        }
        *reinterpret_cast<vec4f*>(dst + (size_t)row * n + 4 * g) = o;
        g += dg;
        row += drow;
        if (g >= ng) {
            g -= ng;
            row++;
        }
    }
}

// The wake kick and then the RF kick along y on rows [r0, r1), fused: each
// thread makes four consecutive outputs of the second kick from the IT+3
// wake-kicked samples they read, and those from 2 IT + 2 consecutive
// inputs.  Every wake-kicked sample is the same sum, in the same order, as
// kick_rows makes it (a sample out of range reads as zero in both), so the
// result equals the two kicks one after the other bit for bit, without the
// intermediate rows in LDS or the barrier between the two.
template <int IT>
__device__ __forceinline__ void kick2_rows(const float* __restrict__ src,
                                           float* __restrict__ dst, uint32_t n,
                                           int r0, int r1,
                                           const float* __restrict__ taps1,
                                           const float* __restrict__ taps2)
{
    if (r1 <= r0) {
        return;
    }
    const uint32_t bs = blockDim.x;
    const uint32_t ng = n >> 2;
    const uint32_t drow = bs / ng, dg = bs % ng;
    uint32_t row = (uint32_t)r0 + threadIdx.x / ng;
    uint32_t g = threadIdx.x % ng;
    while (row < (uint32_t)r1) {
        const float* t1 = taps1 + row * STEP_TAP_WORDS;
        const float* t2 = taps2 + row * STEP_TAP_WORDS;
        const vec4f w1 = *reinterpret_cast<const vec4f*>(t1);
        const vec4f w2 = *reinterpret_cast<const vec4f*>(t2);
        const int32_t base1 = __float_as_int(t1[4]);
        const int32_t base2 = __float_as_int(t2[4]);
        const float* s = src + (size_t)row * n;
        const int32_t y0 = (int32_t)(4 * g);
        float v[2 * IT + 2];
#pragma unroll
        for (int k = 0; k < 2 * IT + 2; k++) {
            const uint32_t ys = (uint32_t)(y0 + base2 + base1 + k);
            v[k] = ys < n ? s[ys] : 0.0f;
        }
        float c[IT + 3];
#pragma unroll
        for (int k = 0; k < IT + 3; k++) {
            const uint32_t yc = (uint32_t)(y0 + base2 + k);
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < IT; j++) {
                acc = acc + v[k + j] * w1[j];
            }
            c[k] = yc < n ? acc : 0.0f;
        }
        vec4f o;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < IT; j++) {
                acc = acc + c[e + j] * w2[j];
            }
            o[e] = acc;
        }
        *reinterpret_cast<vec4f*>(dst + (size_t)row * n + 4 * g) = o;
        g += dg;
        row += drow;
        if (g >= ng) {
            g -= ng;
            row++;
        }
    }
}

// Workgroup-shared scalars of one tile
struct StepShared {
    double red[16];
    float s00;
    float norm;   // renormalisation factor of this step (1: none)
    uint32_t slot;
    uint32_t cnt;
};

// Close a step of the direct form: count it, integrate the profile that
// entered it (FP64, Simpson weights) and sum the previous launch's CSR
// shares in a fixed order.  It reads only what earlier launches wrote, so
// no workgroup of the launch waits for another.  cnt: the steps completed
// before this one; red: 16 doubles of LDS.  Uniform over the workgroup.
__device__ __forceinline__ void step_close(const InstDesc& d, uint32_t parity,
                                           uint32_t cnt, double* red)
{
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t n = d.n;
    const bool csr = d.use_csr != 0;
    const gfloat* pin = (const gfloat*)d.proj[parity];
    const gfloat* ws = (const gfloat*)d.ws;
    double f = 0.0, p = 0.0;
    for (uint32_t x = tid; x < n; x += bs) {
        f += (double)pin[x] * (double)ws[x];
    }
    const bool prev = csr && cnt >= 1 &&
                      ((const guint*)d.csr_valid)[parity ^ 1u] == cnt;
    if (prev) {
        const gdouble* shares =
            (const gdouble*)d.csr_part + (size_t)(parity ^ 1u) * n;
        for (uint32_t k = tid; k < d.tiles; k += bs) {
            p += shares[k];
        }
    }
    block_sum2(f, p, red);
    if (tid == 0) {
        if (prev) {
            *d.csr_power = (float)p;
            if (d.csr_hist && cnt - 1u < d.hist_cap) {
                d.csr_hist[cnt - 1u] = (float)p;
            }
        }
        if (d.filling) {
            d.filling[0] = f;
            d.filling[1] = f;
        }
        if (d.counter) {
            d.counter[parity ^ 1u] = cnt + 1u;
        }
        d.csr_valid[parity] = cnt + 1u;
    }
}

// One row tile of one instance: phases A-H.  Every early return is uniform
// over the workgroup.  MAXC: most mesh columns one thread handles
// (ceil(n / threads) of every instance of the launch), which sizes the
// per-column register tables.  ONE_MESH trades registers for latency, as
// suits a launch of one mesh (a batch keeps its occupancy instead): the
// first float4 loads of the input rows are held in registers through
// phase B, the kernel windows load into LDS, and the projection's adds
// keep their loads scheduled ahead.
template <int IT, int IPFP, bool DIRECT, int SCHEME, bool CLAMP, int MAXC,
          bool ONE_MESH>
__device__ __forceinline__ void step_tile(const InstDesc& d, uint32_t inst,
                                          uint32_t tile, uint32_t logical,
                                          uint32_t parity, float* lds,
                                          StepShared& sh)
{
    constexpr int EARLY = ONE_MESH ? STEP_EARLY_LOADS : 0;
    double* red = sh.red;
    float& s00_sh = sh.s00;
    uint32_t& slot_sh = sh.slot;
    uint32_t& cnt_sh = sh.cnt;
    const unsigned long long t_entry = wall_clock64();
    const uint32_t n = d.n;
    const uint32_t T = d.T;
    const uint32_t stride = d.stride;
    const int x0 = (int)(tile * T);
    if (x0 >= (int)n) {
        return;  // not launched: grids hold exactly the tiles
    }
    VFPS_STAMP(0);
    const int rows_out = min((int)T, (int)n - x0);
    const int hlo = (int)d.hlo;
    const int rbase = x0 - hlo;
    const int R = rows_out + hlo + (int)d.hhi;
    // rows of the buffer that exist in the mesh
    const int lr0 = max(0, -rbase);
    const int lr1 = min(R, (int)n - rbase);
    const StepLds L = step_lds(n, T, stride, DIRECT, IT, SCHEME, ONE_MESH);
    float* bufA = lds + L.bufA;
    float* bufB = lds + L.bufB;
    // placed after the drift/rotation, which frees the buffer it read
    float* tile_out = nullptr;
    float* offw = lds + L.offw;
    float* offr = lds + L.offr;
    float* tapw = lds + L.tapw;
    float* tapr = lds + L.tapr;
    double* gwin = reinterpret_cast<double*>(lds + L.gwin);
    double* hwin = reinterpret_cast<double*>(lds + L.hwin);
    double* csr_rows = reinterpret_cast<double*>(lds + L.csr_rows);
    float* pprev = lds + L.pprev;
    const gfloat* in = (const gfloat*)d.mesh[parity];
    gfloat* out = (gfloat*)d.mesh[parity ^ 1u];
    const gfloat* proj_in = (const gfloat*)d.proj[parity];
    gfloat* proj_out = (gfloat*)d.proj[parity ^ 1u];
    gfloat* prof_hist = (gfloat*)d.prof_hist;
    const gfloat* ws_g = (const gfloat*)d.ws;
    const gfloat* rf_off = (const gfloat*)d.rf_off;
    const gfloat* drift_off = (const gfloat*)d.drift_off;
    const gfloat* wake_off = (const gfloat*)d.wake_off;
    const gfloat* fp_w = (const gfloat*)d.fp_w;
    const guint* fp_i = (const guint*)d.fp_i;
    const gdouble* wake_g = (const gdouble*)d.wake_g;
    const gdouble* csr_h = (const gdouble*)d.csr_h;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6, nw = bs >> 6;
    const int32_t half = (int32_t)(n / 2);
    const bool wake = d.use_wake != 0;
    const bool csr = DIRECT && d.use_csr != 0;

    // ---- A: every global read -------------------------------------------
    //  - the first EARLY float4 of the input rows per thread go out first
    //    and stay in registers through phase B (stored to LDS after the
    //    wake products): the tables' stores below wait for them too (the
    //    load counter drains in order), but the compiler keeps them here
    //    instead of sinking them to their use after B
    const bool vec_rows = (n & 3u) == 0u;
    const size_t in_first = (size_t)(rbase + lr0) * n;
    const uint32_t in_total = (uint32_t)(lr1 - lr0) * n;
    const uint32_t in_total4 = in_total >> 2;
    const gfloat4* in4 = (const gfloat4*)(in + in_first);
    vec4f early[EARLY > 0 ? EARLY : 1];
    if (vec_rows) {
#pragma unroll
        for (int b = 0; b < EARLY; b++) {
            const uint32_t i = tid + (uint32_t)b * bs;
            early[b] = in4[i < in_total4 ? i : in_total4 - 1u];
        }
    }
    // The loads of the step counter, the previous profile, the kernel
    // windows, row 0 and the row offsets go out before any of them is
    // stored to LDS: the hardware counts outstanding loads in order, so a
    // store that waits for its own load waits for every load issued before
    // it (a load-store pair per table would cost a memory round trip each).
    // Indices are clamped into the tables, the stores bounded.
    //  - steps completed before this one (read by one thread, shared with
    //    the workgroup after the barrier that ends phase A)
    //  - previous profile and the kernel windows (direct wake): wake rows
    //    [rbase+lr0, rbase+lr1) need g[r - x' + n - 1]: gwin[k] =
    //    g[rbase + lr0 + k], k < (lr1 - lr0) + n - 1; CSR rows
    //    [x0, x0+rows_out) need hwin[k] = h[x0 + k]
    //  - row 0 of the input, for the out-of-range taps of the rotation in
    //    tiles that do not hold it
    //  - per-row kick offsets (their taps follow with the rows)
    // window values per thread: gl, hl < n + stride <= (MAXC + 1) bs, as the
    // host keeps stride <= threads (one thread per buffer row)
    constexpr int WV = MAXC + 1;
    const uint32_t g0 = (uint32_t)(rbase + lr0);
    const uint32_t gl = (uint32_t)(lr1 - lr0) + n - 1;
    const uint32_t hl = (uint32_t)rows_out + n - 1;
    const bool need_row0 = DIRECT && SCHEME == 1 && rbase > 0 && wake;
    const bool row_thread = (int)tid >= lr0 && (int)tid < lr1;
    uint32_t cnt_v = 0;
    if (tid == 0 && d.counter) {
        cnt_v = d.counter[parity];
    }
    float pv[MAXC];
    double gv[WV], hv[WV];
    float r0v[MAXC];
    double g0v[MAXC];
    if (DIRECT) {
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t x = tid + (uint32_t)c * bs;
            pv[c] = proj_in[x < n ? x : n - 1u];
        }
        if (ONE_MESH && wake) {
#pragma unroll
            for (int c = 0; c < WV; c++) {
                const uint32_t k = tid + (uint32_t)c * bs;
                gv[c] = wake_g[g0 + (k < gl ? k : gl - 1u)];
            }
        }
        if (ONE_MESH && csr) {
#pragma unroll
            for (int c = 0; c < WV; c++) {
                const uint32_t k = tid + (uint32_t)c * bs;
                hv[c] = csr_h[(uint32_t)x0 + (k < hl ? k : hl - 1u)];
            }
        }
        if (need_row0) {
#pragma unroll
            for (int c = 0; c < MAXC; c++) {
                const uint32_t y = tid + (uint32_t)c * bs;
                const uint32_t yy = y < n ? y : n - 1u;
                r0v[c] = in[yy];
                if (ONE_MESH) g0v[c] = wake_g[yy];
            }
        }
    }
    float offr_v = 0.0f, offw_v = 0.0f;
    if (row_thread) {
        const int r = rbase + (int)tid;
        if (SCHEME == 0) {
            offr_v = rf_off[r];
        }
        if (!DIRECT && wake) {
            offw_v = wake_off[r];
        }
    }
    //  - per-column tables in registers (the drift taps are made from the
    //    raw offsets once the stores below have waited for them)
    const int ncols = (int)((n + bs - 1) / bs);
    float col_ws[MAXC];
    float col_doff[MAXC];
    uint32_t col_didx[MAXC][IT];   // drift taps of the column
    float col_dw[MAXC][IT];
    uint32_t col_fi[MAXC][IPFP > 0 ? IPFP : 1];
    float col_fw[MAXC][IPFP > 0 ? IPFP : 1];
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
        const uint32_t y = tid + (uint32_t)c * bs;
        if (c < ncols && y < n) {
            if (SCHEME == 0) {
                col_doff[c] = drift_off[y];
            }
            col_ws[c] = ws_g[y];
#pragma unroll
            for (int k = 0; k < IPFP; k++) {
                col_fi[c][k] = fp_i[(size_t)k * n + y];
                col_fw[c][k] = fp_w[(size_t)k * n + y];
            }
        }
    }
    if (tid == 0) {
        cnt_sh = cnt_v;
    }
    if (DIRECT) {
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t x = tid + (uint32_t)c * bs;
            if (x < n) pprev[x] = pv[c];
        }
        if (ONE_MESH && wake) {
#pragma unroll
            for (int c = 0; c < WV; c++) {
                const uint32_t k = tid + (uint32_t)c * bs;
                if (k < gl) gwin[k] = gv[c];
            }
        }
        if (ONE_MESH && csr) {
#pragma unroll
            for (int c = 0; c < WV; c++) {
                const uint32_t k = tid + (uint32_t)c * bs;
                if (k < hl) hwin[k] = hv[c];
            }
        }
        if (need_row0) {
            float* row0 = lds + L.row0;
            double* g0win = reinterpret_cast<double*>(lds + L.g0win);
#pragma unroll
            for (int c = 0; c < MAXC; c++) {
                const uint32_t y = tid + (uint32_t)c * bs;
                if (y < n) {
                    row0[y] = r0v[c];
                    if (ONE_MESH) g0win[y] = g0v[c];
                }
            }
        }
    }
    if (row_thread) {
        if (SCHEME == 0) {
            offr[tid] = offr_v;
        }
        if (!DIRECT && wake) {
            offw[tid] = offw_v;
        }
    }
    if (SCHEME == 0) {
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t y = tid + (uint32_t)c * bs;
            if (c < ncols && y < n) {
                kick_taps<IT>(col_doff[c], n, col_didx[c], col_dw[c]);
            }
        }
        // the drift reads bufB (the kicks write its rows [lr0, lr1)): in an
        // edge tile its rows outside the mesh hold zeros, so a tap there
        // adds +0 or -0 to a sum that starts at +0, as skipping it does,
        // and the drift needs no per-tap check
        if (lr0 > 0 || lr1 < R) {
            for (uint32_t i = tid; i < (uint32_t)lr0 * n; i += bs) {
                bufB[i] = 0.0f;
            }
            for (uint32_t i = (uint32_t)lr1 * n + tid; i < (uint32_t)R * n;
                 i += bs) {
                bufB[i] = 0.0f;
            }
        }
    }
    //  - input rows [rbase+lr0, rbase+lr1) to LDS: the early registers, then
    //    the rest in batches; the kick taps of the RF (and FFT wake) rows.
    //    Without early loads this runs before the barrier that ends phase A
    //    (after phase B it would keep the batch registers live through B).
    auto fill_rows = [&]() {
        float* dst = bufA + (size_t)lr0 * n;
        if (vec_rows) {
            vec4f* a4 = reinterpret_cast<vec4f*>(dst);
#pragma unroll
            for (int b = 0; b < EARLY; b++) {
                const uint32_t i = tid + (uint32_t)b * bs;
                if (i < in_total4) a4[i] = early[b];
            }
            for (uint32_t base = (uint32_t)EARLY * bs; base < in_total4;
                 base += bs * STEP_LOAD_BATCH) {
                // unconditional loads (clamped index) keep v in registers
                vec4f v[STEP_LOAD_BATCH];
#pragma unroll
                for (int b = 0; b < STEP_LOAD_BATCH; b++) {
                    const uint32_t i = base + tid + (uint32_t)b * bs;
                    v[b] = in4[i < in_total4 ? i : in_total4 - 1u];
                }
#pragma unroll
                for (int b = 0; b < STEP_LOAD_BATCH; b++) {
                    const uint32_t i = base + tid + (uint32_t)b * bs;
                    if (i < in_total4) a4[i] = v[b];
                }
            }
        } else {
            for (uint32_t i = tid; i < in_total; i += bs) {
                dst[i] = in[in_first + i];
            }
        }
        if ((int)tid >= lr0 && (int)tid < lr1) {
            if (SCHEME == 0) {
                store_row_taps<IT>(offr[tid], n, tapr + tid * STEP_TAP_WORDS);
            }
            if (!DIRECT && wake) {
                store_row_taps<IT>(offw[tid], n, tapw + tid * STEP_TAP_WORDS);
            }
        }
    };
    if constexpr (EARLY == 0) {
        fill_rows();
    }
    __syncthreads();
    const uint32_t cnt = cnt_sh;
    // PhaseSpace::integrateAndNormalize (src/PS/PhaseSpace.cpp:342-372) on
    // the steps the reference renormalises: the charge of the profile that
    // entered the step (FP64, summed by wave 0 in a fixed order, so every
    // tile of the instance finds the same factor), then every input sample
    // times fill_set/filling before the maps.  The profile itself stays,
    // as in the reference.
    const bool renorm = DIRECT && d.norm_every != 0 && cnt % d.norm_every == 0;
    float norm = 1.0f;
    if (renorm) {
        if (wave == 0) {
            double f = 0.0;
            for (uint32_t x = lane; x < n; x += 64) {
                f += (double)pprev[x] * (double)ws_g[x];
            }
            f = wave_sum_dpp(f);
            if (lane == 0) {
                sh.norm = d.fill_set > 0.0f ? d.fill_set / (float)f : 0.0f;
            }
        }
        __syncthreads();
        norm = sh.norm;
        if (need_row0) {
            float* row0 = lds + L.row0;
            for (uint32_t y = tid; y < n; y += bs) {
                row0[y] = row0[y] * norm;
            }
        }
        __syncthreads();
    }
    VFPS_STAMP(1);

    // ---- B: direct wake offsets and the CSR share --------------------------
    // row 0 of the wake-kicked input (scheme 1 reads it for taps out of
    // range); tiles that do not hold row 0 compute it themselves
    const bool need_s00 = SCHEME == 1 && rbase > 0;
    double csr_q = 0.0;
    if (DIRECT && (wake || csr)) {
        const int nwk = wake ? lr1 - lr0 : 0;
        const int ncsr = csr ? rows_out : 0;
        const int J = nwk + ncsr + ((need_s00 && wake) ? 1 : 0);
        for (int job = (int)wave; job < J; job += (int)nw) {
            uint32_t base = 0;
            int which = 0;  // 0 wake row, 1 CSR row, 2 row 0's wake
            if (job < nwk) {
                base = (uint32_t)job;
            } else if (job < nwk + ncsr) {
                which = 1;
                base = (uint32_t)(job - nwk);
            } else {
                which = 2;
            }
            double acc = 0.0;
            if constexpr (ONE_MESH) {
                const double* win =
                    which == 0 ? gwin
                               : (which == 1 ? hwin
                                             : reinterpret_cast<const double*>(
                                                   lds + L.g0win));
#pragma unroll 4
                for (uint32_t x = lane; x < n; x += 64) {
                    acc += win[base + n - 1 - x] * (double)pprev[x];
                }
            } else {
                // the same windows, read from the tables in global memory
                const gdouble* win =
                    which == 0 ? wake_g + g0
                               : (which == 1 ? csr_h + (uint32_t)x0 : wake_g);
#pragma unroll 4
                for (uint32_t x = lane; x < n; x += 64) {
                    acc += win[base + n - 1 - x] * (double)pprev[x];
                }
            }
            acc = wave_sum_dpp(acc);
            if (lane != 0) {
                continue;
            }
            if (job < nwk) {
                const int row = lr0 + job;
                const float off = d.wake_scale * (float)acc;
                offw[row] = off;
                store_row_taps<IT>(off, n, tapw + row * STEP_TAP_WORDS);
            } else if (job < nwk + ncsr) {
                const int o = job - nwk;
                csr_rows[o] = (double)pprev[x0 + o] * acc;
            } else {
                // kick row 0 by its wake offset; s00 is its sample y = 0
                const float* row0 = lds + L.row0;
                const float off = d.wake_scale * (float)acc;
                uint32_t idx[IT];
                float w[IT];
                kick_taps<IT>(off, n, idx, w);
                float v = 0.0f;
#pragma unroll
                for (int t = 0; t < IT; t++) {
                    const uint32_t ys = (uint32_t)((int32_t)idx[t] - half);
                    if (ys < n) {
                        v = v + row0[ys] * w[t];
                    }
                }
                s00_sh = v;
            }
        }
    }
    if constexpr (EARLY > 0) {
        fill_rows();
    }
    __syncthreads();
    if (renorm) {
        // the buffer rows of the mesh, rescaled before any map reads them
        float* dst = bufA + (size_t)lr0 * n;
        const uint32_t tot = (uint32_t)(lr1 - lr0) * n;
        for (uint32_t i = tid; i < tot; i += bs) {
            dst[i] = dst[i] * norm;
        }
        __syncthreads();
    }
    if (DIRECT && csr && tid == 0) {
        for (int o = 0; o < rows_out; o++) {
            csr_q += csr_rows[o];
        }
    }
    VFPS_STAMP(2);

    float* cur = bufA;
    float* oth = bufB;
    // ---- C, D: wake kick and RF kick along y -------------------------------
    if (SCHEME == 0 && wake) {
        // both kicks in one pass (scheme 0)
        kick2_rows<IT>(cur, oth, n, lr0, lr1, tapw, tapr);
        __syncthreads();
        float* t = cur;
        cur = oth;
        oth = t;
    } else if (wake) {
        kick_rows<IT>(cur, oth, n, lr0, lr1, tapw);
        __syncthreads();
        float* t = cur;
        cur = oth;
        oth = t;
    }
    VFPS_STAMP(3);
    if constexpr (SCHEME == 0) {
        if (!wake) {
            // ---- D: RF kick along y ----------------------------------------
            kick_rows<IT>(cur, oth, n, lr0, lr1, tapr);
            __syncthreads();
            float* t = cur;
            cur = oth;
            oth = t;
        }
        VFPS_STAMP(4);
        // ---- E: drift along x, one column per thread ----------------------
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t y = tid + (uint32_t)c * bs;
            if (c >= ncols || y >= n) {
                continue;
            }
            const uint32_t* idx = col_didx[c];
            const float* w = col_dw[c];
            // fast path: every tap of every output row lies in the buffer
            // rows (those outside the mesh are zeros), so no tap needs a
            // check
            int32_t tmin = (int32_t)idx[0], tmax = (int32_t)idx[0];
#pragma unroll
            for (int j = 1; j < IT; j++) {
                tmin = min(tmin, (int32_t)idx[j]);
                tmax = max(tmax, (int32_t)idx[j]);
            }
            const int32_t xlo = x0 + tmin - half;
            const int32_t xhi = x0 + rows_out - 1 + tmax - half;
            if (xlo - rbase >= 0 && xhi - rbase < R) {
                for (int orow = 0; orow < rows_out; orow++) {
                    const int32_t xb = x0 + orow - half - rbase;
                    float v = 0.0f;
#pragma unroll
                    for (int j = 0; j < IT; j++) {
                        v = v + cur[(size_t)(xb + (int32_t)idx[j]) * n + y] * w[j];
                    }
                    oth[(size_t)orow * n + y] = v;
                }
                continue;
            }
            for (int orow = 0; orow < rows_out; orow++) {
                const uint32_t x = (uint32_t)(x0 + orow);
                float v = 0.0f;
#pragma unroll
                for (int j = 0; j < IT; j++) {
                    const uint32_t xs = (uint32_t)((int32_t)(x + idx[j]) - half);
                    if (xs < n) {
                        int lr = (int)xs - rbase;
                        if (lr < lr0 || lr >= lr1) {
                            atomicOr(d.err, 1u);
                            lr = lr < lr0 ? lr0 : lr1 - 1;
                        }
                        v = v + cur[(size_t)lr * n + y] * w[j];
                    }
                }
                oth[(size_t)orow * n + y] = v;
            }
        }
    } else {
        // ---- E: rotation (RotationMap::apply).  Taps out of range read
        // sample 0 of the (wake-kicked) input, as the reference does.
        float s00;
        if (rbase <= 0) {
            s00 = cur[(size_t)(-rbase) * n];
        } else if (wake) {
            if (DIRECT) {
                s00 = s00_sh;
            } else {
                uint32_t idx[IT];
                float w[IT];
                kick_taps<IT>(wake_off[0], n, idx, w);
                float v = 0.0f;
#pragma unroll
                for (int j = 0; j < IT; j++) {
                    const uint32_t ys = (uint32_t)((int32_t)idx[j] - half);
                    if (ys < n) {
                        v = v + in[ys] * w[j];
                    }
                }
                s00 = v;
            }
        } else {
            s00 = in[0] * norm;
        }
        VFPS_STAMP(4);
        const RotGeom g = d.rot;
        const float* src = cur;
        const uint32_t drow = bs / n, dy = bs % n;
        uint32_t orow = tid / n, y0 = tid % n;
        while (orow < (uint32_t)rows_out) {
            const float v = rotate_cell_src<IT, CLAMP>(
                n, g, (uint32_t)x0 + orow, y0,
                [&](bool ok, uint32_t i0, uint32_t j0) -> float {
                    if (!ok) {
                        return s00;
                    }
                    int lr = (int)i0 - rbase;
                    if (lr < lr0 || lr >= lr1) {
                        atomicOr(d.err, 2u);
                        lr = lr < lr0 ? lr0 : lr1 - 1;
                    }
                    return src[(size_t)lr * n + j0];
                },
                [&](uint32_t i0) -> const float* {
                    // the IT rows from i0 on, if the buffer holds them all
                    const int lr = (int)i0 - rbase;
                    return lr >= lr0 && lr + IT <= lr1 ? src + (size_t)lr * n
                                                       : nullptr;
                },
                n);
            oth[(size_t)orow * n + y0] = v;
            y0 += dy;
            orow += drow;
            if (y0 >= n) {
                y0 -= n;
                orow++;
            }
        }
    }
    __syncthreads();
    VFPS_STAMP(5);
    // ---- F: Fokker-Planck along y, store, projection products ----------------
    // (one mesh: the output rows go to LDS and out to global memory during
    // the projection, by the threads that do not sum it: a global store
    // here would make the summing lanes wait for its completion first)
    tile_out = L.tile_out == STEP_TILE_IN_BUF ? cur : lds + L.tile_out;
    float* fout = lds + L.fout;
    const uint32_t tstride = n + 4;
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
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
            if constexpr (ONE_MESH) {
                fout[(size_t)orow * n + y] = v;
            } else {
                out[(size_t)(x0 + orow) * n + y] = v;
            }
            tile_out[(size_t)orow * tstride + y] = v * col_ws[c];
        }
    }
    if (tid == 0) {
        // history slot of this step: the steps completed before it
        uint32_t slot = 0xFFFFFFFFu;
        if (d.prof_hist && d.counter && cnt < d.hist_cap) {
            slot = cnt;
        }
        slot_sh = slot;
    }
    __syncthreads();
    const uint32_t slot = slot_sh;
    VFPS_STAMP(6);
    // ---- G: x projection ------------------------------------------------------
    // Stores deferred to here, by threads i0, i0 + nt, ...: the wake offsets
    // of the output rows (direct wake; issued in phase B they held up the
    // next waits on the load counter, which drains in order) and, one mesh,
    // the output rows.
    auto deferred_stores = [&](uint32_t i0, uint32_t nt) {
        if (DIRECT && wake && d.wake_out) {
            for (uint32_t r = i0; r < (uint32_t)rows_out; r += nt) {
                d.wake_out[x0 + (int)r] = offw[hlo + (int)r];
            }
        }
        if constexpr (ONE_MESH) {
            const uint32_t tot = (uint32_t)rows_out * n;
            gfloat* dst = out + (size_t)x0 * n;
            if (vec_rows) {
                const vec4f* s4 = reinterpret_cast<const vec4f*>(fout);
                gfloat4* d4 = (gfloat4*)dst;
                for (uint32_t i = i0; i < (tot >> 2); i += nt) {
                    d4[i] = s4[i];
                }
            } else {
                for (uint32_t i = i0; i < tot; i += nt) {
                    dst[i] = fout[i];
                }
            }
        }
    };
    if (d.seq_proj) {
        // wave 0's first lanes sum; the other waves store (one wave: after)
        if (bs > 64u && tid >= 64u) {
            deferred_stores(tid - 64u, bs - 64u);
        }
        // the reference's std::inner_product order: y = 0..n-1, one thread
        // per row, loads batched ahead of the dependent adds
        if ((int)tid < rows_out) {
            const float* r = tile_out + (size_t)tid * tstride;
            float acc = 0.0f;
            uint32_t y = 0;
            if ((n & 63u) == 0u) {
                // two register buffers of 32 values: one loads while the
                // other is added, so 32 dependent adds cover each LDS trip
                const vec4f* r4 = reinterpret_cast<const vec4f*>(r);
                vec4f qa[8], qb[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    qa[k] = r4[k];
                }
                // (one mesh: scheduling barriers keep each group of loads
                // ahead of the adds it overlaps; merged, every 64 adds wait
                // for a full LDS round trip.  They cost registers, which a
                // batch keeps for its occupancy.)
                for (; y < n; y += 64) {
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        qb[k] = r4[((y + 32) >> 2) + k];
                    }
                    if constexpr (ONE_MESH) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        acc = acc + qa[k].x;
                        acc = acc + qa[k].y;
                        acc = acc + qa[k].z;
                        acc = acc + qa[k].w;
                    }
                    // the last pass reloads the row's first values (unused)
                    const uint32_t yn = y + 64 < n ? y + 64 : 0u;
                    if constexpr (ONE_MESH) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        qa[k] = r4[(yn >> 2) + k];
                    }
                    if constexpr (ONE_MESH) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        acc = acc + qb[k].x;
                        acc = acc + qb[k].y;
                        acc = acc + qb[k].z;
                        acc = acc + qb[k].w;
                    }
                }
            } else if ((n & 15u) == 0u) {
                // the next 16 values load while the current 16 are added
                const vec4f* r4 = reinterpret_cast<const vec4f*>(r);
                vec4f q[4], nq[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    q[k] = r4[k];
                }
                for (; y < n; y += 16) {
                    const uint32_t yn = y + 16 < n ? y + 16 : y;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        nq[k] = r4[(yn >> 2) + k];
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        acc = acc + q[k].x;
                        acc = acc + q[k].y;
                        acc = acc + q[k].z;
                        acc = acc + q[k].w;
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        q[k] = nq[k];
                    }
                }
            }
            for (; y < n; y++) {
                acc = acc + r[y];
            }
            const uint32_t x = (uint32_t)x0 + tid;
            proj_out[x] = acc;
            if (slot != 0xFFFFFFFFu) {
                prof_hist[(size_t)slot * n + x] = acc;
            }
        }
        if (bs == 64u) {
            deferred_stores(tid, 64u);
        }
    } else {
        deferred_stores(tid, bs);
        for (int orow = (int)wave; orow < rows_out; orow += (int)nw) {
            const float* r = tile_out + (size_t)orow * tstride;
            float acc = 0.0f;
            for (uint32_t y = lane; y < n; y += 64) {
                acc = acc + r[y];
            }
            acc = wave_sum(acc);
            if (lane == 0) {
                const uint32_t x = (uint32_t)(x0 + orow);
                proj_out[x] = acc;
                if (slot != 0xFFFFFFFFu) {
                    prof_hist[(size_t)slot * n + x] = acc;
                }
            }
        }
    }
    VFPS_STAMP(7);
    // ---- H: close the step (direct form) --------------------------------------
    // Every tile leaves its CSR share.  In a batch, tile 0 of each instance
    // also closes the step (step_close); a one-mesh launch gives that to a
    // workgroup of its own.
    if (DIRECT) {
        if (csr && tid == 0) {
            d.csr_part[(size_t)parity * n + tile] = csr_q;
        }
        if (!ONE_MESH && tile == 0) {
            step_close(d, parity, cnt, red);
        }
    }
    if (d.tstamp) {
        // per-tile entry and exit times
        __syncthreads();
        if (threadIdx.x == 0) {
            d.tstamp[16 + 2 * logical] = t_entry;
            d.tstamp[17 + 2 * logical] = wall_clock64();
        }
    }
}

#undef VFPS_STAMP

// Launch forms of the step kernel
enum StepForm : int {
    STEP_SINGLE = 0,      // one mesh, descriptor by value, one workgroup per tile
    STEP_PER_TILE = 1,    // a batch, one workgroup per block-map entry
    STEP_PERSISTENT = 2,  // a batch, a resident grid walking the block map
};

// The step of one launch.
//  STEP_SINGLE: one mesh, its descriptor passed by value (the launch's own
//  arguments, so no load of a descriptor or block map precedes the data
//  loads), one workgroup per tile; the direct form adds a last workgroup
//  that closes the step (step_close).
//  Batches: blockmap[nmap] holds (instance << 16 | tile) per logical tile
//  (STEP_NO_TILE: padding); nmap and the grid are multiples of 8.
// XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so the
// map is cut into 8 contiguous runs, one per XCD (neighbouring tiles share
// halo rows, which then hit in that XCD's L2).  STEP_PER_TILE: workgroup w
// takes entry (w % 8) * (nmap / 8) + w / 8.  STEP_PERSISTENT: the
// workgroups of an XCD walk its run with a stride of grid / 8 (the loop is
// a form of its own: around the per-tile body it costs registers).
template <int IT, int IPFP, bool DIRECT, int SCHEME, bool CLAMP, int FORM,
          int MAXC>
__global__ void __launch_bounds__(STEP_MAX_THREADS)
step_kernel(const InstDesc* __restrict__ descs,
            const uint32_t* __restrict__ blockmap, uint32_t nmap,
            uint32_t parity, InstDesc d0)
{
    extern __shared__ __align__(16) float lds[];
    __shared__ StepShared sh;
    const uint32_t G = gridDim.x;
    if constexpr (FORM == STEP_SINGLE) {
        // direct form: the last workgroup closes the step, beside the tiles
        // (in a tile it would lengthen the launch's slowest workgroup)
        const uint32_t tiles = DIRECT ? G - 1u : G;
        if (DIRECT && blockIdx.x == tiles) {
            const uint32_t cnt = d0.counter ? ((const guint*)d0.counter)[parity] : 0u;
            step_close(d0, parity, cnt, sh.red);
            return;
        }
        uint32_t logical = blockIdx.x;
        if ((tiles & 7u) == 0u) {
            logical = (blockIdx.x & 7u) * (tiles >> 3) + (blockIdx.x >> 3);
        }
        step_tile<IT, IPFP, DIRECT, SCHEME, CLAMP, MAXC, true>(
            d0, 0u, logical, logical, parity, lds, sh);
    } else if constexpr (FORM == STEP_PER_TILE) {
        const uint32_t logical = (blockIdx.x & 7u) * (nmap >> 3) + (blockIdx.x >> 3);
        const uint32_t m = blockmap[logical];
        if (m == STEP_NO_TILE) {
            return;  // padding of the map to a multiple of the XCD count
        }
        step_tile<IT, IPFP, DIRECT, SCHEME, CLAMP, MAXC, false>(
            descs[m >> 16], m >> 16, m & 0xFFFFu, logical, parity, lds, sh);
    } else {
        const uint32_t per = nmap >> 3;
        const uint32_t wgs = G >> 3;
        const uint32_t xcd = blockIdx.x & 7u;
        for (uint32_t i = blockIdx.x >> 3; i < per; i += wgs) {
            const uint32_t logical = xcd * per + i;
            const uint32_t m = blockmap[logical];
            if (m == STEP_NO_TILE) {
                continue;
            }
            step_tile<IT, IPFP, DIRECT, SCHEME, CLAMP, MAXC, false>(
                descs[m >> 16], m >> 16, m & 0xFFFFu, logical, parity, lds, sh);
            // the next tile reuses the LDS
            __syncthreads();
        }
    }
}

}  // namespace vfps_impl
