// Strip form of the batched kick-drift step (scheme 0, direct wake).
//
// A workgroup advances a strip of S = K * T consecutive output rows of one
// instance, in K sub-tiles of T rows, instead of one tile of T rows:
//   A  once per strip: the previous x projection, the per-column drift,
//      Fokker-Planck and Simpson tables, the RF taps of every row the strip
//      kicks and (when they fit the staging rows) the rows the first drift
//      window kicks, all loads issued before any store;
//   B  once per strip: the direct wake offsets of the kicked rows and the
//      CSR dot products of the output rows, STRIP_DOT_ROWS rows per
//      wavefront sharing each load of the profile (each row's sum keeps the
//      lane order and the wave_sum_dpp of the tile kernel, so the bits are
//      the same), then the tiles' CSR shares;
//   then per sub-tile, two barriers apart:
//   C, D  the wake and RF kicks of the rows that enter the drift's window,
//      from the input rows staged in LDS into a ring of kicked rows (a row
//      is kicked once per strip, not once per tile that reads it), beside
//      the previous sub-tile's F;
//   E  the drift along x from the ring, a sliding window of consecutive
//      rows per column, beside the loads of the next entering rows;
//   F  Fokker-Planck along y and the store of the output rows;
//   and at the end:
//   G  the x projection of the strip's rows, one row per lane of one
//      wavefront, summed in std::inner_product order from the rows just
//      stored;
//   H  the CSR shares were stored in B; strip 0 closes the step.
// Every output value is the same sum, in the same order, as in the tile
// kernel, so the two forms give the same bits (pinned by the batch tests).
#pragma once

#include "step_kernel.hpp"

namespace vfps_impl {

// most output rows of one strip: its x projection is summed by one
// wavefront, one row per lane
constexpr uint32_t STRIP_MAX_ROWS = 64;
// most rows of one sub-tile
constexpr uint32_t STRIP_MAX_T = 8;
// most rows of the direct-form dot products one wavefront sums together
constexpr int STRIP_DOT_ROWS = 10;
// float4 loads of the input rows one thread keeps in flight
constexpr int STRIP_LOAD_BATCH = 2;
// float4 loads of stored rows one thread keeps in flight in the projection
constexpr int STRIP_PROJ_BATCH = 4;
// loads per thread of each strip table (profile, weights, kernel windows)
// issued together before any is stored
constexpr int STRIP_PRE = 3;
// float4 loads per thread of the first drift window's rows issued with the
// tables (when the rows fit the two staging buffers)
constexpr int STRIP_EARLY4 = 4;

// rows of the ring of kicked rows: a power of two that holds the drift's
// window of T + halo rows
__host__ __device__ inline uint32_t strip_ring_rows(uint32_t T, uint32_t halo)
{
    uint32_t r = 1;
    while (r < T + halo) r <<= 1;
    return r;
}

// LDS layout of the strip kernel (4-byte words; FP64 arrays first).  The
// FP64 kernel windows serve the dot products before the first kick only,
// so they share their words with the ring and the staging rows.
struct StripLds {
    uint32_t wsum, csum;   // FP64 dot products of the wake rows, CSR rows
    uint32_t pprev;        // previous x projection
    uint32_t wsl;          // Simpson weights
    uint32_t tapw, tapr;   // kick taps of every kicked row of the strip
    uint32_t ring;         // ring of kicked rows, row r at (r + hlo) % rows
    uint32_t stage_a;      // [T][n] input rows a kick reads
    uint32_t stage_b;      // [T][n] the drift's output rows
    uint32_t gwin, hwin;   // FP64 windows of the wake and CSR kernels
    uint32_t shared;       // words of ring + staging rows (or the windows)
    uint32_t words;
};

__host__ __device__ inline StripLds strip_lds(uint32_t n, uint32_t T, uint32_t K,
                                              uint32_t halo)
{
    StripLds L{};
    const uint32_t S = T * K;
    const uint32_t nk = S + halo;  // rows one strip kicks, at most
    const uint32_t n4 = step_round4(n);
    uint32_t w = 0;
    L.wsum = w;
    w += 2 * nk;
    L.csum = w;
    w += 2 * S;
    w = step_round4(w);
    L.pprev = w;
    w += n4;
    L.wsl = w;
    w += n4;
    L.tapw = w;
    w += STEP_TAP_WORDS * nk;
    L.tapr = w;
    w += STEP_TAP_WORDS * nk;
    L.ring = w;
    L.stage_a = L.ring + strip_ring_rows(T, halo) * n4;
    L.stage_b = L.stage_a + T * n4;
    const uint32_t rows_words = L.stage_b + T * n4 - L.ring;
    // (each window is followed by STRIP_DOT_ROWS doubles that the dot
    // products' unused row slots may read)
    L.gwin = w;
    L.hwin = L.gwin + 2 * (nk + n + STRIP_DOT_ROWS);
    const uint32_t win_words = 2 * (nk + n + STRIP_DOT_ROWS) + 2 * (S + n + STRIP_DOT_ROWS);
    L.shared = rows_words > win_words ? rows_words : win_words;
    w += L.shared;
    // the kick's 16-byte window reads run up to 6 words past a staged row
    w += 8;
    L.words = w;
    return L;
}

// The wake kick and then the RF kick along y (or the RF kick alone) of the
// rows [a, a + c) staged in LDS, into their ring slots.  Each thread makes
// GW consecutive outputs of a row (GW = 4 or 8; n % GW == 0) from the
// GW + IT - 1 wake-kicked samples they read, each the same sum in the same
// order as kick2_rows / kick_rows make it; a wavefront whose groups all
// read inside their rows skips the per-sample checks, and reads its
// windows as 16-byte words when their starts agree modulo 4 (lds_window16).
template <int IT, int GW>
__device__ __forceinline__ void strip_kick(const float* __restrict__ stage,
                                           float* __restrict__ ring, uint32_t n,
                                           int a, int c, int k0, uint32_t hlo,
                                           uint32_t rmask,
                                           const float* __restrict__ tapw,
                                           const float* __restrict__ tapr,
                                           bool wake)
{
    constexpr int NC = GW + IT - 1;      // wake-kicked samples a group reads
    constexpr int NV = GW + 2 * IT - 2;  // input samples behind them
    const uint32_t bs = blockDim.x;
    const uint32_t ng = n / GW;
    const uint32_t drow = bs / ng, dg = bs % ng;
    uint32_t row = threadIdx.x / ng;
    uint32_t g = threadIdx.x % ng;
    const int in = (int)n;
    while (row < (uint32_t)c) {
        const int r = a + (int)row;
        const float* t2 = tapr + (uint32_t)(r - k0) * STEP_TAP_WORDS;
        const vec4f w2 = *reinterpret_cast<const vec4f*>(t2);
        const int32_t base2 = __float_as_int(t2[4]);
        const float* s = stage + (size_t)row * n;
        const int32_t y0 = (int32_t)(GW * g);
        float o[GW];
        if (wake) {
            const float* t1 = tapw + (uint32_t)(r - k0) * STEP_TAP_WORDS;
            const vec4f w1 = *reinterpret_cast<const vec4f*>(t1);
            const int32_t base1 = __float_as_int(t1[4]);
            const int32_t clo = y0 + base2;
            const int32_t lo = clo + base1;
            const bool inside = lo >= 0 && lo + NV - 1 < in && clo >= 0 &&
                                clo + NC - 1 < in;
            float v[NV];
            float cc[NC];
            // the wake-kicked samples of a group that lies inside its row
            auto kick1 = [&](const float* vv) {
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    float acc = 0.0f;
#pragma unroll
                    for (int j = 0; j < IT; j++) {
                        acc = acc + vv[k + j] * w1[j];
                    }
                    cc[k] = acc;
                }
            };
            if (__all(inside)) {
                // every group of the wavefront lies inside its row
                const uint32_t sft = (uint32_t)lo & 3u;
                const uint32_t sft0 = __builtin_amdgcn_readfirstlane(sft);
                if (__all(sft == sft0)) {
                    lds_window16<NV>(s, lo, sft0, kick1);
                } else {
#pragma unroll
                    for (int k = 0; k < NV; k++) {
                        v[k] = s[lo + k];
                    }
                    kick1(v);
                }
            } else {
#pragma unroll
                for (int k = 0; k < NV; k++) {
                    const uint32_t ys = (uint32_t)(lo + k);
                    v[k] = ys < n ? s[ys] : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    float acc = 0.0f;
#pragma unroll
                    for (int j = 0; j < IT; j++) {
                        acc = acc + v[k + j] * w1[j];
                    }
                    cc[k] = (uint32_t)(clo + k) < n ? acc : 0.0f;
                }
            }
#pragma unroll
            for (int e = 0; e < GW; e++) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < IT; j++) {
                    acc = acc + cc[e + j] * w2[j];
                }
                o[e] = acc;
            }
        } else {
            const int32_t lo = y0 + base2;
            const bool inside = lo >= 0 && lo + NC - 1 < in;
            float v[NC];
            auto kick2 = [&](const float* vv) {
#pragma unroll
                for (int e = 0; e < GW; e++) {
                    float acc = 0.0f;
#pragma unroll
                    for (int j = 0; j < IT; j++) {
                        acc = acc + vv[e + j] * w2[j];
                    }
                    o[e] = acc;
                }
            };
            const uint32_t sft = (uint32_t)lo & 3u;
            const uint32_t sft0 = __builtin_amdgcn_readfirstlane(sft);
            if (__all(inside && sft == sft0)) {
                lds_window16<NC>(s, lo, sft0, kick2);
            } else if (__all(inside)) {
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    v[k] = s[lo + k];
                }
                kick2(v);
            } else {
#pragma unroll
                for (int k = 0; k < NC; k++) {
                    const uint32_t ys = (uint32_t)(lo + k);
                    v[k] = ys < n ? s[ys] : 0.0f;
                }
                kick2(v);
            }
        }
        const uint32_t slot = ((uint32_t)r + hlo) & rmask;
        float* dst = ring + __umul24(slot, n) + GW * g;
#pragma unroll
        for (int q = 0; q < GW / 4; q++) {
            vec4f ov;
            ov.x = o[4 * q];
            ov.y = o[4 * q + 1];
            ov.z = o[4 * q + 2];
            ov.w = o[4 * q + 3];
            *reinterpret_cast<vec4f*>(dst + 4 * q) = ov;
        }
        g += dg;
        row += drow;
        if (g >= ng) {
            g -= ng;
            row++;
        }
    }
}

template <int IT, int IPFP, int MAXC>
__device__ __forceinline__ void strip_body(const InstDesc& d, uint32_t strip,
                                           uint32_t parity, float* lds,
                                           StepShared& sh)
{
    const uint32_t n = d.n;
    const uint32_t T = d.T;
    const uint32_t K = d.K;
    const uint32_t S = T * K;
    const int in = (int)n;
    const int s0 = (int)(strip * S);
    if (s0 >= in) {
        return;  // not launched: maps hold exactly the strips
    }
    const int srows = min((int)S, in - s0);
    const int hlo = (int)d.hlo, hhi = (int)d.hhi;
    const uint32_t rmask = d.ringw - 1u;
    // the rows the strip kicks: every row a drift window of it reads
    const int k0 = max(0, s0 - hlo);
    const int k1 = min(in, s0 + srows + hhi);
    const int nk = k1 - k0;
    const StripLds L = strip_lds(n, T, K, d.hlo + d.hhi);
    double* gwin = reinterpret_cast<double*>(lds + L.gwin);
    double* hwin = reinterpret_cast<double*>(lds + L.hwin);
    double* wsum = reinterpret_cast<double*>(lds + L.wsum);
    double* csum = reinterpret_cast<double*>(lds + L.csum);
    float* pprev = lds + L.pprev;
    float* wsl = lds + L.wsl;
    float* tapw = lds + L.tapw;
    float* tapr = lds + L.tapr;
    float* ring = lds + L.ring;
    float* stage_a = lds + L.stage_a;
    float* stage_b = lds + L.stage_b;
    const gfloat* in_mesh = (const gfloat*)d.mesh[parity];
    gfloat* out = (gfloat*)d.mesh[parity ^ 1u];
    const gfloat* proj_in = (const gfloat*)d.proj[parity];
    gfloat* proj_out = (gfloat*)d.proj[parity ^ 1u];
    gfloat* prof_hist = (gfloat*)d.prof_hist;
    const gfloat* ws_g = (const gfloat*)d.ws;
    const gfloat* rf_off = (const gfloat*)d.rf_off;
    const gfloat* drift_off = (const gfloat*)d.drift_off;
    const gfloat* fp_w = (const gfloat*)d.fp_w;
    const guint* fp_i = (const guint*)d.fp_i;
    const gdouble* wake_g = (const gdouble*)d.wake_g;
    const gdouble* csr_h = (const gdouble*)d.csr_h;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6, nw = bs >> 6;
    const int32_t half = (int32_t)(n / 2);
    const bool wake = d.use_wake != 0;
    const bool csr = d.use_csr != 0;

    // ---- A: the strip's tables ---------------------------------------------
    // The first STRIP_PRE loads per thread of every table, the per-column
    // tables and the rows the first drift window kicks are all issued before
    // any is stored, so that they share one memory round trip.
    uint32_t cnt_v = 0;
    if (tid == 0 && d.counter) {
        cnt_v = d.counter[parity];
    }
    const uint32_t n4 = n >> 2;
    // the first window kicks rows [k0, e0); they go to stage_a and on into
    // stage_b, which the kernel windows below do not overlap
    const int ro0 = min((int)T, in - s0);
    const int e0 = min(k1, s0 + ro0 + hhi);
    const uint32_t etot4 = (uint32_t)(e0 - k0) * n4;
    const bool early = e0 - k0 <= 2 * (int)T &&
                       etot4 <= (uint32_t)STRIP_EARLY4 * bs &&
                       L.hwin + 2u * (S + n + STRIP_DOT_ROWS) <= L.stage_a;
    vec4f er[STRIP_EARLY4];
    if (early) {
        const gfloat4* src4 = (const gfloat4*)(in_mesh + (size_t)k0 * n);
#pragma unroll
        for (int b = 0; b < STRIP_EARLY4; b++) {
            const uint32_t i = tid + (uint32_t)b * bs;
            er[b] = src4[i < etot4 ? i : etot4 - 1u];
        }
    }
    // the kernel windows the strip's dot products read: row r of the wake
    // needs g[r + n - 1 - x'] (r in [k0, k1)), CSR row r h[r + n - 1 - x']
    // (r in [s0, s0 + srows))
    const uint32_t gl = wake ? (uint32_t)nk + n - 1u : 0u;
    const uint32_t hl = csr ? (uint32_t)srows + n - 1u : 0u;
    float pv[STRIP_PRE], wv[STRIP_PRE];
    double gv[STRIP_PRE], hv[STRIP_PRE];
#pragma unroll
    for (int b = 0; b < STRIP_PRE; b++) {
        const uint32_t i = tid + (uint32_t)b * bs;
        pv[b] = i < n ? proj_in[i] : 0.0f;
        wv[b] = i < n ? ws_g[i] : 0.0f;
        gv[b] = i < gl ? wake_g[(uint32_t)k0 + i] : 0.0;
        hv[b] = i < hl ? csr_h[(uint32_t)s0 + i] : 0.0;
    }
    // per-column tables: the drift taps' first row and weights (the taps
    // are consecutive unless one falls outside the mesh), the
    // Fokker-Planck stencil's first index and weights (consecutive away from
    // the mesh edges and the row-floor(ycenter) quirk); columns whose taps
    // are not consecutive reread their tables in the sub-tiles
    const int ncols = (int)((n + bs - 1) / bs);
    uint32_t col_d0[MAXC];
    float col_dw[MAXC][IT];
    uint32_t col_f0[MAXC];
    float col_fw[MAXC][IPFP > 0 ? IPFP : 1];
    uint32_t dcons = 0, fcons = 0;  // bit c: column c's taps are consecutive
    // their loads (and the first RF offset per thread) go out before any
    // is used; a column past the mesh reads column 0, unused
    float col_doff[MAXC];
    uint32_t col_fi[MAXC][IPFP > 0 ? IPFP : 1];
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
        const uint32_t y = tid + (uint32_t)c * bs;
        const uint32_t yy = (c < ncols && y < n) ? y : 0u;
        col_doff[c] = drift_off[yy];
#pragma unroll
        for (int k = 0; k < IPFP; k++) {
            col_fi[c][k] = fp_i[(size_t)k * n + yy];
            col_fw[c][k] = fp_w[(size_t)k * n + yy];
        }
    }
    const float rf0 = rf_off[k0 + min((int)tid, nk - 1)];
#pragma unroll
    for (int c = 0; c < MAXC; c++) {
        const uint32_t y = tid + (uint32_t)c * bs;
        col_d0[c] = 0;
        col_f0[c] = 0;
        if (c < ncols && y < n) {
            uint32_t idx[IT];
            kick_taps<IT>(col_doff[c], n, idx, col_dw[c]);
            bool cons = true;
#pragma unroll
            for (int j = 1; j < IT; j++) {
                cons = cons && idx[j] == idx[0] + (uint32_t)j;
            }
            dcons |= cons ? (1u << c) : 0u;
            col_d0[c] = idx[0];
            bool fc = true;
#pragma unroll
            for (int k = 0; k < IPFP; k++) {
                const uint32_t fi = col_fi[c][k];
                if (k == 0) col_f0[c] = fi;
                fc = fc && fi == col_f0[c] + (uint32_t)k;
            }
            fcons |= fc ? (1u << c) : 0u;
        }
    }
    if ((int)tid < nk) {
        store_row_taps<IT>(rf0, n, tapr + tid * STEP_TAP_WORDS);
    }
    for (int i = (int)(tid + bs); i < nk; i += (int)bs) {
        store_row_taps<IT>(rf_off[k0 + i], n, tapr + (uint32_t)i * STEP_TAP_WORDS);
    }
#pragma unroll
    for (int b = 0; b < STRIP_PRE; b++) {
        const uint32_t i = tid + (uint32_t)b * bs;
        if (i < n) {
            pprev[i] = pv[b];
            wsl[i] = wv[b];
        }
        if (i < gl) gwin[i] = gv[b];
        if (i < hl) hwin[i] = hv[b];
    }
    for (uint32_t i = tid + STRIP_PRE * bs; i < n; i += bs) {
        pprev[i] = proj_in[i];
        wsl[i] = ws_g[i];
    }
    for (uint32_t i = tid + STRIP_PRE * bs; i < gl; i += bs) gwin[i] = wake_g[(uint32_t)k0 + i];
    for (uint32_t i = tid + STRIP_PRE * bs; i < hl; i += bs) hwin[i] = csr_h[(uint32_t)s0 + i];
    if (early) {
        vec4f* st4 = reinterpret_cast<vec4f*>(stage_a);
#pragma unroll
        for (int b = 0; b < STRIP_EARLY4; b++) {
            const uint32_t i = tid + (uint32_t)b * bs;
            if (i < etot4) st4[i] = er[b];
        }
    }
    if (tid == 0) {
        sh.cnt = cnt_v;
    }
    __syncthreads();
    const uint32_t cnt = sh.cnt;
    // PhaseSpace::integrateAndNormalize on the steps the reference
    // renormalises, as in the tile kernel (the same FP64 sum, so every
    // strip and every tile finds the same factor)
    const bool renorm = d.norm_every != 0 && cnt % d.norm_every == 0;
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
        if (early) {
            // the first window's rows, staged before the factor was known;
            // each thread scales the words it stored
            vec4f* st4 = reinterpret_cast<vec4f*>(stage_a);
            for (uint32_t i = tid; i < etot4; i += bs) {
                vec4f w = st4[i];
                w.x = w.x * norm;
                w.y = w.y * norm;
                w.z = w.z * norm;
                w.w = w.w * norm;
                st4[i] = w;
            }
        }
    }

    // ---- B: wake offsets of the kicked rows, CSR rows of the output rows ----
    // A job is up to STRIP_DOT_ROWS consecutive rows of one table, sized so
    // that the jobs spread over the wavefronts: lane l takes x' = l, l + 64,
    // ... and row r accumulates g[r + n - 1 - x'] p[x'] (the tile kernel's
    // window order), then wave_sum_dpp.
    {
        const int tot_rows = (wake ? nk : 0) + (csr ? srows : 0);
        const int rj = min(STRIP_DOT_ROWS, max(1, (tot_rows + (int)nw - 1) / (int)nw));
        const int nwj = wake ? (nk + rj - 1) / rj : 0;
        const int ncj = csr ? (srows + rj - 1) / rj : 0;
        for (int job = (int)wave; job < nwj + ncj; job += (int)nw) {
            const bool isw = job < nwj;
            const int r_first = isw ? k0 + rj * job : s0 + rj * (job - nwj);
            const int rows = isw ? min(rj, k1 - r_first) : min(rj, s0 + srows - r_first);
            const double* tbl =
                isw ? gwin + (r_first - k0 + in - 1) : hwin + (r_first - s0 + in - 1);
            double acc[STRIP_DOT_ROWS];
#pragma unroll
            for (int i = 0; i < STRIP_DOT_ROWS; i++) {
                acc[i] = 0.0;
            }
            // every row slot reads its window value (past a job's last row
            // they read words the windows' padding keeps inside LDS, and
            // their sums are dropped), so that the loads of one x' go out
            // together instead of one branch and one wait per row
            for (uint32_t x = lane; x < n; x += 64) {
                const double p = (double)pprev[x];
                const double* t = tbl - x;
                double tv[STRIP_DOT_ROWS];
#pragma unroll
                for (int i = 0; i < STRIP_DOT_ROWS; i++) {
                    tv[i] = t[i];
                }
#pragma unroll
                for (int i = 0; i < STRIP_DOT_ROWS; i++) {
                    acc[i] += tv[i] * p;
                }
            }
            double* dst = isw ? wsum + (r_first - k0) : csum + (r_first - s0);
#pragma unroll
            for (int i = 0; i < STRIP_DOT_ROWS; i++) {
                if (i < rows) {
                    const double v = wave_sum_dpp(acc[i]);
                    if (lane == 0) dst[i] = v;
                }
            }
        }
    }
    __syncthreads();
    // wake taps of the kicked rows, the wake offsets of the output rows
    if (wake) {
        for (int i = (int)tid; i < nk; i += (int)bs) {
            const float off = d.wake_scale * (float)wsum[i];
            store_row_taps<IT>(off, n, tapw + (uint32_t)i * STEP_TAP_WORDS);
            const int r = k0 + i;
            if (d.wake_out && r >= s0 && r < s0 + srows) {
                d.wake_out[r] = off;
            }
        }
    }
    // each tile's CSR share: sum over its rows of p[x] * (H p)[x], in order
    if (csr) {
        for (uint32_t k = tid; k < K; k += bs) {
            const int x0 = s0 + (int)(k * T);
            if (x0 >= in) continue;
            const int rows_out = min((int)T, in - x0);
            double q = 0.0;
            for (int o = 0; o < rows_out; o++) {
                q += (double)pprev[x0 + o] * csum[x0 + o - s0];
            }
            d.csr_part[(size_t)parity * n + strip * K + k] = q;
        }
    }
    __syncthreads();

    // ---- sub-tiles ----------------------------------------------------------
    // Two barriers per sub-tile: the kick of the rows entering the drift's
    // window (from stage_a into the ring) runs beside the previous
    // sub-tile's Fokker-Planck (stage_b to global memory); after a barrier
    // the drift (ring to stage_b) runs beside the loads of the next
    // sub-tile's entering rows into stage_a; then a barrier.  A slot the
    // kick writes held a row below the drift window it follows, so the
    // ring needs no more rows than one window.
    // rows [a, a + c) of the input (renormalised when due) into dst
    auto load_rows = [&](int a, int c, float* dst) {
        const uint32_t tot4 = (uint32_t)c * n4;
        const gfloat4* src4 = (const gfloat4*)(in_mesh + (size_t)a * n);
        vec4f* st4 = reinterpret_cast<vec4f*>(dst);
        for (uint32_t base = 0; base < tot4; base += bs * STRIP_LOAD_BATCH) {
            vec4f v[STRIP_LOAD_BATCH];
#pragma unroll
            for (int b = 0; b < STRIP_LOAD_BATCH; b++) {
                const uint32_t i = base + tid + (uint32_t)b * bs;
                v[b] = src4[i < tot4 ? i : tot4 - 1u];
            }
#pragma unroll
            for (int b = 0; b < STRIP_LOAD_BATCH; b++) {
                const uint32_t i = base + tid + (uint32_t)b * bs;
                if (i < tot4) {
                    vec4f w = v[b];
                    if (renorm) {
                        w.x = w.x * norm;
                        w.y = w.y * norm;
                        w.z = w.z * norm;
                        w.w = w.w * norm;
                    }
                    st4[i] = w;
                }
            }
        }
    };
    // four outputs per thread: the 16-byte window reads of neighbouring
    // lanes then fall on different banks
    auto kick_rows = [&](int a, int c) {
        strip_kick<IT, 4>(stage_a, ring, n, a, c, k0, (uint32_t)hlo, rmask,
                          tapw, tapr, wake);
    };
    // ring rows of [lo, hi) outside the mesh read as zeros
    auto zero_rows = [&](int lo, int hi) {
        for (int r = lo; r < hi; r++) {
            if (r >= 0 && r < in) continue;
            float* dst = ring + (size_t)(((uint32_t)r + (uint32_t)hlo) & rmask) * n;
            for (uint32_t y = tid; y < n; y += bs) dst[y] = 0.0f;
        }
    };
    // F: Fokker-Planck along y of sub-tile x0's rows, to global memory
    auto fp_rows = [&](int x0, int rows_out) {
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t y = tid + (uint32_t)c * bs;
            if (c >= ncols || y >= n) {
                continue;
            }
            const bool fc = (fcons >> c) & 1u;
            const uint32_t f0 = col_f0[c];
            if (IPFP > 0 && rows_out == 4 && __all(fc)) {
                // a full tile of four rows: every stencil value loads
                // before the sums (the same sums in the same order)
                float sv[4][IPFP > 0 ? IPFP : 1];
#pragma unroll
                for (int o = 0; o < 4; o++) {
#pragma unroll
                    for (int kk = 0; kk < IPFP; kk++) {
                        sv[o][kk] = stage_b[(size_t)o * n + f0 + (uint32_t)kk];
                    }
                }
#pragma unroll
                for (int o = 0; o < 4; o++) {
                    float v = 0.0f;
#pragma unroll
                    for (int kk = 0; kk < IPFP; kk++) {
                        v = v + sv[o][kk] * col_fw[c][kk];
                    }
                    out[(size_t)(x0 + o) * n + y] = v;
                }
                continue;
            }
            if (IPFP == 0 || __all(fc)) {
                // every column of the wavefront has a consecutive stencil
                for (int o = 0; o < rows_out; o++) {
                    const float* src = stage_b + (size_t)o * n;
                    float v;
                    if constexpr (IPFP == 0) {
                        v = src[y];
                    } else {
                        v = 0.0f;
#pragma unroll
                        for (int kk = 0; kk < IPFP; kk++) {
                            v = v + src[f0 + (uint32_t)kk] * col_fw[c][kk];
                        }
                    }
                    out[(size_t)(x0 + o) * n + y] = v;
                }
                continue;
            }
            uint32_t fi[IPFP > 0 ? IPFP : 1];
            if (IPFP > 0 && !fc) {
#pragma unroll
                for (int kk = 0; kk < IPFP; kk++) {
                    fi[kk] = fp_i[(size_t)kk * n + y];
                }
            }
            for (int o = 0; o < rows_out; o++) {
                const float* src = stage_b + (size_t)o * n;
                float v = 0.0f;
                if (fc) {
#pragma unroll
                    for (int kk = 0; kk < IPFP; kk++) {
                        v = v + src[f0 + (uint32_t)kk] * col_fw[c][kk];
                    }
                } else {
#pragma unroll
                    for (int kk = 0; kk < IPFP; kk++) {
                        v = v + src[fi[kk]] * col_fw[c][kk];
                    }
                }
                out[(size_t)(x0 + o) * n + y] = v;
            }
        }
    };
    // E: drift along x of sub-tile x0's rows from the ring into stage_b
    auto drift_rows = [&](int x0, int rows_out) {
        const int wlo = x0 - hlo, whi = x0 + rows_out + hhi;
#pragma unroll
        for (int c = 0; c < MAXC; c++) {
            const uint32_t y = tid + (uint32_t)c * bs;
            if (c >= ncols || y >= n) {
                continue;
            }
            const float* w = col_dw[c];
            // the taps of output row x0 + o are rows r0 + o .. r0 + o + IT - 1
            const int r0 = x0 + (int)col_d0[c] - half;
            const bool win_ok = ((dcons >> c) & 1u) && r0 >= wlo &&
                                r0 + rows_out + IT - 2 < whi;
            auto window = [&]() {
                // every slot is read, so that the loads go out together
                // (one past the window reads some ring row, in bounds, and
                // is not used)
                float win[STRIP_MAX_T + IT - 1];
#pragma unroll
                for (int t = 0; t < (int)STRIP_MAX_T + IT - 1; t++) {
                    const uint32_t slot =
                        ((uint32_t)(r0 + t) + (uint32_t)hlo) & rmask;
                    win[t] = ring[__umul24(slot, n) + y];
                }
#pragma unroll
                for (int o = 0; o < (int)STRIP_MAX_T; o++) {
                    if (o < rows_out) {
                        float v = 0.0f;
#pragma unroll
                        for (int j = 0; j < IT; j++) {
                            v = v + win[o + j] * w[j];
                        }
                        stage_b[(size_t)o * n + y] = v;
                    }
                }
            };
            // a wavefront whose columns all read windows takes no per-lane
            // branch
            if (__all(win_ok)) {
                window();
                continue;
            }
            if (win_ok) {
                window();
                continue;
            }
            // taps not consecutive, or beyond the window (an error)
            uint32_t idx[IT];
            {
                float wdummy[IT];
                kick_taps<IT>(drift_off[y], n, idx, wdummy);
            }
            for (int o = 0; o < rows_out; o++) {
                const int x = x0 + o;
                float v = 0.0f;
#pragma unroll
                for (int j = 0; j < IT; j++) {
                    int xs = x + (int)idx[j] - half;
                    if (xs < wlo || xs >= whi) {
                        if (xs < 0 || xs >= in) {
                            continue;  // outside the mesh: skipped
                        }
                        atomicOr(d.err, 1u);
                        xs = xs < wlo ? wlo : whi - 1;
                    }
                    const uint32_t slot = ((uint32_t)xs + (uint32_t)hlo) & rmask;
                    v = v + ring[__umul24(slot, n) + y] * w[j];
                }
                stage_b[(size_t)o * n + y] = v;
            }
        }
    };

    // the first window: its rows outside the mesh, then its rows [k0, e0):
    // staged in phase A and kicked at once, or loaded and kicked below
    // sub-tile 0's own entering rows at most T at a time, sub-tile 0 kicking
    // its own
    zero_rows(s0 - hlo, s0 + ro0 + hhi);
    // sub-tile k kicks rows [ea, eb)
    int ea = 0, eb = 0;
    if (early) {
        kick_rows(k0, e0 - k0);
    } else {
        const int p1 = min(k1, s0 + hhi);
        for (int a = k0; a < p1; a += (int)T) {
            const int c = min((int)T, p1 - a);
            load_rows(a, c, stage_a);
            __syncthreads();
            kick_rows(a, c);
            __syncthreads();
        }
        ea = p1;
        eb = e0;
        if (eb > ea) {
            load_rows(ea, eb - ea, stage_a);
        }
        __syncthreads();
    }
    int prev_x0 = -1, prev_rows = 0;
    for (uint32_t k = 0; k < K; k++) {
        const int x0 = s0 + (int)(k * T);
        if (x0 >= in) break;
        const int rows_out = min((int)T, in - x0);
        const int whi = x0 + rows_out + hhi;
        // C, D beside the previous sub-tile's F
        if (prev_x0 >= 0) {
            fp_rows(prev_x0, prev_rows);
            // rows above the mesh entering this window
            zero_rows(max(x0 + hhi, in), whi);
        }
        if (eb > ea) {
            kick_rows(ea, eb - ea);
        }
        __syncthreads();
        // E beside the loads of the next sub-tile's entering rows
        drift_rows(x0, rows_out);
        ea = eb = 0;
        const int x0n = x0 + (int)T;
        if (k + 1 < K && x0n < in) {
            const int ron = min((int)T, in - x0n);
            ea = min(k1, x0n + hhi);
            eb = min(k1, x0n + ron + hhi);
            if (eb > ea) {
                load_rows(ea, eb - ea, stage_a);
            }
        }
        __syncthreads();
        prev_x0 = x0;
        prev_rows = rows_out;
    }
    if (prev_x0 >= 0) {
        fp_rows(prev_x0, prev_rows);
    }
    __syncthreads();  // the projection below reuses the staging rows

    // ---- H: strip 0 closes the step -----------------------------------------
    if (strip == 0) {
        step_close(d, parity, cnt, sh.red);
    }
    // ---- G: x projection of the strip's rows --------------------------------
    // The rows were stored by this workgroup before the barrier above.  In
    // chunks of C columns they are loaded (coalesced) into the ring and
    // staging rows, free by now, as the products v * w, two chunks in turn:
    // while one lane per row adds a chunk, continuing the reference's
    // std::inner_product order, the other wavefronts load the next.  Rows
    // in LDS are C + 4 words apart, so the lanes' reads of one column fall
    // in different banks.
    {
        const uint32_t room = (d.ringw + 2u * T) * n;
        uint32_t C = 4;
        while (C * 2 <= n && 2u * (uint32_t)srows * (C * 2 + 4) <= room) {
            C *= 2;
        }
        const uint32_t pst = C + 4;
        float* pb[2] = {ring, ring + (uint32_t)srows * pst};
        const uint32_t c4 = C >> 2;
        const uint32_t lc4 = (uint32_t)__builtin_ctz(c4);
        const uint32_t tot = (uint32_t)srows << lc4;
        // the loading threads: all but the summing wavefront, if any
        const uint32_t lt0 = bs > 64u ? 64u : 0u;
        const uint32_t nl = bs - lt0;
        // a thread's loads of one chunk are issued together; a load with
        // nothing to fetch reads the chunk's first word of row 0 instead
        auto load_chunk = [&](uint32_t c0, float* buf) {
            const uint32_t cw4 = min(C, n - c0) >> 2;
            const gfloat* src = out + (size_t)s0 * n + c0;
            for (uint32_t i0 = tid - lt0; i0 < tot; i0 += STRIP_PROJ_BATCH * nl) {
                vec4f v[STRIP_PROJ_BATCH];
#pragma unroll
                for (int b = 0; b < STRIP_PROJ_BATCH; b++) {
                    const uint32_t i = i0 + (uint32_t)b * nl;
                    const uint32_t r = i >> lc4, q = i & (c4 - 1u);
                    const bool ok = i < tot && q < cw4;
                    v[b] = *(const gfloat4*)(src + (ok ? (size_t)r * n + 4u * q : 0u));
                }
#pragma unroll
                for (int b = 0; b < STRIP_PROJ_BATCH; b++) {
                    const uint32_t i = i0 + (uint32_t)b * nl;
                    const uint32_t r = i >> lc4, q = i & (c4 - 1u);
                    if (i < tot && q < cw4) {
                        const vec4f w =
                            *reinterpret_cast<const vec4f*>(wsl + c0 + 4u * q);
                        vec4f p;
                        p.x = v[b].x * w.x;
                        p.y = v[b].y * w.y;
                        p.z = v[b].z * w.z;
                        p.w = v[b].w * w.w;
                        *reinterpret_cast<vec4f*>(buf + r * pst + 4u * q) = p;
                    }
                }
            }
        };
        if (tid >= lt0) {
            load_chunk(0u, pb[0]);
        }
        __syncthreads();
        float acc = 0.0f;
        uint32_t j = 0;
        for (uint32_t c0 = 0; c0 < n; c0 += C, j++) {
            if (c0 + C < n && tid >= lt0) {
                load_chunk(c0 + C, pb[(j + 1u) & 1u]);
            }
            if ((int)tid < srows) {
                const uint32_t cw4 = min(C, n - c0) >> 2;
                const vec4f* p4 =
                    reinterpret_cast<const vec4f*>(pb[j & 1u] + tid * pst);
#pragma unroll 8
                for (uint32_t q = 0; q < cw4; q++) {
                    const vec4f p = p4[q];
                    acc = acc + p.x;
                    acc = acc + p.y;
                    acc = acc + p.z;
                    acc = acc + p.w;
                }
            }
            __syncthreads();
        }
        if ((int)tid < srows) {
            const int x = s0 + (int)tid;
            proj_out[x] = acc;
            if (d.prof_hist && d.counter && cnt < d.hist_cap) {
                prof_hist[(size_t)cnt * n + x] = acc;
            }
        }
    }
}

// One workgroup per strip of a batch: blockmap[nmap] holds (instance << 16
// | strip) (STEP_NO_TILE: padding); nmap and the grid are multiples of 8,
// cut into 8 contiguous runs, one per XCD.  Registers are held to what two
// workgroups of 512 threads per CU leave (128 VGPRs), or, with four
// columns per thread (workgroups of at most 256 threads), to 256.
template <int IT, int IPFP, int MAXC>
__global__ void __launch_bounds__(STEP_MAX_THREADS, MAXC >= 4 ? 2 : 4)
strip_kernel(const InstDesc* __restrict__ descs,
             const uint32_t* __restrict__ blockmap, uint32_t nmap,
             uint32_t parity)
{
    extern __shared__ __align__(16) float lds[];
    __shared__ StepShared sh;
    const uint32_t logical = (blockIdx.x & 7u) * (nmap >> 3) + (blockIdx.x >> 3);
    const uint32_t m = blockmap[logical];
    if (m == STEP_NO_TILE) {
        return;
    }
    strip_body<IT, IPFP, MAXC>(descs[m >> 16], m & 0xFFFFu, parity, lds, sh);
}

}  // namespace vfps_impl
