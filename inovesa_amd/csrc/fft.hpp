// In-LDS FFT building blocks for the wake-potential convolution.
//
// A real sequence of length N = 2M is transformed through one complex FFT
// of length M (even samples in the real parts, odd samples in the
// imaginary parts).  The complex FFT is a Stockham autosort with radix-8
// passes (and one radix-4 or radix-2 pass), in place in one LDS buffer
// with a spare slot after every 8 elements, with no bit reversal.
//
// The butterflies run in FP64.  Inputs and outputs stay FP32 exactly where
// the reference's arrays are float (profile, spectrum F, W = Z F, wake
// potential): an FP32 FFT leaves a white rounding floor of about
// eps*|profile| in every bin, which the steep high-frequency impedance
// amplifies into ~1e-5 relative errors of the wake; computing the
// transform in FP64 gives the correctly rounded FP32 spectrum instead.
#pragma once

#include <hip/hip_runtime.h>

namespace vfps_impl {

__host__ __device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

constexpr uint32_t FFT_THREADS = 1024;
// longest complex transform (real length 8192)
constexpr uint32_t FFT_MAX_M = 4096;

// LDS slot of element i of a power-of-two transform buffer: one spare slot
// after every 8 elements, so that the stride-8 stores of a radix-8 pass
// spread over the banks
__host__ __device__ __forceinline__ uint32_t fft_slot(uint32_t i)
{
    return i + (i >> 3);
}
// slots a buffer of M elements spans
__host__ __device__ constexpr uint32_t fft_slots(uint32_t M)
{
    return M + M / 8;
}

// The butterflies' DFT of 4 (forward exp(-i..), INV exp(+i..)), in place
template <bool INV>
__device__ __forceinline__ void dft4(double2& c0, double2& c1, double2& c2,
                                     double2& c3)
{
    const double2 t0 = cadd(c0, c2), t1 = csub(c0, c2);
    const double2 t2 = cadd(c1, c3), d = csub(c1, c3);
    // forward: -i d; inverse: +i d
    const double2 t3 = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
    c0 = cadd(t0, t2);
    c1 = cadd(t1, t3);
    c2 = csub(t0, t2);
    c3 = csub(t1, t3);
}

// One radix-R Stockham pass (R = 2, 4, 8) over M = 2^lgM elements in the
// slots of `a`, in place: every butterfly reads its R inputs, the
// workgroup synchronises, then each writes its R outputs.  Butterfly j
// (j < M/R) of the pass with Ns = 2^lgNs takes inputs j + k M/R, scales
// input k by w^(k jm M/(Ns R)) (jm = j mod Ns, w = exp(-2 pi i / M)), and
// writes output q to (j - jm) R + jm + q Ns.  The twiddle w^(jm M/(Ns R))
// comes from the half table tw (its index is below M/2), its powers by
// complex products in FP64.
template <bool INV, int R>
__device__ __forceinline__ void fft_pass(double2* a, uint32_t lgM,
                                         uint32_t lgNs, const double2* tw)
{
    constexpr int LGR = R == 8 ? 3 : (R == 4 ? 2 : 1);
    // butterflies per thread: M/R <= FFT_MAX_M/R over FFT_THREADS threads
    constexpr int NB = (int)((FFT_MAX_M / R + FFT_THREADS - 1) / FFT_THREADS);
    const uint32_t nbf = 1u << (lgM - LGR);
    const uint32_t Ns = 1u << lgNs;
    const uint32_t lgts = lgM - lgNs - LGR;
    double2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint32_t j = threadIdx.x + (uint32_t)b * FFT_THREADS;
        if (j < nbf) {
#pragma unroll
            for (int k = 0; k < R; k++) {
                v[b][k] = a[fft_slot(j + (uint32_t)k * nbf)];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint32_t j = threadIdx.x + (uint32_t)b * FFT_THREADS;
        if (j >= nbf) {
            continue;
        }
        double2* x = v[b];
        const uint32_t jm = j & (Ns - 1u);
        if (jm != 0u) {
            double2 w1 = tw[jm << lgts];
            if (INV) {
                w1 = cconj(w1);
            }
            x[1] = cmul(x[1], w1);
            if (R >= 4) {
                const double2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
                x[2] = cmul(x[2], w2);
                x[3] = cmul(x[3], w3);
                if (R == 8) {
                    const double2 w4 = cmul(w2, w2), w5 = cmul(w4, w1);
                    const double2 w6 = cmul(w3, w3), w7 = cmul(w4, w3);
                    x[4] = cmul(x[4], w4);
                    x[5] = cmul(x[5], w5);
                    x[6] = cmul(x[6], w6);
                    x[7] = cmul(x[7], w7);
                }
            }
        }
        double2 o[R];
        if (R == 2) {
            o[0] = cadd(x[0], x[1]);
            o[1] = csub(x[0], x[1]);
        } else if (R == 4) {
            dft4<INV>(x[0], x[1], x[2], x[3]);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                o[q] = x[q];
            }
        } else {
            // DFT of 8 as two DFTs of 4: e_k = x_k + x_{k+4} gives the even
            // outputs, (x_k - x_{k+4}) exp(-+ i pi k / 4) the odd ones
            const double r = 0.70710678118654752440;
            double2 e[4], d[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                e[k] = cadd(x[k], x[k + 4]);
                d[k] = csub(x[k], x[k + 4]);
            }
            d[1] = INV ? make_double2(r * (d[1].x - d[1].y), r * (d[1].x + d[1].y))
                       : make_double2(r * (d[1].x + d[1].y), r * (d[1].y - d[1].x));
            d[2] = INV ? make_double2(-d[2].y, d[2].x) : make_double2(d[2].y, -d[2].x);
            d[3] = INV ? make_double2(-r * (d[3].x + d[3].y), r * (d[3].x - d[3].y))
                       : make_double2(r * (d[3].y - d[3].x), -r * (d[3].x + d[3].y));
            dft4<INV>(e[0], e[1], e[2], e[3]);
            dft4<INV>(d[0], d[1], d[2], d[3]);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                o[2 * q] = e[q];
                o[2 * q + 1] = d[q];
            }
        }
        const uint32_t dst = ((j - jm) << LGR) + jm;
#pragma unroll
        for (int q = 0; q < R; q++) {
            a[fft_slot(dst + (uint32_t)q * Ns)] = o[q];
        }
    }
    __syncthreads();
}

// Complex FFT of length M (power of two, 2 <= M <= FFT_MAX_M) in place in
// the slots (fft_slot) of the LDS buffer `a`: Stockham autosort, radix-8
// passes with a radix-4 or radix-2 pass where log2(M) is not a multiple of
// 3, so the output is in natural order without a bit reversal.  tw: LDS
// half table exp(-2 pi i t / M), t < M/2.  Forward uses exp(-i..), INV
// exp(+i..); unnormalised.  Must be called by the whole workgroup of
// FFT_THREADS threads.
template <bool INV>
__device__ void fft_lds(double2* a, uint32_t M, const double2* tw)
{
    const uint32_t lgM = 31u - (uint32_t)__clz(M);
    uint32_t lgNs = 0;
    while (lgNs < lgM) {
        const uint32_t left = lgM - lgNs;
        if (left >= 3u && (left % 3u == 0u || left > 4u)) {
            fft_pass<INV, 8>(a, lgM, lgNs, tw);
            lgNs += 3;
        } else if (left >= 2u) {
            fft_pass<INV, 4>(a, lgM, lgNs, tw);
            lgNs += 2;
        } else {
            fft_pass<INV, 2>(a, lgM, lgNs, tw);
            lgNs += 1;
        }
    }
}

// Complex DFT of length M (any M) in place in the LDS buffer `a`, by the
// direct sum out[k] = sum_j a[j] w^(j k) in FP64, for the lengths the
// radix-4/2 FFT does not take (the reference's FFTW plans any length, e.g.
// nmax = 768 = 3 * 256 when RoundPadding is off).  tmp: LDS scratch of M
// entries; tw: the full table exp(-2 pi i t / M), t < M (global memory: the
// powers of w are picked by index, (j k) mod M, so no rounding accumulates
// across the sum).  Forward uses exp(-i..), INV exp(+i..); unnormalised.
// Must be called by the whole workgroup.
template <bool INV>
__device__ void dft_lds(double2* a, double2* tmp, uint32_t M,
                        const double2* __restrict__ tw)
{
    for (uint32_t i = threadIdx.x; i < M; i += FFT_THREADS) {
        tmp[i] = a[i];
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < M; k += FFT_THREADS) {
        double re = 0.0, im = 0.0;
        uint32_t t = 0;
        for (uint32_t j = 0; j < M; j++) {
            const double2 x = tmp[j];
            const double2 w = tw[t];
            const double wy = INV ? -w.y : w.y;
            re += x.x * w.x - x.y * wy;
            im += x.x * wy + x.y * w.x;
            t += k;
            if (t >= M) t -= M;
        }
        a[k] = make_double2(re, im);
    }
    __syncthreads();
}

__host__ __device__ __forceinline__ bool fft_pow2(uint32_t M)
{
    return M >= 1u && (M & (M - 1u)) == 0u;
}

// Spectrum F_k (k = 0..M) of the real input from Z = FFT_M(z):
// for the pair (k, M-k), 0 < k <= M/2.  wk = exp(-2 pi i k / N).
__device__ __forceinline__ void r2c_pair(double2 zk, double2 zkp, double2 wk,
                                         double2* fk, double2* fkp)
{
    const double2 e = make_double2(0.5 * (zk.x + zkp.x), 0.5 * (zk.y - zkp.y));
    // o = -i/2 (zk - conj(zkp))
    const double2 dd = make_double2(zk.x - zkp.x, zk.y + zkp.y);
    const double2 o = make_double2(0.5 * dd.y, -0.5 * dd.x);
    *fk = cadd(e, cmul(wk, o));
    // F_{M-k} = conj(e) + (-conj(wk)) conj(o)
    *fkp = csub(cconj(e), cmul(cconj(wk), cconj(o)));
}

// Inverse pre-pass: Y_k = (W_k + conj W_{M-k}) + i conj(wk) (W_k - conj W_{M-k})
// and Y_{M-k} = (W_{M-k} + conj W_k) - i wk (W_{M-k} - conj W_k).
__device__ __forceinline__ void c2r_pair(double2 wk_, double2 wkp_, double2 wk,
                                         double2* yk, double2* ykp)
{
    const double2 s = make_double2(wk_.x + wkp_.x, wk_.y - wkp_.y);
    const double2 d = make_double2(wk_.x - wkp_.x, wk_.y + wkp_.y);
    const double2 t = cmul(cconj(wk), d);  // conj(wk) d
    *yk = make_double2(s.x - t.y, s.y + t.x);  // s + i t
    const double2 s2 = make_double2(wkp_.x + wk_.x, wkp_.y - wk_.y);
    const double2 d2 = make_double2(wkp_.x - wk_.x, wkp_.y + wk_.y);
    const double2 t2 = cmul(wk, d2);  // wk d2
    *ykp = make_double2(s2.x + t2.y, s2.y - t2.x);  // s2 - i t2
}

}  // namespace vfps_impl
