// In-LDS FFT building blocks for the wake-potential convolution.
//
// A real sequence of length N = 2M is transformed through one complex FFT
// of length M (even samples in the real parts, odd samples in the
// imaginary parts).  The complex FFT is a Stockham autosort with radix-4
// stages and one radix-2 stage when log2(M) is odd, in place in one LDS
// buffer, with no bit reversal.
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

// twiddle exp(-2 pi i t / M) for t < M from the half table
// tw[t] (t < M/2), using exp(-2 pi i (t + M/2) / M) = -exp(-2 pi i t / M)
__device__ __forceinline__ double2 twiddle(const double2* tw, uint32_t t,
                                           uint32_t half)
{
    if (t < half) {
        return tw[t];
    }
    const double2 w = tw[t - half];
    return make_double2(-w.x, -w.y);
}

// Largest number of butterflies one thread handles per stage (the field
// kernels run 1024 threads: M <= 8192)
constexpr int FFT_MAX_BF = 2;
constexpr uint32_t FFT_THREADS = 1024;

// Complex FFT of length M (power of two, 2 <= M <= 4 * FFT_MAX_BF *
// blockDim.x) in place in the LDS buffer `a`.  Each stage reads all its
// inputs into registers, synchronises, then writes the outputs to their
// Stockham positions, so one buffer suffices.  tw: LDS half table
// exp(-2 pi i t / M), t < M/2.  Forward uses exp(-i..), INV exp(+i..);
// unnormalised.  Must be called by the whole workgroup.
template <bool INV>
__device__ void fft_lds(double2* a, uint32_t M, const double2* tw)
{
    const uint32_t half = M / 2;
    uint32_t Ns = 1;
    while (Ns < M) {
        const uint32_t R = ((M / Ns) % 4u == 0u) ? 4u : 2u;
        const uint32_t nbf = M / R;
        const uint32_t tscale = M / (Ns * R);
        double2 v[FFT_MAX_BF][4];
#pragma unroll
        for (int b = 0; b < FFT_MAX_BF; b++) {
            const uint32_t j = threadIdx.x + b * blockDim.x;
            if (j < nbf) {
                v[b][0] = a[j];
                v[b][1] = a[j + nbf];
                if (R == 4) {
                    v[b][2] = a[j + 2 * nbf];
                    v[b][3] = a[j + 3 * nbf];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < FFT_MAX_BF; b++) {
            const uint32_t j = threadIdx.x + b * blockDim.x;
            if (j >= nbf) {
                continue;
            }
            const uint32_t jm = j & (Ns - 1);
            const uint32_t tstep = jm * tscale;
            const uint32_t idxD = (j - jm) * R + jm;
            if (R == 4) {
                double2 v0 = v[b][0], v1 = v[b][1], v2 = v[b][2], v3 = v[b][3];
                if (tstep != 0) {
                    double2 w1 = twiddle(tw, tstep, half);
                    double2 w2 = twiddle(tw, 2 * tstep, half);
                    double2 w3 = twiddle(tw, 3 * tstep, half);
                    if (INV) {
                        w1 = cconj(w1);
                        w2 = cconj(w2);
                        w3 = cconj(w3);
                    }
                    v1 = cmul(v1, w1);
                    v2 = cmul(v2, w2);
                    v3 = cmul(v3, w3);
                }
                const double2 a0 = cadd(v0, v2), a1 = csub(v0, v2);
                const double2 a2 = cadd(v1, v3), d = csub(v1, v3);
                // forward: -i d; inverse: +i d
                const double2 a3 = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
                a[idxD] = cadd(a0, a2);
                a[idxD + Ns] = cadd(a1, a3);
                a[idxD + 2 * Ns] = csub(a0, a2);
                a[idxD + 3 * Ns] = csub(a1, a3);
            } else {
                double2 v0 = v[b][0], v1 = v[b][1];
                if (tstep != 0) {
                    double2 w1 = twiddle(tw, tstep, half);
                    if (INV) {
                        w1 = cconj(w1);
                    }
                    v1 = cmul(v1, w1);
                }
                a[idxD] = cadd(v0, v1);
                a[idxD + Ns] = csub(v0, v1);
            }
        }
        __syncthreads();
        Ns *= R;
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
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        tmp[i] = a[i];
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < M; k += blockDim.x) {
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
