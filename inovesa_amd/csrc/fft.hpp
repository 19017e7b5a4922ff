// In-LDS FFT building blocks for the wake-potential convolution.
//
// A real sequence of length N = 2M is transformed through one complex FFT
// of length M (even samples in the real parts, odd samples in the
// imaginary parts).  The complex FFT is a Stockham autosort with radix-4
// stages and one radix-2 stage when log2(M) is odd; it ping-pongs between
// two LDS buffers and needs no bit reversal.
#pragma once

#include <hip/hip_runtime.h>

namespace vfps_impl {

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 addf2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 subf2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Complex FFT of length M (power of two, >= 2) held in LDS buffer `a`;
// `b` is scratch of the same size.  tw[t] = exp(-2 pi i t / M), t < M.
// Forward uses exp(-i..), INV uses exp(+i..); unnormalised.  Returns the
// buffer holding the result.  Must be called by the whole workgroup.
template <bool INV>
__device__ float2* fft_lds(float2* a, float2* b, uint32_t M,
                           const float2* __restrict__ tw)
{
    uint32_t Ns = 1;
    while (Ns < M) {
        const uint32_t R = ((M / Ns) % 4u == 0u) ? 4u : 2u;
        const uint32_t nbf = M / R;
        const uint32_t tscale = M / (Ns * R);
        for (uint32_t j = threadIdx.x; j < nbf; j += blockDim.x) {
            const uint32_t jm = j & (Ns - 1);
            const uint32_t tstep = jm * tscale;
            const uint32_t idxD = (j - jm) * R + jm;
            if (R == 4) {
                float2 v0 = a[j];
                float2 v1 = a[j + nbf];
                float2 v2 = a[j + 2 * nbf];
                float2 v3 = a[j + 3 * nbf];
                if (tstep != 0) {
                    float2 w1 = tw[tstep], w2 = tw[2 * tstep], w3 = tw[3 * tstep];
                    if (INV) {
                        w1 = conjf2(w1);
                        w2 = conjf2(w2);
                        w3 = conjf2(w3);
                    }
                    v1 = cmulf(v1, w1);
                    v2 = cmulf(v2, w2);
                    v3 = cmulf(v3, w3);
                }
                const float2 a0 = addf2(v0, v2), a1 = subf2(v0, v2);
                const float2 a2 = addf2(v1, v3), d = subf2(v1, v3);
                // forward: -i d; inverse: +i d
                const float2 a3 = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
                b[idxD] = addf2(a0, a2);
                b[idxD + Ns] = addf2(a1, a3);
                b[idxD + 2 * Ns] = subf2(a0, a2);
                b[idxD + 3 * Ns] = subf2(a1, a3);
            } else {
                float2 v0 = a[j];
                float2 v1 = a[j + nbf];
                if (tstep != 0) {
                    float2 w1 = tw[tstep];
                    if (INV) {
                        w1 = conjf2(w1);
                    }
                    v1 = cmulf(v1, w1);
                }
                b[idxD] = addf2(v0, v1);
                b[idxD + Ns] = subf2(v0, v1);
            }
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// Spectrum F_k (k = 0..M) of the real input from Z = FFT_M(z):
// for the pair (k, M-k), 0 < k <= M/2.  wk = exp(-2 pi i k / N).
__device__ __forceinline__ void r2c_pair(float2 zk, float2 zkp, float2 wk,
                                         float2* fk, float2* fkp)
{
    const float2 e = make_float2(0.5f * (zk.x + zkp.x), 0.5f * (zk.y - zkp.y));
    // o = -i/2 (zk - conj(zkp))
    const float2 dd = make_float2(zk.x - zkp.x, zk.y + zkp.y);
    const float2 o = make_float2(0.5f * dd.y, -0.5f * dd.x);
    *fk = addf2(e, cmulf(wk, o));
    // F_{M-k} = conj(e) + (-conj(wk)) conj(o)
    *fkp = subf2(conjf2(e), cmulf(conjf2(wk), conjf2(o)));
}

// Inverse pre-pass: Y_k = (W_k + conj W_{M-k}) + i conj(wk) (W_k - conj W_{M-k})
// and Y_{M-k} = (W_{M-k} + conj W_k) - i wk (W_{M-k} - conj W_k).
__device__ __forceinline__ void c2r_pair(float2 wk_, float2 wkp_, float2 wk,
                                         float2* yk, float2* ykp)
{
    const float2 s = make_float2(wk_.x + wkp_.x, wk_.y - wkp_.y);
    const float2 d = make_float2(wk_.x - wkp_.x, wk_.y + wkp_.y);
    const float2 t = cmulf(conjf2(wk), d);  // conj(wk) d
    *yk = make_float2(s.x - t.y, s.y + t.x);  // s + i t
    const float2 s2 = make_float2(wkp_.x + wk_.x, wkp_.y - wk_.y);
    const float2 d2 = make_float2(wkp_.x - wk_.x, wkp_.y + wk_.y);
    const float2 t2 = cmulf(wk, d2);  // wk d2
    *ykp = make_float2(s2.x + t2.y, s2.y - t2.x);  // s2 - i t2
}

}  // namespace vfps_impl
