// Body of the ElectricField kernel (wake potential and CSR spectrum),
// shared by the per-field kernel and the fused step.
#pragma once

#include "fft.hpp"
#include "field.hpp"
#include "reduce.hpp"

namespace vfps_impl {

// WAKE: wake potential of all bunches (profiles at bucket*spacing).
// CSR: CSR spectrum/power; alone it handles bunch blockIdx.x with its
// profile at offset 0 (src/PS/ElectricField.cpp:152-163); fused with WAKE
// (one bunch at offset 0) it reuses the wake's forward transform.
template <bool WAKE, bool CSR>
__device__ void field_body(const FieldParams& p, uint32_t block_bunch,
                           float2* lds)
{
    __shared__ float red[16];
    const uint32_t M = p.M;
    float2* A = lds;
    float2* B = lds + M;
    float* Af = reinterpret_cast<float*>(A);
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        A[i] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // padBunchProfiles (src/PS/ElectricField.cpp:255-263)
    if (WAKE) {
        for (uint32_t b = 0; b < p.nb; b++) {
            const uint32_t base = p.buckets[b] * p.spacing;
            for (uint32_t x = threadIdx.x; x < p.n; x += blockDim.x) {
                Af[base + x] = p.proj[(size_t)b * p.n + x];
            }
        }
    } else {
        const uint32_t b = block_bunch;
        for (uint32_t x = threadIdx.x; x < p.n; x += blockDim.x) {
            Af[x] = p.proj[(size_t)b * p.n + x];
        }
    }
    __syncthreads();
    if (p.padded) {
        for (uint32_t i = threadIdx.x; i < 2 * M; i += blockDim.x) {
            p.padded[i] = Af[i];
        }
    }
    if (p.filling) {
        // charge integral in FP64 (src/PS/PhaseSpace.cpp:200-239)
        __shared__ double dred[16];
        double tot = 0.0;
        for (uint32_t b = 0; b < p.nb; b++) {
            double acc = 0.0;
            for (uint32_t x = threadIdx.x; x < p.n; x += blockDim.x) {
                acc += (double)p.proj[(size_t)b * p.n + x] * (double)p.ws[x];
            }
            acc = block_sum(acc, dred);
            if (threadIdx.x == 0) {
                p.filling[b] = acc;
            }
            tot += acc;
        }
        if (threadIdx.x == 0) {
            p.filling[p.nb] = tot;
        }
    }
    if (!WAKE && !CSR) {
        // no impedance: the step only needs the integral and the counter
        if (threadIdx.x == 0 && p.counter) {
            *p.counter = *p.counter + 1;
        }
        return;
    }
    float2* Zf = fft_lds<false>(A, B, M, p.tw_m);
    float2* other = (Zf == A) ? B : A;
    float csr = 0.0f;
    const size_t sbase = CSR ? (size_t)(WAKE ? 0 : block_bunch) * 2 * M : 0;
    for (uint32_t k = threadIdx.x; k <= M / 2; k += blockDim.x) {
        if (k == 0) {
            const float2 z0 = Zf[0];
            const float f0 = z0.x + z0.y;  // F_0 (real)
            const float fm = z0.x - z0.y;  // F_M (real)
            if (CSR) {
                const float a0 = fabsf(f0), am = fabsf(fm);
                const float s0 = p.renorm[0] * p.Zc[0].x * (a0 * a0);
                const float sm = p.renorm[M] * p.Zc[M].x * (am * am);
                if (p.spectrum) {
                    p.spectrum[sbase] = s0;
                    p.spectrum[sbase + M] = sm;
                }
                csr += p.fdelta * s0 + p.fdelta * sm;
            }
            if (WAKE) {
                // W_0 = Z_0 F_0 with its imaginary part dropped (c2r
                // ignores it); W_M is never assigned by the reference and
                // is taken as zero.
                const float w0 = p.Z[0].x * f0;
                Zf[0] = make_float2(w0, w0);
            }
        } else {
            const uint32_t kp = M - k;
            float2 fk, fkp;
            r2c_pair(Zf[k], Zf[kp], p.tw_n[k], &fk, &fkp);
            if (CSR) {
                const float ak = hypotf(fk.x, fk.y);
                const float sk = p.renorm[k] * p.Zc[k].x * (ak * ak);
                csr += p.fdelta * sk;
                if (p.spectrum) p.spectrum[sbase + k] = sk;
                if (kp != k) {
                    const float akp = hypotf(fkp.x, fkp.y);
                    const float skp = p.renorm[kp] * p.Zc[kp].x * (akp * akp);
                    csr += p.fdelta * skp;
                    if (p.spectrum) p.spectrum[sbase + kp] = skp;
                }
            }
            if (WAKE) {
                const float2 wk = cmulf(p.Z[k], fk);
                const float2 wkp = cmulf(p.Z[kp], fkp);
                float2 yk, ykp;
                if (kp != k) {
                    c2r_pair(wk, wkp, p.tw_n[k], &yk, &ykp);
                    Zf[k] = yk;
                    Zf[kp] = ykp;
                } else {
                    Zf[k] = make_float2(2.0f * wk.x, -2.0f * wk.y);
                }
            }
        }
    }
    if (CSR) {
        if (p.spectrum) {
            for (uint32_t k = M + 1 + threadIdx.x; k < 2 * M; k += blockDim.x) {
                p.spectrum[sbase + k] = 0.0f;
            }
        }
        csr = block_sum(csr, red);
        if (threadIdx.x == 0 && p.power) {
            p.power[WAKE ? 0 : block_bunch] = csr;
        }
    }
    if (threadIdx.x == 0 && p.counter) {
        // per-step history: this step's CSR power, then count the step
        const uint32_t c = *p.counter;
        if (CSR && p.csr_hist && c < p.hist_cap) {
            p.csr_hist[c] = csr;
        }
        *p.counter = c + 1;
    }
    if (!WAKE) {
        return;
    }
    __syncthreads();
    float2* y = fft_lds<true>(Zf, other, M, p.tw_m);
    const float* yf = reinterpret_cast<const float*>(y);
    for (uint32_t b = 0; b < p.nb; b++) {
        const uint32_t base = p.buckets[b] * p.spacing;
        for (uint32_t x = threadIdx.x; x < p.n; x += blockDim.x) {
            const float v = p.scaling * yf[base + x];
            p.wake[(size_t)b * p.n + x] = v;
            if (p.wake2) {
                p.wake2[(size_t)b * p.n + x] = v;
            }
        }
    }
}


}  // namespace vfps_impl
