// Body of the ElectricField kernel (wake potential and CSR spectrum),
// shared by the per-field kernel and the fused step.
#pragma once

#include "fft.hpp"
#include "field.hpp"
#include "reduce.hpp"

namespace vfps_impl {

// WAKE: wake potential of all bunches (profiles at bucket*spacing,
// src/PS/ElectricField.cpp:186-263).
// CSR: CSR spectrum/power (src/PS/ElectricField.cpp:133-184); alone it
// handles bunch `block_bunch` with its profile at offset 0; fused with WAKE
// (one bunch at offset 0) it reuses the wake's forward transform.
// The transforms run in FP64 (fft.hpp); the spectrum F, the product
// W = Z F and the wake potential are rounded to FP32 where the reference
// holds them as float, and the float arithmetic on them follows the
// reference.  lds: field_lds_bytes(M): fft_slots(M) + M/2 double2 (data
// with a spare slot after every 8 elements, twiddle half table) for a
// power-of-two M, else 2 M (data, DFT scratch).  Runs with FFT_THREADS
// threads (the loops step by that constant: blockDim would be a load from
// the dispatch packet, waited for after every barrier).
constexpr int FIELD_PAIRS = 3;  // (k, M-k) pairs per thread: M <= 4096

#define VFPS_FSTAMP(k)                                                        \
    do {                                                                      \
        if (p.tstamp && blockIdx.x == 0) {                                    \
            __syncthreads();                                                  \
            if (threadIdx.x == 0) p.tstamp[k] = wall_clock64();               \
        }                                                                     \
    } while (0)

// A loaded value made opaque at this point: arithmetic on it cannot be
// moved up next to its load, where the wait for the load would then sit
// (before the barrier it is meant to overlap).
template <typename T>
__device__ __forceinline__ T opaque_after(T v)
{
    asm volatile("" : "+v"(v));
    return v;
}

template <bool WAKE, bool CSR>
__device__ void field_body(const FieldParams& p, uint32_t block_bunch,
                           double2* lds)
{
    __shared__ double red[16];
    VFPS_FSTAMP(0);
    const uint32_t M = p.M;
    // power-of-two M: radix-8 FFT with the half twiddle table in LDS;
    // other M: direct DFT with M entries of LDS scratch
    const bool pow2 = fft_pow2(M);
    // element i of the transform buffer: a spare slot after every 8
    // elements for the FFT (fft_slot), dense for the DFT
    const uint32_t psh = pow2 ? 3u : 31u;
    auto slot = [psh](uint32_t i) { return i + (i >> psh); };
    // real sample i of the padded array: element i/2, real part for even i
    auto rslot = [&](uint32_t i) { return 2u * slot(i >> 1) + (i & 1u); };
    double2* A = lds;
    double2* twl = lds + (pow2 ? fft_slots(M) : M);
    double* Ad = reinterpret_cast<double*>(A);
    // one bunch (the step's case): its profile and Simpson weights are
    // fetched first, with the tables and the step counter below, so that
    // one memory round trip serves the whole set-up
    constexpr int PV = 2;
    const bool one = p.nb == 1 && p.n <= PV * FFT_THREADS;
    float pv[PV] = {}, wv[PV] = {};
    uint32_t bk0 = WAKE && p.nb == 1 ? p.buckets[0] : 0u;
    if (one) {
        const uint32_t b = WAKE ? 0u : block_bunch;
#pragma unroll
        for (int k = 0; k < PV; k++) {
            const uint32_t x = threadIdx.x + (uint32_t)k * FFT_THREADS;
            if (x < p.n) {
                pv[k] = p.proj[(size_t)b * p.n + x];
                if (p.filling) {
                    wv[k] = p.ws[x];
                }
            }
        }
    }
    const uint32_t pin = p.counter_pingpong ? p.parity : 0u;
    const uint32_t pout = p.counter_pingpong ? (p.parity ^ 1u) : 0u;
    uint32_t cnt = 0;
    if (threadIdx.x == 0 && p.counter) {
        cnt = p.counter[pin];
    }
    for (uint32_t i = threadIdx.x; i < M; i += FFT_THREADS) {
        A[slot(i)] = make_double2(0.0, 0.0);
    }
    if ((WAKE || CSR) && pow2) {
        for (uint32_t i = threadIdx.x; i < M / 2; i += FFT_THREADS) {
            twl[i] = p.tw_m[i];
        }
    }
    // this thread's (k, M-k) pairs: tables fetched before the transform so
    // their latency hides behind it
    float2 zk[FIELD_PAIRS], zkp[FIELD_PAIRS];
    float ck[FIELD_PAIRS], ckp[FIELD_PAIRS], rk[FIELD_PAIRS], rkp[FIELD_PAIRS];
    double2 wn[FIELD_PAIRS];
#pragma unroll
    for (int j = 0; j < FIELD_PAIRS; j++) {
        const uint32_t k = threadIdx.x + j * FFT_THREADS;
        const uint32_t kp = M - k;
        if ((WAKE || CSR) && k <= M / 2) {
            wn[j] = p.tw_n[k];
            if (WAKE) {
                zk[j] = p.Z[k];
                zkp[j] = p.Z[kp];
            }
            if (CSR) {
                ck[j] = p.Zc[k].x;
                ckp[j] = p.Zc[kp].x;
                rk[j] = p.renorm[k];
                rkp[j] = p.renorm[kp];
            }
        }
    }
    __syncthreads();
    // the early loads are used from here on
    if (one) {
#pragma unroll
        for (int k = 0; k < PV; k++) {
            pv[k] = opaque_after(pv[k]);
            wv[k] = opaque_after(wv[k]);
        }
    }
    bk0 = opaque_after(bk0);
    cnt = opaque_after(cnt);
    const uint32_t base1 = bk0 * p.spacing;
    // padBunchProfiles (src/PS/ElectricField.cpp:255-263)
    if (one) {
        const uint32_t base = base1;
#pragma unroll
        for (int k = 0; k < PV; k++) {
            const uint32_t x = threadIdx.x + (uint32_t)k * FFT_THREADS;
            if (x < p.n) {
                Ad[rslot(base + x)] = (double)pv[k];
            }
        }
    } else if (WAKE) {
        for (uint32_t b = 0; b < p.nb; b++) {
            const uint32_t base = p.buckets[b] * p.spacing;
            for (uint32_t x = threadIdx.x; x < p.n; x += FFT_THREADS) {
                Ad[rslot(base + x)] = (double)p.proj[(size_t)b * p.n + x];
            }
        }
    } else {
        const uint32_t b = block_bunch;
        for (uint32_t x = threadIdx.x; x < p.n; x += FFT_THREADS) {
            Ad[rslot(x)] = (double)p.proj[(size_t)b * p.n + x];
        }
    }
    __syncthreads();
    if (p.padded) {
        for (uint32_t i = threadIdx.x; i < 2 * M; i += FFT_THREADS) {
            p.padded[i] = (float)Ad[rslot(i)];
        }
    }
    if (p.filling) {
        // charge integral in FP64 (src/PS/PhaseSpace.cpp:200-239)
        __shared__ double dred[16];
        double tot = 0.0;
        for (uint32_t b = 0; b < p.nb; b++) {
            double acc = 0.0;
            if (one) {
#pragma unroll
                for (int k = 0; k < PV; k++) {
                    const uint32_t x = threadIdx.x + (uint32_t)k * FFT_THREADS;
                    if (x < p.n) {
                        acc += (double)pv[k] * (double)wv[k];
                    }
                }
            } else {
                for (uint32_t x = threadIdx.x; x < p.n; x += FFT_THREADS) {
                    acc += (double)p.proj[(size_t)b * p.n + x] * (double)p.ws[x];
                }
            }
            acc = block_sum_dpp<FFT_THREADS>(acc, dred);
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
            p.counter[pout] = cnt + 1;
        }
        return;
    }
    VFPS_FSTAMP(1);
    if (pow2) {
        fft_lds<false>(A, M, twl);
    } else {
        dft_lds<false>(A, twl, M, p.tw_m);
    }
    VFPS_FSTAMP(2);
    double2* Zf = A;
    // CSR power: each bin's float term as the reference forms it, summed
    // in FP64 (the reference's sequential float sum carries a rounding
    // error of up to nmax*eps; the FP64 sum is the exact sum of the terms)
    double csr = 0.0;
    const size_t sbase = CSR ? (size_t)(WAKE ? 0 : block_bunch) * 2 * M : 0;
    // CSR spectrum term of one bin from the float spectrum value
    auto csr_bin = [&](uint32_t k, float2 f, float rn, float zc) {
        const float a = hypotf(f.x, f.y);  // std::norm via std::abs
        const float s = rn * zc * (a * a);
        if (p.spectrum) {
            p.spectrum[sbase + k] = s;
        }
        return (double)(p.fdelta * s);
    };
#pragma unroll
    for (int j = 0; j < FIELD_PAIRS; j++) {
        const uint32_t k = threadIdx.x + j * FFT_THREADS;
        if (k > M / 2) {
            continue;
        }
        if (k == 0) {
            const double2 z0 = Zf[0];  // slot 0 is element 0
            const float f0 = (float)(z0.x + z0.y);  // F_0 (real)
            const float fm = (float)(z0.x - z0.y);  // F_M (real)
            if (CSR) {
                csr += csr_bin(0, make_float2(f0, 0.0f), rk[j], ck[j]);
                csr += csr_bin(M, make_float2(fm, 0.0f), rkp[j], ckp[j]);
            }
            if (WAKE) {
                // W_0 = Z_0 F_0; c2r ignores its imaginary part.  W_M is
                // never assigned by the reference and is taken as zero.
                const float2 w0 = cmulf(zk[j], make_float2(f0, 0.0f));
                Zf[0] = make_double2((double)w0.x, (double)w0.x);
            }
        } else {
            const uint32_t kp = M - k;
            double2 fk, fkp;
            r2c_pair(Zf[slot(k)], Zf[slot(kp)], wn[j], &fk, &fkp);
            const float2 f1 = make_float2((float)fk.x, (float)fk.y);
            const float2 f2 = make_float2((float)fkp.x, (float)fkp.y);
            if (CSR) {
                csr += csr_bin(k, f1, rk[j], ck[j]);
                if (kp != k) {
                    csr += csr_bin(kp, f2, rkp[j], ckp[j]);
                }
            }
            if (WAKE) {
                const float2 w1 = cmulf(zk[j], f1);
                const float2 w2 = cmulf(zkp[j], f2);
                const double2 wk = make_double2(w1.x, w1.y);
                const double2 wkp = make_double2(w2.x, w2.y);
                if (kp != k) {
                    double2 yk, ykp;
                    c2r_pair(wk, wkp, wn[j], &yk, &ykp);
                    Zf[slot(k)] = yk;
                    Zf[slot(kp)] = ykp;
                } else {
                    Zf[slot(k)] = make_double2(2.0 * wk.x, -2.0 * wk.y);
                }
            }
        }
    }
    if (CSR) {
        if (p.spectrum) {
            for (uint32_t k = M + 1 + threadIdx.x; k < 2 * M; k += FFT_THREADS) {
                p.spectrum[sbase + k] = 0.0f;
            }
        }
        csr = block_sum_dpp<FFT_THREADS>(csr, red);
        if (threadIdx.x == 0 && p.power) {
            p.power[WAKE ? 0 : block_bunch] = (float)csr;
        }
    }
    if (threadIdx.x == 0 && p.counter) {
        // per-step history: this step's CSR power, then count the step
        if (CSR && p.csr_hist && cnt < p.hist_cap) {
            p.csr_hist[cnt] = (float)csr;
        }
        p.counter[pout] = cnt + 1;
    }
    if (!WAKE) {
        return;
    }
    __syncthreads();
    VFPS_FSTAMP(3);
    if (pow2) {
        fft_lds<true>(A, M, twl);
    } else {
        dft_lds<true>(A, twl, M, p.tw_m);
    }
    VFPS_FSTAMP(4);
    const double* yd = reinterpret_cast<const double*>(A);
    if (p.wake_padded) {
        for (uint32_t i = threadIdx.x; i < 2 * M; i += FFT_THREADS) {
            p.wake_padded[i] = (float)yd[rslot(i)];
        }
    }
    for (uint32_t b = 0; b < p.nb; b++) {
        const uint32_t base = p.nb == 1 ? base1 : p.buckets[b] * p.spacing;
        for (uint32_t x = threadIdx.x; x < p.n; x += FFT_THREADS) {
            const float v = p.scaling * (float)yd[rslot(base + x)];
            p.wake[(size_t)b * p.n + x] = v;
            if (p.wake2) {
                p.wake2[(size_t)b * p.n + x] = v;
            }
        }
    }
    VFPS_FSTAMP(5);
}

}  // namespace vfps_impl
