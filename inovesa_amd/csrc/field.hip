// ElectricField (inc/PS/ElectricField.hpp, src/PS/ElectricField.cpp):
// padded bunch profile -> real FFT -> x Z(n) -> inverse real FFT -> wake
// potential, and the CSR spectrum/power from the same forward transform.
//
// One workgroup runs a whole transform pair in LDS (fft.hpp); the complex
// multiply by Z, the Hermitian post/pre passes, the scaling and the CSR
// spectrum are fused into the pass between the two FFTs.
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "field_body.hpp"
#include "maps.hpp"
#include "reduce.hpp"

namespace vfps_impl {

template <bool WAKE, bool CSR>
__global__ void __launch_bounds__(1024)
field_kernel(FieldParams p)
{
    extern __shared__ double2 lds[];
    field_body<WAKE, CSR>(p, blockIdx.x, lds);
}

size_t field_lds_bytes(uint32_t M)
{
    return sizeof(double2) * (fft_pow2(M) ? fft_slots(M) + M / 2 : 2 * M);
}

int launch_field(const FieldParams& p, bool wake, bool csr, uint32_t blocks,
                 hipStream_t st)
{
    const size_t lds = field_lds_bytes(p.M);
    if (wake && csr) {
        hipLaunchKernelGGL((field_kernel<true, true>), dim3(1), dim3(FFT_THREADS), lds,
                           st, p);
    } else if (wake) {
        hipLaunchKernelGGL((field_kernel<true, false>), dim3(1), dim3(FFT_THREADS),
                           lds, st, p);
    } else {
        hipLaunchKernelGGL((field_kernel<false, true>), dim3(blocks),
                           dim3(FFT_THREADS), lds, st, p);
    }
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int field_set_lds_limit(size_t bytes)
{
    if (bytes <= 65536) {
        return VFPS_OK;
    }
    VFPS_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&field_kernel<true, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    VFPS_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&field_kernel<true, false>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    VFPS_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&field_kernel<false, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return VFPS_OK;
}

// twiddle tables: e^{-2 pi i t/M} (t < M) and e^{-2 pi i k/N} (k < M),
// in double
void make_twiddles(uint32_t M, std::vector<double2>& tw_m,
                   std::vector<double2>& tw_n)
{
    const double pi = 3.141592653589793238462643383279502884;
    tw_m.resize(M);
    tw_n.resize(M);
    for (uint32_t t = 0; t < M; t++) {
        double a = -2.0 * pi * (double)t / (double)M;
        tw_m[t] = make_double2(std::cos(a), std::sin(a));
        double b = -2.0 * pi * (double)t / (2.0 * (double)M);
        tw_n[t] = make_double2(std::cos(b), std::sin(b));
    }
}

// CSR renormalisation per frequency (src/PS/ElectricField.cpp:166-172)
std::vector<float> csr_renorm(uint32_t nmax, float delta0, double qscale,
                              float fc, float* fdelta)
{
    Ruler fax(nmax, 0.0f, 1.0f / delta0);
    const float scale_hz = (float)(2.99792458e8 / (double)(float)qscale);
    const float renorm0 = delta0 * delta0;
    std::vector<float> r(nmax);
    for (uint32_t i = 0; i < nmax; i++) {
        float rn = renorm0;
        if (fc > 0.0f) {
            float q = scale_hz * fax.data[i] / fc;
            double e = 1.0 - std::exp(-std::pow((double)q, 2.0));
            rn = (float)((double)rn * e);
        }
        r[i] = rn;
    }
    *fdelta = fax.delta;
    return r;
}

FieldParams field_params(vfps_field* f)
{
    FieldParams p{};
    p.nb = f->nb;
    p.n = f->ps->n;
    p.M = f->nmax / 2;
    p.spacing = f->spacing;
    p.buckets = f->buckets_d;
    p.scaling = f->wakescaling;
    p.proj = f->ps->proj;
    p.ws = f->ps->ws;
    p.Z = f->z->z;
    p.tw_m = f->tw_m;
    p.tw_n = f->tw_n;
    p.wake = f->wake;
    p.Zc = f->z->z;
    p.renorm = f->csr_renorm;
    p.fdelta = f->fdelta;
    return p;
}

int launch_wake_potential(vfps_field* f, float* dst)
{
    FieldParams p = field_params(f);
    p.wake2 = dst;
    p.padded = f->padded;
    p.wake_padded = f->wake_padded;
    return launch_field(p, true, false, 1, f->dev->stream);
}

// ---- the wake potential without allocation ---------------------------------
// Scratch of vfps_wake_potential_noalloc: the FFT twiddle tables and the
// bucket numbers, in a file-scope device array (one copy per GPU, held until
// the process ends).  Host code never uses the array's name as an address:
// its device address is taken once per device (noalloc_scratch) and that
// pointer goes to the kernels.
constexpr uint32_t NOALLOC_MAX_NMAX = 8192;
constexpr uint32_t NOALLOC_MAX_BUNCHES = 64;
constexpr size_t NOALLOC_TW_BYTES = sizeof(double2) * (NOALLOC_MAX_NMAX / 2);
constexpr size_t NOALLOC_BYTES =
    2 * NOALLOC_TW_BYTES + sizeof(uint32_t) * NOALLOC_MAX_BUNCHES;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static void* noalloc_ptr[NOALLOC_MAX_DEVICES];

// the device address of `scratch` on dev (taken on the first call for that
// device, with that device current)
static int noalloc_scratch(const vfps_dev* dev, unsigned char** out)
{
    VFPS_CHECK_ARG(dev->ordinal >= 0 && dev->ordinal < NOALLOC_MAX_DEVICES,
                   "vfps_wake_potential_noalloc: device ordinal");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev->ordinal]) {
        int cur = 0;
        VFPS_HIP(hipGetDevice(&cur));
        if (cur != dev->ordinal) VFPS_HIP(hipSetDevice(dev->ordinal));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev->ordinal) hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            return fail(VFPS_ERR_HIP, std::string("hipGetSymbolAddress: ") +
                                          hipGetErrorString(e));
        }
        noalloc_ptr[dev->ordinal] = ptr;
    }
    *out = static_cast<unsigned char*>(noalloc_ptr[dev->ordinal]);
    return VFPS_OK;
}

struct NoAllocBuckets {
    uint32_t b[NOALLOC_MAX_BUNCHES];
};

// the twiddle tables of make_twiddles and the bucket numbers, into the
// scratch (every entry the field kernel reads is written: nothing carries
// over from an earlier call)
__global__ void noalloc_setup_kernel(uint32_t M, double2* tw_m, double2* tw_n,
                                     uint32_t nb, NoAllocBuckets bk,
                                     uint32_t* buckets)
{
    const double pi = 3.141592653589793238462643383279502884;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < M;
         t += gridDim.x * blockDim.x) {
        const double a = -2.0 * pi * (double)t / (double)M;
        const double b = -2.0 * pi * (double)t / (2.0 * (double)M);
        tw_m[t] = make_double2(cos(a), sin(a));
        tw_n[t] = make_double2(cos(b), sin(b));
    }
    if (blockIdx.x == 0 && threadIdx.x < nb) {
        buckets[threadIdx.x] = bk.b[threadIdx.x];
    }
}

}  // namespace vfps_impl

using namespace vfps_impl;

extern "C" {

int vfps_wake_potential_noalloc(vfps_dev* dev, const float* proj, uint32_t nb,
                                uint32_t n, const vfps_cplx* z, uint32_t nmax,
                                const uint32_t* buckets, uint32_t spacing_bins,
                                float wakescaling, float* wake)
{
    VFPS_CHECK_ARG(dev && proj && z && buckets && wake,
                   "vfps_wake_potential_noalloc: NULL argument");
    VFPS_CHECK_ARG(nb >= 1 && nb <= NOALLOC_MAX_BUNCHES && n >= 1,
                   "vfps_wake_potential_noalloc: 1 to 64 bunches of n >= 1 "
                   "bins");
    VFPS_CHECK_ARG(nmax >= 4 && (nmax & (nmax - 1)) == 0,
                   "vfps_wake_potential_noalloc: FFT length must be a power "
                   "of two >= 4");
    VFPS_CHECK_ARG(nmax <= NOALLOC_MAX_NMAX,
                   "vfps_wake_potential_noalloc: FFT length exceeds the "
                   "scratch maximum of 8192");
    for (uint32_t b = 0; b < nb; b++) {
        VFPS_CHECK_ARG((size_t)buckets[b] * spacing_bins + n <= nmax,
                       "vfps_wake_potential_noalloc: bunch does not fit the "
                       "padded array");
    }
    const uint32_t M = nmax / 2;
    VFPS_CHECK_ARG(field_lds_bytes(M) <= dev->lds_max,
                   "vfps_wake_potential_noalloc: FFT length exceeds the LDS "
                   "of one CU");
    VFPS_TRY_BEGIN
    unsigned char* base = nullptr;
    int rc = noalloc_scratch(dev, &base);
    if (rc != VFPS_OK) return rc;
    if ((rc = field_set_lds_limit(field_lds_bytes(M))) != VFPS_OK) return rc;
    double2* tw_m = reinterpret_cast<double2*>(base);
    double2* tw_n = reinterpret_cast<double2*>(base + NOALLOC_TW_BYTES);
    uint32_t* bk_d = reinterpret_cast<uint32_t*>(base + 2 * NOALLOC_TW_BYTES);
    NoAllocBuckets bk{};
    for (uint32_t b = 0; b < nb; b++) bk.b[b] = buckets[b];
    hipStream_t st = dev->stream;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(ceil_div(M, 256u)),
                       dim3(256), 0, st, M, tw_m, tw_n, nb, bk, bk_d);
    VFPS_HIP(hipGetLastError());
    FieldParams p{};
    p.nb = nb;
    p.n = n;
    p.M = M;
    p.spacing = spacing_bins;
    p.buckets = bk_d;
    // _wakescaling(wakescaling/_nmax) (src/PS/ElectricField.cpp:58)
    p.scaling = wakescaling / (float)nmax;
    p.proj = proj;
    p.Z = reinterpret_cast<const float2*>(z);
    p.tw_m = tw_m;
    p.tw_n = tw_n;
    p.wake = wake;
    return launch_field(p, true, false, 1, st);
    VFPS_TRY_END
}

float vfps_field_wakescaling(double Ib, double dt, double qscale,
                             float delta1, double sigma_delta, double E0)
{
    // ps->getScale(0,"Meter") is the float stored in the Ruler's scale map
    return (float)(Ib * dt * 2.99792458e8 / (double)(float)qscale
                   / ((double)delta1 * sigma_delta * E0));
}

int vfps_field_create(vfps_mesh* ps, vfps_impedance* z,
                      const uint32_t* buckets, uint32_t nbuckets,
                      uint32_t spacing_bins, double f_rev,
                      float revolutionpart, float wakescaling,
                      vfps_field** out)
{
    VFPS_CHECK_ARG(ps && z && out && buckets, "ElectricField: NULL argument");
    VFPS_CHECK_ARG(nbuckets == ps->nb,
                   "ElectricField: one bucket number per bunch expected");
    const uint32_t nmax = z->n;
    VFPS_CHECK_ARG(nmax >= 4 && nmax <= 8192 && nmax % 2 == 0,
                   "ElectricField: FFT length (impedance size) must be even, "
                   "in [4, 8192]");
    for (uint32_t b = 0; b < nbuckets; b++) {
        VFPS_CHECK_ARG((size_t)buckets[b] * spacing_bins + ps->n <= nmax,
                       "ElectricField: bunch does not fit the padded array");
    }
    const uint32_t M = nmax / 2;
    VFPS_CHECK_ARG(field_lds_bytes(M) <= ps->dev->lds_max,
                   "ElectricField: FFT length exceeds the LDS of one CU");
    VFPS_TRY_BEGIN
    int rc = field_set_lds_limit(field_lds_bytes(M));
    if (rc != VFPS_OK) return rc;
    auto* f = new vfps_field();
    f->dev = ps->dev;
    f->ps = ps;
    f->z = z;
    f->nmax = nmax;
    f->nb = ps->nb;
    f->spacing = spacing_bins;
    f->buckets.assign(buckets, buckets + nbuckets);
    // _wakescaling(wakescaling/_nmax) (src/PS/ElectricField.cpp:58)
    f->wakescaling = wakescaling / (float)nmax;
    f->f_rev = f_rev;
    f->revolutionpart = revolutionpart;
    std::vector<double2> twm, twn;
    make_twiddles(M, twm, twn);
    auto renorm = csr_renorm(nmax, ps->ax0.delta, ps->qscale, 0.0f,
                             &f->fdelta);
    f->csr_fc = 0.0f;
    hipError_t e = hipSuccess;
    const uint32_t n = ps->n, nb = ps->nb;
    e = e ? e : hipMalloc(&f->buckets_d, sizeof(uint32_t) * nb);
    e = e ? e : hipMalloc(&f->wake, sizeof(float) * nb * n);
    e = e ? e : hipMalloc(&f->padded, sizeof(float) * nmax);
    e = e ? e : hipMalloc(&f->wake_padded, sizeof(float) * nmax);
    e = e ? e : hipMalloc(&f->tw_m, sizeof(double2) * M);
    e = e ? e : hipMalloc(&f->tw_n, sizeof(double2) * M);
    e = e ? e : hipMalloc(&f->csr_renorm, sizeof(float) * nmax);
    e = e ? e : hipMalloc(&f->spectrum, sizeof(float) * nb * nmax);
    e = e ? e : hipMalloc(&f->power, sizeof(float) * nb);
    if (e != hipSuccess) {
        vfps_field_destroy(f);
        return fail(VFPS_ERR_NOMEM, hipGetErrorString(e));
    }
    hipStream_t st = ps->dev->stream;
    hipMemcpyAsync(f->buckets_d, buckets, sizeof(uint32_t) * nb,
                   hipMemcpyHostToDevice, st);
    hipMemsetAsync(f->wake, 0, sizeof(float) * nb * n, st);
    hipMemsetAsync(f->padded, 0, sizeof(float) * nmax, st);
    hipMemsetAsync(f->wake_padded, 0, sizeof(float) * nmax, st);
    hipMemcpyAsync(f->tw_m, twm.data(), sizeof(double2) * M,
                   hipMemcpyHostToDevice, st);
    hipMemcpyAsync(f->tw_n, twn.data(), sizeof(double2) * M,
                   hipMemcpyHostToDevice, st);
    hipMemcpyAsync(f->csr_renorm, renorm.data(), sizeof(float) * nmax,
                   hipMemcpyHostToDevice, st);
    hipMemsetAsync(f->spectrum, 0, sizeof(float) * nb * nmax, st);
    hipMemsetAsync(f->power, 0, sizeof(float) * nb, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        vfps_field_destroy(f);
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    *out = f;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_field_destroy(vfps_field* f)
{
    if (!f) return VFPS_OK;
    hipStreamSynchronize(f->dev->stream);
    hipFree(f->buckets_d);
    hipFree(f->wake);
    hipFree(f->padded);
    hipFree(f->wake_padded);
    hipFree(f->tw_m);
    hipFree(f->tw_n);
    hipFree(f->csr_renorm);
    hipFree(f->spectrum);
    hipFree(f->power);
    delete f;
    return VFPS_OK;
}

__global__ void pad_kernel(uint32_t nb, uint32_t n, uint32_t nmax,
                           uint32_t spacing, const uint32_t* buckets,
                           const float* proj, float* padded)
{
    for (uint32_t i = threadIdx.x; i < nmax; i += blockDim.x) {
        padded[i] = 0.0f;
    }
    __syncthreads();
    for (uint32_t b = 0; b < nb; b++) {
        for (uint32_t x = threadIdx.x; x < n; x += blockDim.x) {
            padded[buckets[b] * spacing + x] = proj[(size_t)b * n + x];
        }
    }
}

int vfps_field_pad(vfps_field* f)
{
    VFPS_CHECK_ARG(f, "padBunchProfiles: NULL field");
    hipLaunchKernelGGL(pad_kernel, dim3(1), dim3(256), 0, f->dev->stream,
                       f->nb, f->ps->n, f->nmax, f->spacing, f->buckets_d,
                       f->ps->proj, f->padded);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_field_get_padded(vfps_field* f, float* out)
{
    VFPS_CHECK_ARG(f && out, "getPaddedBunchProfiles: NULL argument");
    VFPS_HIP(hipMemcpyAsync(out, f->padded, sizeof(float) * f->nmax,
                            hipMemcpyDeviceToHost, f->dev->stream));
    VFPS_HIP(hipStreamSynchronize(f->dev->stream));
    return VFPS_OK;
}

int vfps_field_get_padded_wake(vfps_field* f, float* out)
{
    VFPS_CHECK_ARG(f && out, "getPaddedWakePotential: NULL argument");
    VFPS_HIP(hipMemcpyAsync(out, f->wake_padded, sizeof(float) * f->nmax,
                            hipMemcpyDeviceToHost, f->dev->stream));
    VFPS_HIP(hipStreamSynchronize(f->dev->stream));
    return VFPS_OK;
}

int vfps_field_wake_potential(vfps_field* f, float* wp)
{
    VFPS_CHECK_ARG(f, "wakePotential: NULL field");
    int rc = launch_wake_potential(f, nullptr);
    if (rc != VFPS_OK) return rc;
    if (wp) {
        VFPS_HIP(hipMemcpyAsync(wp, f->wake, sizeof(float) * f->nb * f->ps->n,
                                hipMemcpyDeviceToHost, f->dev->stream));
        VFPS_HIP(hipStreamSynchronize(f->dev->stream));
    }
    return VFPS_OK;
}

int vfps_field_get_wakescaling(const vfps_field* f, float* s)
{
    VFPS_CHECK_ARG(f && s, "getWakeScaling: NULL argument");
    *s = f->wakescaling;
    return VFPS_OK;
}

int vfps_field_update_csr(vfps_field* f, float fc, float* power,
                          float* spectrum)
{
    VFPS_CHECK_ARG(f, "updateCSR: NULL field");
    VFPS_TRY_BEGIN
    if (fc != f->csr_fc) {
        auto renorm = csr_renorm(f->nmax, f->ps->ax0.delta, f->ps->qscale, fc,
                                 &f->fdelta);
        VFPS_HIP(hipMemcpyAsync(f->csr_renorm, renorm.data(),
                                sizeof(float) * f->nmax,
                                hipMemcpyHostToDevice, f->dev->stream));
        VFPS_HIP(hipStreamSynchronize(f->dev->stream));
        f->csr_fc = fc;
    }
    FieldParams p = field_params(f);
    p.spectrum = f->spectrum;
    p.power = f->power;
    int rc = launch_field(p, false, true, f->nb, f->dev->stream);
    if (rc != VFPS_OK) return rc;
    if (power) {
        VFPS_HIP(hipMemcpyAsync(power, f->power, sizeof(float) * f->nb,
                                hipMemcpyDeviceToHost, f->dev->stream));
    }
    if (spectrum) {
        VFPS_HIP(hipMemcpyAsync(spectrum, f->spectrum,
                                sizeof(float) * f->nb * f->nmax,
                                hipMemcpyDeviceToHost, f->dev->stream));
    }
    if (power || spectrum) {
        VFPS_HIP(hipStreamSynchronize(f->dev->stream));
    }
    return VFPS_OK;
    VFPS_TRY_END
}

}  // extern "C"
