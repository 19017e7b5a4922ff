// Impedance tables (inc/Z/*.hpp, src/Z/*.cpp), built on the device in FP64
// and rounded to complex<float>, then added to the table like
// Impedance::operator+= (src/Z/Impedance.cpp:66-76).
//
// Airy functions for x >= 0 (ParallelPlatesCSR):
//  - x < AIRY_SPLIT: Taylor expansion about the nearest node of a table of
//    (Ai, Ai', Bi, Bi') at x_k = k h.  The table itself is generated on the
//    device by Taylor stepping of y'' = x y: Bi forward from its values at
//    0 (DLMF 9.2.4, 9.2.6), Ai backward from AIRY_SPLIT where the
//    asymptotic series is exact to double precision; both directions keep
//    the stepped solution dominant, so the stepping is stable.
//    Derivatives follow y^(n+2) = x y^(n) + n y^(n-1).
//  - x >= AIRY_SPLIT: the asymptotic expansions DLMF 9.7.5-9.7.8 with the
//    coefficients u_k, v_k of DLMF 9.7.2.
#include <cmath>

#include "common.hpp"
#include "reduce.hpp"

namespace vfps_impl {

constexpr double AIRY_H = 0.25;
constexpr double AIRY_SPLIT = 12.0;
constexpr int AIRY_NODES = 49;  // x_k = k h, k = 0..48
constexpr double PI_D = 3.141592653589793238462643383279502884;
constexpr double LN_DBL_MAX = 709.782712893383973096;

// Taylor step of y'' = x y from x0 by d: returns y(x0+d), y'(x0+d)
__device__ inline void airy_taylor(double x0, double y, double yp, double d,
                                   double* y1, double* yp1)
{
    double am1 = 0.0;      // a_{n-1}
    double a0 = y;         // a_n   (a_k = y^(k)(x0) d^k / k!)
    double a1 = yp * d;    // a_{n+1}
    double sy = a0 + a1;
    double sp = a1;        // sum of k a_k (k = 1)
    const double d2 = d * d, d3 = d2 * d;
    for (int n = 0; n < 60; n++) {
        // a_{n+2} = (x0 d^2 a_n + d^3 a_{n-1}) / ((n+1)(n+2))
        double a2 = (x0 * d2 * a0 + d3 * am1) / ((double)(n + 1) * (double)(n + 2));
        sy += a2;
        sp += (double)(n + 2) * a2;
        am1 = a0;
        a0 = a1;
        a1 = a2;
        if (n > 4 && fabs(a2) <= 1e-18 * fabs(sy) && fabs(a0) <= 1e-18 * fabs(sy)) {
            break;
        }
    }
    *y1 = sy;
    *yp1 = (d != 0.0) ? sp / d : yp;
}

// asymptotic expansions for x >= AIRY_SPLIT; returns 1 if Bi or Bi'
// exceed the double range
__device__ inline int airy_asym(double x, double* ai, double* aip, double* bi,
                                double* bip)
{
    const double z = 2.0 / 3.0 * x * sqrt(x);
    const double x14 = sqrt(sqrt(x));
    double uk = 1.0, su_p = 1.0, su_m = 1.0, sv_p = 1.0, sv_m = 1.0;
    double zk = 1.0;
    for (int k = 1; k < 80; k++) {
        uk = uk * (double)((6 * k - 5) * (6 * k - 3) * (6 * k - 1))
             / ((double)(2 * k - 1) * 216.0 * (double)k);
        const double vk = -(double)(6 * k + 1) / (double)(6 * k - 1) * uk;
        zk = zk / z;
        const double tu = uk * zk, tv = vk * zk;
        const double sg = (k & 1) ? -1.0 : 1.0;
        su_p += tu;
        su_m += sg * tu;
        sv_p += tv;
        sv_m += sg * tv;
        if (fabs(tu) < 1e-18 && fabs(tv) < 1e-18) {
            break;
        }
        if ((double)k > 2.0 * z) {
            break;
        }
    }
    const double rpi = 1.0 / sqrt(PI_D);
    const double em = exp(-z);
    *ai = 0.5 * rpi / x14 * em * su_m;
    *aip = -0.5 * rpi * x14 * em * sv_m;
    // Bi' = x^{1/4} e^z / sqrt(pi) * sv_p: test against the double range in
    // logarithms before forming it
    const double lbip = z + log(x14 * rpi * sv_p);
    const double lbi = z + log(rpi / x14 * su_p);
    if (lbip > LN_DBL_MAX || lbi > LN_DBL_MAX) {
        *bi = INFINITY;
        *bip = INFINITY;
        return 1;
    }
    *bi = exp(lbi);
    *bip = exp(lbip);
    return 0;
}

// node table: [AIRY_NODES][4] = Ai, Ai', Bi, Bi'
__global__ void airy_nodes_kernel(double* tab)
{
    if (threadIdx.x == 0) {
        // Bi(0) = 3^{-1/6}/Gamma(2/3), Bi'(0) = 3^{1/6}/Gamma(1/3)
        double y = 0.614926627446000735150922369093613553594728188648596;
        double yp = 0.448288357353826357914823710398828390866226799212262;
        tab[2] = y;
        tab[3] = yp;
        for (int k = 1; k < AIRY_NODES; k++) {
            airy_taylor((k - 1) * AIRY_H, y, yp, AIRY_H, &y, &yp);
            tab[4 * k + 2] = y;
            tab[4 * k + 3] = yp;
        }
    } else if (threadIdx.x == 1) {
        double ai, aip, bi, bip;
        const int last = AIRY_NODES - 1;
        airy_asym(last * AIRY_H, &ai, &aip, &bi, &bip);
        tab[4 * last + 0] = ai;
        tab[4 * last + 1] = aip;
        for (int k = last - 1; k >= 0; k--) {
            airy_taylor((k + 1) * AIRY_H, ai, aip, -AIRY_H, &ai, &aip);
            tab[4 * k + 0] = ai;
            tab[4 * k + 1] = aip;
        }
    }
}

__device__ inline int airy_eval(double x, const double* __restrict__ tab,
                                double* ai, double* aip, double* bi,
                                double* bip)
{
    if (x >= AIRY_SPLIT) {
        return airy_asym(x, ai, aip, bi, bip);
    }
    int k = (int)floor(x / AIRY_H + 0.5);
    k = k < 0 ? 0 : (k > AIRY_NODES - 1 ? AIRY_NODES - 1 : k);
    const double x0 = k * AIRY_H;
    const double d = x - x0;
    airy_taylor(x0, tab[4 * k + 0], tab[4 * k + 1], d, ai, aip);
    airy_taylor(x0, tab[4 * k + 2], tab[4 * k + 3], d, bi, bip);
    return 0;
}

__global__ void airy_kernel(uint32_t n, const double* x, const double* tab,
                            double* out, int32_t* ovf)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    double a, ap, b, bp;
    ovf[i] = airy_eval(x[i], tab, &a, &ap, &b, &bp);
    out[4 * i + 0] = a;
    out[4 * i + 1] = ap;
    out[4 * i + 2] = b;
    out[4 * i + 3] = bp;
}

// ParallelPlatesCSR (src/Z/ParallelPlatesCSR.cpp:26-69): one workgroup per
// frequency index i = 1..nfreqs/2; odd p = 1, 3, ... <= maxp spread over
// the threads; terms from the first p where Bi or Bi' overflow on are
// dropped (the reference's Boost overflow exception ends its p loop there;
// Bi, Bi' increase with u, so that is exactly the set dropped).
__global__ void __launch_bounds__(256)
pp_csr_kernel(uint32_t nfreqs, double delta, double r_bend, double f0,
              double g, const double* __restrict__ tab, float2* z,
              unsigned long long* nterms)
{
    __shared__ double red[16];
    const uint32_t i = blockIdx.x + 1;
    const double pi2 = PI_D * PI_D;
    const double z0 = 1.0 / (8.854187817e-12 * 2.99792458e8);  // physcons::Z0
    const double n = i * delta;
    const double m = n * pow(g / r_bend, 3. / 2.);
    const uint32_t maxp = (uint32_t)(2 * m * pow(r_bend / g, 3. / 2.) * f0 * g
                                     / 2.99792458e8);
    const double b = pow(m, -4. / 3.);
    const double p23 = pow(2.0, 2. / 3.);
    const uint32_t nodd = (maxp + 1) / 2;  // p = 2t+1 <= maxp
    double zr = 0.0, zi = 0.0;
    unsigned long long cnt = 0;
    for (uint32_t t = threadIdx.x; t < nodd; t += blockDim.x) {
        const uint32_t p = 2 * t + 1;
        const double u = pi2 * p * p / p23 * b;
        if (u > 200.0) {
            continue;  // far beyond the overflow of Bi'
        }
        double ai, aip, bi, bip;
        if (airy_eval(u, tab, &ai, &aip, &bi, &bip)) {
            continue;
        }
        const double uai = u * ai;
        zr += aip * aip + uai * ai;
        zi += aip * (-bip) + uai * (-bi);
        cnt++;
    }
    zr = block_sum(zr, red);
    zi = block_sum(zi, red);
    double c = (double)cnt;
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        const double sc = 4.0 * b * n * g / r_bend * pi2 * pow(2.0, 1. / 3.) * z0;
        const float2 old = z[i];
        z[i] = make_float2(old.x + (float)(zr * sc), old.y + (float)(zi * sc));
        atomicAdd(nterms, (unsigned long long)c);
    }
}

// FreeSpaceCSR (src/Z/FreeSpaceCSR.cpp:19-42) in FP64:
// Z_i = (306.3 + 176.9 j) (i delta)^{1/3}, i <= n/2
__global__ void fs_csr_kernel(uint32_t n, double delta, float2* z)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n / 2) {
        return;
    }
    const double p = cbrt((double)i * delta);
    const float2 old = z[i];
    z[i] = make_float2(old.x + (float)(306.3 * p), old.y + (float)(176.9 * p));
}

// ResistiveWall (src/Z/ResistiveWall.cpp:25-55) in FP64:
// Z_i = Z1 (1 - j) sqrt(i delta), i <= n/2
__global__ void rw_kernel(uint32_t n, double z1, double delta, float2* z)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n / 2) {
        return;
    }
    const double r = z1 * sqrt((double)i * delta);
    const float2 old = z[i];
    z[i] = make_float2(old.x + (float)r, old.y + (float)(-r));
}

// ConstImpedance (src/Z/ConstImpedance.cpp:19-27): Z for i < n/2
__global__ void const_kernel(uint32_t n, float re, float im, float2* z)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n / 2) {
        return;
    }
    const float2 old = z[i];
    z[i] = make_float2(old.x + re, old.y + im);
}

__global__ void add_kernel(uint32_t n, const float2* a, float2* z)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    z[i] = make_float2(z[i].x + a[i].x, z[i].y + a[i].y);
}

static int airy_table(hipStream_t st, double** tab)
{
    VFPS_HIP(hipMalloc(tab, sizeof(double) * 4 * AIRY_NODES));
    hipLaunchKernelGGL(airy_nodes_kernel, dim3(1), dim3(64), 0, st, *tab);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

}  // namespace vfps_impl

using namespace vfps_impl;

extern "C" {

int vfps_impedance_create(vfps_dev* dev, uint32_t nfreqs, float f_max,
                          vfps_impedance** out)
{
    VFPS_CHECK_ARG(dev && out && nfreqs >= 2, "Impedance: bad argument");
    VFPS_TRY_BEGIN
    auto* z = new vfps_impedance();
    z->dev = dev;
    z->n = nfreqs;
    z->fmax = f_max;
    hipError_t e = hipMalloc(&z->z, sizeof(float2) * nfreqs);
    if (e != hipSuccess) {
        delete z;
        return fail(VFPS_ERR_NOMEM, hipGetErrorString(e));
    }
    hipMemsetAsync(z->z, 0, sizeof(float2) * nfreqs, dev->stream);
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    *out = z;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_impedance_create_table(vfps_dev* dev, const vfps_cplx* zt,
                                uint32_t nfreqs, float f_max,
                                vfps_impedance** out)
{
    VFPS_CHECK_ARG(zt, "Impedance: NULL table");
    int rc = vfps_impedance_create(dev, nfreqs, f_max, out);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipMemcpyAsync((*out)->z, zt, sizeof(float2) * nfreqs,
                            hipMemcpyHostToDevice, dev->stream));
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    return VFPS_OK;
}

int vfps_impedance_destroy(vfps_impedance* z)
{
    if (!z) return VFPS_OK;
    hipStreamSynchronize(z->dev->stream);
    hipFree(z->z);
    delete z;
    return VFPS_OK;
}

int vfps_impedance_add_freespace_csr(vfps_impedance* z, float f_rev,
                                     float f_max)
{
    VFPS_CHECK_ARG(z && f_rev > 0, "FreeSpaceCSR: bad argument");
    const double delta = (double)f_max / (double)f_rev / (double)(z->n - 1);
    hipLaunchKernelGGL(fs_csr_kernel, dim3(ceil_div(z->n / 2 + 1, 256)),
                       dim3(256), 0, z->dev->stream, z->n, delta, z->z);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_impedance_add_parallel_plates(vfps_impedance* z, float f0,
                                       float f_max, double g)
{
    VFPS_CHECK_ARG(z && f0 > 0 && g > 0, "ParallelPlatesCSR: bad argument");
    double* tab = nullptr;
    unsigned long long* cnt = nullptr;
    hipStream_t st = z->dev->stream;
    int rc = airy_table(st, &tab);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipMalloc(&cnt, sizeof(unsigned long long)));
    hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st);
    // delta = f_max/f0/(nfreqs-1.0): float quotient, then double
    const double delta = (double)(f_max / f0) / ((double)z->n - 1.0);
    const double r_bend = 2.99792458e8 / (2 * PI_D * (double)f0);
    if (z->n / 2 >= 1) {
        hipLaunchKernelGGL(pp_csr_kernel, dim3(z->n / 2), dim3(256), 0, st,
                           z->n, delta, r_bend, (double)f0, g, tab, z->z, cnt);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(tab);
    hipFree(cnt);
    if (e != hipSuccess) {
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    return VFPS_OK;
}

int vfps_impedance_add_resistive_wall(vfps_impedance* z, float f0,
                                      float f_max, double L, double s,
                                      double xi, double b)
{
    VFPS_CHECK_ARG(z && f0 > 0 && s > 0 && b > 0,
                   "ResistiveWall: bad argument");
    const double mu_r = 1 + xi;
    const double z1 = std::sqrt(376.730313461 * mu_r * (double)f0 / s / PI_D
                                / 2.99792458e8) * L / 2 / b;
    const double delta = (double)f_max / (double)f0 / ((double)z->n - 1.0);
    hipLaunchKernelGGL(rw_kernel, dim3(ceil_div(z->n / 2 + 1, 256)), dim3(256),
                       0, z->dev->stream, z->n, z1, delta, z->z);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_impedance_add_const(vfps_impedance* z, float re, float im)
{
    VFPS_CHECK_ARG(z, "ConstImpedance: NULL table");
    hipLaunchKernelGGL(const_kernel, dim3(ceil_div(z->n, 256)), dim3(256), 0,
                       z->dev->stream, z->n, re, im, z->z);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_impedance_add_collimator(vfps_impedance* z, double outer,
                                  double inner)
{
    VFPS_CHECK_ARG(z && outer > 0 && inner > 0,
                   "CollimatorImpedance: bad argument");
    // Z0/pi*log(outer/inner) with Impedance::Z0
    const float re = (float)(376.730313461 / PI_D * std::log(outer / inner));
    return vfps_impedance_add_const(z, re, 0.0f);
}

int vfps_impedance_add(vfps_impedance* z, const vfps_impedance* rhs)
{
    VFPS_CHECK_ARG(z && rhs && z->n == rhs->n, "Impedance +=: size mismatch");
    hipLaunchKernelGGL(add_kernel, dim3(ceil_div(z->n, 256)), dim3(256), 0,
                       z->dev->stream, z->n, rhs->z, z->z);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_impedance_download(vfps_impedance* z, vfps_cplx* out)
{
    VFPS_CHECK_ARG(z && out, "Impedance: NULL argument");
    VFPS_HIP(hipMemcpyAsync(out, z->z, sizeof(float2) * z->n,
                            hipMemcpyDeviceToHost, z->dev->stream));
    VFPS_HIP(hipStreamSynchronize(z->dev->stream));
    return VFPS_OK;
}

int vfps_impedance_nfreqs(const vfps_impedance* z, uint32_t* n)
{
    VFPS_CHECK_ARG(z && n, "Impedance: NULL argument");
    *n = z->n;
    return VFPS_OK;
}

int vfps_airy(vfps_dev* dev, const double* x, uint32_t n, double* out,
              int32_t* ovf)
{
    VFPS_CHECK_ARG(dev && x && out && ovf && n > 0, "vfps_airy: bad argument");
    hipStream_t st = dev->stream;
    double* tab = nullptr;
    double* dx = nullptr;
    double* dout = nullptr;
    int32_t* dovf = nullptr;
    int rc = airy_table(st, &tab);
    if (rc != VFPS_OK) return rc;
    hipError_t e = hipMalloc(&dx, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * 4 * n);
    if (e == hipSuccess) e = hipMalloc(&dovf, sizeof(int32_t) * n);
    if (e == hipSuccess) {
        hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, st);
        hipLaunchKernelGGL(airy_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                           st, n, dx, tab, dout, dovf);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipMemcpyAsync(out, dout, sizeof(double) * 4 * n,
                       hipMemcpyDeviceToHost, st);
        hipMemcpyAsync(ovf, dovf, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                       st);
        e = hipStreamSynchronize(st);
    }
    hipFree(tab);
    hipFree(dx);
    hipFree(dout);
    hipFree(dovf);
    if (e != hipSuccess) {
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    return VFPS_OK;
}

}  // extern "C"
