// Device handle, errors and the PhaseSpace object (inc/PS/PhaseSpace.hpp).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

#include "common.hpp"
#include "reduce.hpp"

namespace vfps_impl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

// inc/PS/Ruler.hpp:31-51
Ruler::Ruler(uint32_t s, float mn, float mx) : steps(s), min(mn), max(mx)
{
    if (mn >= mx) {
        throw std::invalid_argument("Tried to set up Ruler with min >= max.");
    }
    if (s < 2) {
        throw std::invalid_argument("Ruler needs at least two points.");
    }
    delta = (max - min) / (float)(steps - 1);
    float zb = ((min + max) / (min - max) + 1.0f) * (float)(steps - 1);
    zerobin = zb / 2.0f;
    data.resize(steps);
    for (uint32_t i = 0; i < steps; i++) {
        data[i] = min + ((float)i * delta);
    }
}

// src/PS/PhaseSpace.cpp:486-501
std::vector<float> simpson_weights(uint32_t n, float delta0)
{
    std::vector<float> rv(n);
    const float ca = 3.0f;
    float dc = 1.0f;
    const float h03 = delta0 / 3.0f;
    rv[0] = h03;
    for (uint32_t x = 1; x < n - 1; x++) {
        rv[x] = h03 * (ca + dc);
        dc = -dc;
    }
    rv[n - 1] = h03;
    return rv;
}

// ---- kernels -----------------------------------------------------------

// gaus (src/PS/PhaseSpace.cpp:471-484): FP64 exponential rounded to float;
// g[0..n) for axis 0, g[n..2n) for axis 1
__global__ void gaus_kernel(uint32_t n, float qmin, float dq, float pmin,
                            float dp, double zoom2, float* g)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n) {
        return;
    }
    const float c = (float)0.398942280401432677939946059934381868475858631164934657665925829670657925899L;
    uint32_t k = (i < n) ? i : i - n;
    float a = (i < n) ? (qmin + ((float)k * dq)) : (pmin + ((float)k * dp));
    double ad = (double)a;
    double e = exp(((-0.5) * ad) * ad / zoom2);
    g[i] = (float)((double)c * e);
}

// createFromProjections (src/PS/PhaseSpace.cpp:457-469): outer product
__global__ void outer_kernel(uint32_t nb, uint32_t n, const float* g,
                             float* data)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nxy = (size_t)n * n;
    if (i >= nxy * nb) {
        return;
    }
    size_t c = i % nxy;
    uint32_t x = (uint32_t)(c / n), y = (uint32_t)(c % n);
    data[i] = g[x] * g[n + y];
}

// updateXProjection (src/PS/PhaseSpace.cpp:291-323): std::inner_product
// order (acc = acc + data*ws for y = 0..n-1, a float accumulator), so the
// projection equals the reference's bit for bit (the wake potential
// amplifies the last bits of the profile through the high-frequency
// impedance).  A workgroup takes PX_ROWS rows: all its threads form the
// products data*ws with coalesced loads into LDS (rows padded by 4 words),
// then one lane per row adds its row in order from LDS.
constexpr uint32_t PX_ROWS = 16;
constexpr uint32_t PX_THREADS = 256;
constexpr int PX_BATCH = 8;   // float4 loads in flight per thread

__global__ void __launch_bounds__(PX_THREADS)
project_x_kernel(uint32_t rows, uint32_t n, uint32_t rows_per_wg,
                 const float* __restrict__ data,
                 const float* __restrict__ ws, float* __restrict__ proj)
{
    extern __shared__ __align__(16) float px[];   // rows_per_wg * (n + 4)
    const uint32_t r0 = blockIdx.x * rows_per_wg;
    const uint32_t nr = min(rows_per_wg, rows - r0);
    const uint32_t stride = n + 4;
    const float* src = data + (size_t)r0 * n;
    if ((n & 3u) == 0u) {
        // PX_BATCH float4s of the rows (and their weights) in flight per
        // thread before any is stored: a load that waits for its own data
        // would wait for every load before it too
        const uint32_t n4 = n >> 2;
        const float4* s4 = reinterpret_cast<const float4*>(src);
        const float4* w4 = reinterpret_cast<const float4*>(ws);
        const uint32_t dr = PX_THREADS / n4, dg = PX_THREADS % n4;
        uint32_t r = threadIdx.x / n4, g = threadIdx.x % n4;
        while (r < nr) {
            float4 v[PX_BATCH], w[PX_BATCH];
            uint32_t rb[PX_BATCH], gb[PX_BATCH];
#pragma unroll
            for (int b = 0; b < PX_BATCH; b++) {
                rb[b] = r;
                gb[b] = g;
                // (past the last row: row 0 again, not stored)
                const uint32_t rs = r < nr ? r : 0u;
                v[b] = s4[(size_t)rs * n4 + g];
                w[b] = w4[g];
                g += dg;
                r += dr;
                if (g >= n4) {
                    g -= n4;
                    r++;
                }
            }
#pragma unroll
            for (int b = 0; b < PX_BATCH; b++) {
                if (rb[b] < nr) {
                    float4 o;
                    o.x = v[b].x * w[b].x;
                    o.y = v[b].y * w[b].y;
                    o.z = v[b].z * w[b].z;
                    o.w = v[b].w * w[b].w;
                    *reinterpret_cast<float4*>(px + rb[b] * stride + 4 * gb[b]) = o;
                }
            }
        }
    } else {
        for (uint32_t i = threadIdx.x; i < nr * n; i += blockDim.x) {
            const uint32_t r = i / n, y = i - r * n;
            px[r * stride + y] = src[i] * ws[y];
        }
    }
    __syncthreads();
    if (threadIdx.x < nr) {
        const float* p = px + threadIdx.x * stride;
        float acc = 0.0f;
        uint32_t y = 0;
        if ((n & 15u) == 0u) {
            // the next 16 products load while the current 16 are added
            const float4* p4 = reinterpret_cast<const float4*>(p);
            float4 q[4], nq[4];
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = p4[k];
            for (; y < n; y += 16) {
                const uint32_t yn = y + 16 < n ? y + 16 : y;
#pragma unroll
                for (int k = 0; k < 4; k++) nq[k] = p4[(yn >> 2) + k];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    acc = acc + q[k].x;
                    acc = acc + q[k].y;
                    acc = acc + q[k].z;
                    acc = acc + q[k].w;
                }
#pragma unroll
                for (int k = 0; k < 4; k++) q[k] = nq[k];
            }
        }
        for (; y < n; y++) {
            acc = acc + p[y];
        }
        proj[r0 + threadIdx.x] = acc;
    }
}

// updateYProjection (src/PS/PhaseSpace.cpp:325-340): for each column y,
// proj[y] = proj[y] + data[x][y]*ws[x] for x = 0..n-1 in order.  A
// workgroup takes 64 columns of one bunch, lane y of wave 0 their sums;
// the other waves stage chunks of PY_ROWS rows (the products) into a
// double buffer in LDS, one chunk ahead of the sums.
constexpr uint32_t PY_COLS = 64;
constexpr uint32_t PY_ROWS = 64;
constexpr uint32_t PY_THREADS = 256;

__global__ void __launch_bounds__(PY_THREADS)
project_y_kernel(uint32_t nb, uint32_t n, const float* __restrict__ data,
                 const float* __restrict__ ws, float* __restrict__ proj)
{
    __shared__ float buf[2][PY_ROWS][PY_COLS + 1];
    const uint32_t b = blockIdx.y;
    const uint32_t c0 = blockIdx.x * PY_COLS;
    const uint32_t ncol = min(PY_COLS, n - c0);
    const float* d = data + (size_t)b * n * n;
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const uint32_t nchunks = (n + PY_ROWS - 1) / PY_ROWS;
    // stage chunk k by the threads [t0, t0 + nt)
    auto stage = [&](uint32_t k, uint32_t t0, uint32_t nt) {
        const uint32_t x0 = k * PY_ROWS;
        const uint32_t nr = min(PY_ROWS, n - x0);
        float(*dst)[PY_COLS + 1] = buf[k & 1u];
        for (uint32_t i = tid - t0; i < nr * PY_COLS; i += nt) {
            const uint32_t r = i / PY_COLS, c = i % PY_COLS;
            if (c < ncol) {
                dst[r][c] = d[(size_t)(x0 + r) * n + c0 + c] * ws[x0 + r];
            }
        }
    };
    stage(0, 0, blockDim.x);
    __syncthreads();
    float acc = 0.0f;
    for (uint32_t k = 0; k < nchunks; k++) {
        if (wave != 0) {
            if (k + 1 < nchunks) stage(k + 1, 64, blockDim.x - 64);
        } else if (tid < ncol) {
            const uint32_t nr = min(PY_ROWS, n - k * PY_ROWS);
            float(*src)[PY_COLS + 1] = buf[k & 1u];
            for (uint32_t r = 0; r < nr; r++) {
                acc = acc + src[r][tid];
            }
        }
        __syncthreads();
    }
    if (tid < ncol) {
        proj[(size_t)b * n + c0 + tid] = acc;
    }
}

// integrate (src/PS/PhaseSpace.cpp:200-239), FP64 accumulation
__global__ void integrate_kernel(uint32_t nb, uint32_t n, const float* proj,
                                 const float* ws, double* filling,
                                 double* integral)
{
    __shared__ double part[16];
    double total = 0.0;
    for (uint32_t b = 0; b < nb; b++) {
        double acc = 0.0;
        for (uint32_t x = threadIdx.x; x < n; x += blockDim.x) {
            acc += (double)proj[(size_t)b * n + x] * (double)ws[x];
        }
        acc = block_sum(acc, part);
        if (threadIdx.x == 0) {
            filling[b] = acc;
        }
        total += acc;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *integral = total;
    }
}

// normalize (src/PS/PhaseSpace.cpp:342-372)
__global__ void normalize_kernel(uint32_t nb, size_t nxy, float* data,
                                 const float* fill_set, const double* filling)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy * nb) {
        return;
    }
    uint32_t b = (uint32_t)(i / nxy);
    float fs = fill_set[b];
    if (fs > 0.0f) {
        float factor = (float)((double)fs / filling[b]);
        data[i] = data[i] * factor;
    } else {
        data[i] = 0.0f;
    }
}

// average + variance (src/PS/PhaseSpace.cpp:241-289), FP64 accumulation;
// one block per bunch
__global__ void variance_kernel(uint32_t n, const float* proj,
                                const float* axis, float delta,
                                const float* fill_set, const double* filling,
                                float* out)
{
    __shared__ double part[16];
    const uint32_t b = blockIdx.x;
    const float* p = proj + (size_t)b * n;
    const bool filled = fill_set[b] > 0.0f;
    double avg = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        avg += (double)p[i] * (double)axis[i];
    }
    avg = block_sum(avg, part);
    float favg = filled ? (float)(avg * ((double)delta / filling[b])) : 0.0f;
    __syncthreads();
    double var = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        double d = (double)(axis[i] - favg);
        var += (double)p[i] * d * d;
    }
    var = block_sum(var, part);
    if (threadIdx.x == 0) {
        float fvar = filled ? (float)(var * ((double)delta / filling[b])) : 0.0f;
        out[3 * b + 0] = favg;
        out[3 * b + 1] = fvar;
        out[3 * b + 2] = sqrtf(fvar);
    }
}

// ---- launch helpers used by other translation units ----------------------
int launch_project_x(vfps_mesh* m, const float* src, float* proj)
{
    const uint32_t rows = m->nb * m->n;
    // as many rows per workgroup as the LDS holds, up to PX_ROWS
    const size_t row_bytes = sizeof(float) * (m->n + 4);
    const uint32_t rpw = (uint32_t)std::max<size_t>(
        1, std::min<size_t>(PX_ROWS, m->dev->lds_max / row_bytes));
    const size_t lds = row_bytes * rpw;
    VFPS_CHECK_ARG(lds <= m->dev->lds_max,
                   "updateXProjection: a mesh row exceeds the LDS");
    if (lds > 65536) {
        VFPS_HIP(hipFuncSetAttribute(
            reinterpret_cast<const void*>(&project_x_kernel),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL(project_x_kernel, dim3(ceil_div(rows, rpw)),
                       dim3(PX_THREADS), lds, m->dev->stream, rows, m->n, rpw,
                       src, m->ws, proj);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int launch_integrate(vfps_mesh* m)
{
    hipLaunchKernelGGL(integrate_kernel, dim3(1), dim3(256), 0, m->dev->stream,
                       m->nb, m->n, m->proj, m->ws, m->filling, m->integral);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

}  // namespace vfps_impl

using namespace vfps_impl;

extern "C" {

const char* vfps_last_error(void) { return g_last_error.c_str(); }

const char* vfps_version(void) { return "vfps-mi355x 0.1 (gfx950)"; }

int vfps_dev_open(int ordinal, vfps_dev** out)
{
    VFPS_CHECK_ARG(out != nullptr, "vfps_dev_open: out is NULL");
    VFPS_TRY_BEGIN
    int count = 0;
    VFPS_HIP(hipGetDeviceCount(&count));
    VFPS_CHECK_ARG(ordinal >= 0 && ordinal < count,
                   "vfps_dev_open: no such device");
    VFPS_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    VFPS_HIP(hipGetDeviceProperties(&prop, ordinal));
    auto* d = new vfps_dev();
    d->ordinal = ordinal;
    d->ncu = prop.multiProcessorCount;
    d->lds_max = prop.sharedMemPerBlock;
    d->arch = prop.gcnArchName;
    if (d->arch.rfind("gfx950", 0) != 0) {
        delete d;
        return fail(VFPS_ERR_UNSUPPORTED,
                    "vfps_dev_open: library is built for gfx950, device is " +
                        std::string(prop.gcnArchName));
    }
    hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return fail(VFPS_ERR_HIP, std::string("hipStreamCreate: ") +
                                      hipGetErrorString(e));
    }
    *out = d;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_dev_close(vfps_dev* dev)
{
    if (!dev) {
        return VFPS_OK;
    }
    hipSetDevice(dev->ordinal);
    hipStreamSynchronize(dev->stream);
    hipStreamDestroy(dev->stream);
    delete dev;
    return VFPS_OK;
}

int vfps_sync(vfps_dev* dev)
{
    VFPS_CHECK_ARG(dev, "vfps_sync: NULL device");
    VFPS_HIP(hipStreamSynchronize(dev->stream));
    return VFPS_OK;
}

static void mesh_free(vfps_mesh* m)
{
    if (!m) {
        return;
    }
    hipFree(m->data);
    hipFree(m->proj);
    hipFree(m->ws);
    hipFree(m->fill_set_d);
    hipFree(m->filling);
    hipFree(m->integral);
    delete m;
}

static int mesh_alloc(vfps_dev* dev, uint32_t nb, uint32_t n, float qmin,
                      float qmax, double qscale, float pmin, float pmax,
                      double pscale, const float* filling, vfps_mesh** out)
{
    VFPS_CHECK_ARG(dev, "PhaseSpace: NULL device");
    VFPS_CHECK_ARG(nb >= 1 && n >= 3, "PhaseSpace: need nb >= 1 and n >= 3");
    // src/PS/PhaseSpace.cpp:47-51
    float s = 0.0f;
    for (uint32_t i = 0; i < nb; i++) {
        s = s + filling[i];
    }
    if (std::round(1e5 * (double)s) != 1e5) {
        return fail(VFPS_ERR_INVALID, "Argument \"filling\" not normalized.");
    }
    auto* m = new vfps_mesh();
    m->dev = dev;
    m->nb = nb;
    m->n = n;
    try {
        m->ax0 = Ruler(n, qmin, qmax);
        m->ax1 = Ruler(n, pmin, pmax);
    } catch (const std::exception& ex) {
        delete m;
        return fail(VFPS_ERR_INVALID, ex.what());
    }
    m->qscale = qscale;
    m->pscale = pscale;
    m->ws_h = simpson_weights(n, m->ax0.delta);
    m->filling_set.assign(filling, filling + nb);
    hipSetDevice(dev->ordinal);
    hipError_t e = hipSuccess;
    e = (e == hipSuccess) ? hipMalloc(&m->data, sizeof(float) * m->cells()) : e;
    e = (e == hipSuccess) ? hipMalloc(&m->proj, sizeof(float) * 2 * nb * n) : e;
    e = (e == hipSuccess) ? hipMalloc(&m->ws, sizeof(float) * n) : e;
    e = (e == hipSuccess) ? hipMalloc(&m->fill_set_d, sizeof(float) * nb) : e;
    e = (e == hipSuccess) ? hipMalloc(&m->filling, sizeof(double) * nb) : e;
    e = (e == hipSuccess) ? hipMalloc(&m->integral, sizeof(double)) : e;
    if (e != hipSuccess) {
        mesh_free(m);
        return fail(VFPS_ERR_NOMEM, std::string("PhaseSpace allocation: ") +
                                        hipGetErrorString(e));
    }
    hipMemsetAsync(m->proj, 0, sizeof(float) * 2 * nb * n, dev->stream);
    hipMemcpyAsync(m->ws, m->ws_h.data(), sizeof(float) * n,
                   hipMemcpyHostToDevice, dev->stream);
    hipMemcpyAsync(m->fill_set_d, filling, sizeof(float) * nb,
                   hipMemcpyHostToDevice, dev->stream);
    e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess) {
        mesh_free(m);
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    *out = m;
    return VFPS_OK;
}

int vfps_mesh_create(vfps_dev* dev, uint32_t nb, uint32_t n, float qmin,
                     float qmax, double qscale, float pmin, float pmax,
                     double pscale, double charge, double current,
                     const float* filling, double zoom, const float* data,
                     vfps_mesh** out)
{
    VFPS_CHECK_ARG(out && filling, "vfps_mesh_create: NULL argument");
    VFPS_TRY_BEGIN
    vfps_mesh* m = nullptr;
    int rc = mesh_alloc(dev, nb, n, qmin, qmax, qscale, pmin, pmax, pscale,
                        filling, &m);
    if (rc != VFPS_OK) {
        return rc;
    }
    m->charge = charge;
    m->current = current;
    hipStream_t st = dev->stream;
    if (data) {
        VFPS_HIP(hipMemcpyAsync(m->data, data, sizeof(float) * m->cells(),
                                hipMemcpyHostToDevice, st));
    } else {
        // gaus on both axes, outer product, then updateXProjection,
        // integrate, normalize (src/PS/PhaseSpace.cpp:56-63, 457-469)
        hipLaunchKernelGGL(gaus_kernel, dim3(ceil_div(2 * n, 256)), dim3(256),
                           0, st, n, m->ax0.min, m->ax0.delta, m->ax1.min,
                           m->ax1.delta, zoom * zoom, m->proj);
        hipLaunchKernelGGL(outer_kernel, dim3(ceil_div(m->cells(), 256)),
                           dim3(256), 0, st, nb, n, m->proj, m->data);
        VFPS_HIP(hipGetLastError());
        if ((rc = launch_project_x(m, m->data, m->proj)) != VFPS_OK) return rc;
        if ((rc = launch_integrate(m)) != VFPS_OK) return rc;
        if ((rc = vfps_mesh_normalize(m)) != VFPS_OK) return rc;
    }
    // constructor tail: updateXProjection, updateYProjection, integrate
    if ((rc = vfps_mesh_project_x(m)) != VFPS_OK) return rc;
    if ((rc = vfps_mesh_project_y(m)) != VFPS_OK) return rc;
    if ((rc = launch_integrate(m)) != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(st));
    *out = m;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_mesh_clone(const vfps_mesh* src, vfps_mesh** out)
{
    VFPS_CHECK_ARG(src && out, "vfps_mesh_clone: NULL argument");
    VFPS_TRY_BEGIN
    vfps_mesh* m = nullptr;
    int rc = mesh_alloc(src->dev, src->nb, src->n, src->ax0.min, src->ax0.max,
                        src->qscale, src->ax1.min, src->ax1.max, src->pscale,
                        src->filling_set.data(), &m);
    if (rc != VFPS_OK) {
        return rc;
    }
    m->charge = src->charge;
    m->current = src->current;
    hipStream_t st = src->dev->stream;
    VFPS_HIP(hipMemcpyAsync(m->data, src->data, sizeof(float) * m->cells(),
                            hipMemcpyDeviceToDevice, st));
    if ((rc = vfps_mesh_project_x(m)) != VFPS_OK) return rc;
    if ((rc = vfps_mesh_project_y(m)) != VFPS_OK) return rc;
    if ((rc = launch_integrate(m)) != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(st));
    *out = m;
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_mesh_destroy(vfps_mesh* m)
{
    if (m) {
        hipStreamSynchronize(m->dev->stream);
    }
    mesh_free(m);
    return VFPS_OK;
}

int vfps_mesh_upload(vfps_mesh* m, const float* data)
{
    VFPS_CHECK_ARG(m && data, "vfps_mesh_upload: NULL argument");
    VFPS_HIP(hipMemcpyAsync(m->data, data, sizeof(float) * m->cells(),
                            hipMemcpyHostToDevice, m->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mesh_download(vfps_mesh* m, float* data)
{
    VFPS_CHECK_ARG(m && data, "vfps_mesh_download: NULL argument");
    VFPS_HIP(hipMemcpyAsync(data, m->data, sizeof(float) * m->cells(),
                            hipMemcpyDeviceToHost, m->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mesh_copy(vfps_mesh* dst, const vfps_mesh* src)
{
    VFPS_CHECK_ARG(dst && src, "vfps_mesh_copy: NULL argument");
    VFPS_CHECK_ARG(dst->nb == src->nb && dst->n == src->n,
                   "vfps_mesh_copy: size mismatch");
    VFPS_HIP(hipMemcpyAsync(dst->data, src->data, sizeof(float) * dst->cells(),
                            hipMemcpyDeviceToDevice, dst->dev->stream));
    return VFPS_OK;
}

int vfps_mesh_device_ptr(vfps_mesh* m, float** dptr)
{
    VFPS_CHECK_ARG(m && dptr, "vfps_mesh_device_ptr: NULL argument");
    *dptr = m->data;
    return VFPS_OK;
}

int vfps_mesh_project_x(vfps_mesh* m)
{
    VFPS_CHECK_ARG(m, "vfps_mesh_project_x: NULL mesh");
    return launch_project_x(m, m->data, m->proj);
}

int vfps_mesh_project_y(vfps_mesh* m)
{
    VFPS_CHECK_ARG(m, "vfps_mesh_project_y: NULL mesh");
    hipLaunchKernelGGL(project_y_kernel, dim3(ceil_div(m->n, PY_COLS), m->nb),
                       dim3(PY_THREADS), 0, m->dev->stream, m->nb, m->n,
                       m->data, m->ws, m->proj + (size_t)m->nb * m->n);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_mesh_get_projection(vfps_mesh* m, int axis, float* out)
{
    VFPS_CHECK_ARG(m && out && (axis == 0 || axis == 1),
                   "vfps_mesh_get_projection: bad argument");
    VFPS_HIP(hipMemcpyAsync(out, m->proj + (size_t)axis * m->nb * m->n,
                            sizeof(float) * m->nb * m->n,
                            hipMemcpyDeviceToHost, m->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mesh_set_projection(vfps_mesh* m, int axis, const float* proj)
{
    VFPS_CHECK_ARG(m && proj && (axis == 0 || axis == 1),
                   "vfps_mesh_set_projection: bad argument");
    VFPS_HIP(hipMemcpyAsync(m->proj + (size_t)axis * m->nb * m->n, proj,
                            sizeof(float) * m->nb * m->n,
                            hipMemcpyHostToDevice, m->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mesh_integrate(vfps_mesh* m, double* filling, double* integral)
{
    VFPS_CHECK_ARG(m, "vfps_mesh_integrate: NULL mesh");
    int rc = launch_integrate(m);
    if (rc != VFPS_OK) {
        return rc;
    }
    if (filling) {
        VFPS_HIP(hipMemcpyAsync(filling, m->filling, sizeof(double) * m->nb,
                                hipMemcpyDeviceToHost, m->dev->stream));
    }
    if (integral) {
        VFPS_HIP(hipMemcpyAsync(integral, m->integral, sizeof(double),
                                hipMemcpyDeviceToHost, m->dev->stream));
    }
    if (filling || integral) {
        VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    }
    return VFPS_OK;
}

int vfps_mesh_normalize(vfps_mesh* m)
{
    VFPS_CHECK_ARG(m, "vfps_mesh_normalize: NULL mesh");
    const size_t nxy = (size_t)m->n * m->n;
    hipLaunchKernelGGL(normalize_kernel, dim3(ceil_div(m->cells(), 256)),
                       dim3(256), 0, m->dev->stream, m->nb, nxy, m->data,
                       m->fill_set_d, m->filling);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int vfps_mesh_variance(vfps_mesh* m, int axis, float* mean, float* var,
                       float* rms)
{
    VFPS_CHECK_ARG(m && (axis == 0 || axis == 1),
                   "vfps_mesh_variance: bad argument");
    VFPS_TRY_BEGIN
    const Ruler& ax = axis == 0 ? m->ax0 : m->ax1;
    float* tmp = nullptr;
    VFPS_HIP(hipMalloc(&tmp, sizeof(float) * (m->n + 3 * m->nb)));
    hipMemcpyAsync(tmp, ax.data.data(), sizeof(float) * m->n,
                   hipMemcpyHostToDevice, m->dev->stream);
    hipLaunchKernelGGL(variance_kernel, dim3(m->nb), dim3(256), 0,
                       m->dev->stream, m->n,
                       m->proj + (size_t)axis * m->nb * m->n, tmp, ax.delta,
                       m->fill_set_d, m->filling, tmp + m->n);
    std::vector<float> h(3 * m->nb);
    hipMemcpyAsync(h.data(), tmp + m->n, sizeof(float) * 3 * m->nb,
                   hipMemcpyDeviceToHost, m->dev->stream);
    hipError_t e = hipStreamSynchronize(m->dev->stream);
    hipFree(tmp);
    if (e != hipSuccess) {
        return fail(VFPS_ERR_HIP, hipGetErrorString(e));
    }
    for (uint32_t b = 0; b < m->nb; b++) {
        if (mean) mean[b] = h[3 * b];
        if (var) var[b] = h[3 * b + 1];
        if (rms) rms[b] = h[3 * b + 2];
    }
    return VFPS_OK;
    VFPS_TRY_END
}

int vfps_mesh_info(const vfps_mesh* m, uint32_t* nb, uint32_t* n,
                   float* delta0, float* delta1, float* zerobin0,
                   float* zerobin1)
{
    VFPS_CHECK_ARG(m, "vfps_mesh_info: NULL mesh");
    if (nb) *nb = m->nb;
    if (n) *n = m->n;
    if (delta0) *delta0 = m->ax0.delta;
    if (delta1) *delta1 = m->ax1.delta;
    if (zerobin0) *zerobin0 = m->ax0.zerobin;
    if (zerobin1) *zerobin1 = m->ax1.zerobin;
    return VFPS_OK;
}

int vfps_mesh_axis(const vfps_mesh* m, int axis, float* out)
{
    VFPS_CHECK_ARG(m && out && (axis == 0 || axis == 1),
                   "vfps_mesh_axis: bad argument");
    const Ruler& ax = axis == 0 ? m->ax0 : m->ax1;
    std::memcpy(out, ax.data.data(), sizeof(float) * m->n);
    return VFPS_OK;
}

int vfps_mesh_simpson_weights(const vfps_mesh* m, float* out)
{
    VFPS_CHECK_ARG(m && out, "vfps_mesh_simpson_weights: bad argument");
    std::memcpy(out, m->ws_h.data(), sizeof(float) * m->n);
    return VFPS_OK;
}

}  // extern "C"
