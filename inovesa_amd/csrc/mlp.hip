// Multilayer perceptron with GELU activations, trained and evaluated with
// hand-written kernels.  No reference counterpart: it is the learned
// surrogate that maps (bunch current, chamber gap, damping time) to the CSR
// power spectrum and the instability verdict of the VFP solver.
//
// Dense layers: y = x W^T + b with W [out][in].  The products run on the
// matrix cores (MFMA 16x16x32, BF16 operands, FP32 accumulation), one
// wavefront per 16x16 output tile, with the bias, GELU and its derivative
// fused into the store.  Master weights, gradients and Adam moments are
// FP32; each step writes the BF16 copy the products read.
//
// Parameter layout (flattened FP32): for each layer W (out x in, row-major)
// then b (out).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "reduce.hpp"

namespace {

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (NaN stays NaN)
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
        return (uint16_t)((u >> 16) | 0x0040u);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// GELU with the exact Gaussian CDF: x Phi(x), and its derivative
__device__ inline float gelu(float x)
{
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ inline float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

enum Epilogue : int {
    EPI_STORE = 0,       // C fp32
    EPI_BIAS_GELU = 1,   // Z = acc + b (fp32), A = gelu(Z) (bf16)
    EPI_BIAS = 2,        // Y = acc + b (fp32)
    EPI_GELU_BACK = 3,   // dZ = acc * gelu'(Zp): fp32 and bf16
};

struct GemmArgs {
    const uint16_t* A;   // bf16, A(i,k) = A[i*asi + k*ask]
    const uint16_t* B;   // bf16, B(k,j) = B[k*bsk + j*bsj]
    int64_t asi, ask, bsk, bsj;
    uint32_t M, N, K;
    float* C;            // fp32 output, C[i*ldc + j]
    uint16_t* Cb;        // bf16 output (GELU / GELU_BACK)
    const float* bias;   // [N]
    const float* Zp;     // GELU_BACK: pre-activations, Zp[i*ldc + j]
    uint32_t ldc;
};

// One wavefront computes a 16x16 tile of C with v_mfma_f32_16x16x32_bf16
// (gfx950: 32 products of K per instruction).  Operand layout of the
// instruction: lane l holds A(l%16, 8(l/16)+0..7) and B(8(l/16)+0..7,
// l%16); it receives C(4(l/16)+r, l%16), r = 0..3.
template <int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t tiles_n = (g.N + 15u) / 16u;
    const uint32_t tiles = ((g.M + 15u) / 16u) * tiles_n;
    if (wave >= tiles) {
        return;
    }
    const uint32_t ti = (wave / tiles_n) * 16u, tj = (wave % tiles_n) * 16u;
    const uint32_t r16 = lane & 15u, kq = (lane >> 4) * 8u;
    const uint32_t ia = ti + r16;   // A row of this lane
    const uint32_t jb = tj + r16;   // B column of this lane
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k0 = 0; k0 < g.K; k0 += 32u) {
        u16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t k = k0 + kq + (uint32_t)t;
            const bool ka = ia < g.M && k < g.K;
            const bool kb = jb < g.N && k < g.K;
            a[t] = ka ? g.A[(int64_t)ia * g.asi + (int64_t)k * g.ask] : (uint16_t)0;
            b[t] = kb ? g.B[(int64_t)k * g.bsk + (int64_t)jb * g.bsj] : (uint16_t)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
            0, 0, 0);
    }
    const uint32_t j = tj + r16;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t i = ti + (lane >> 4) * 4u + (uint32_t)r;
        if (i >= g.M || j >= g.N) {
            continue;
        }
        const size_t o = (size_t)i * g.ldc + j;
        const float v = acc[r];
        if (EPI == EPI_STORE) {
            g.C[o] = v;
        } else if (EPI == EPI_BIAS) {
            g.C[o] = v + g.bias[j];
        } else if (EPI == EPI_BIAS_GELU) {
            const float z = v + g.bias[j];
            g.C[o] = z;
            g.Cb[o] = f2bf(gelu(z));
        } else {
            const float dz = v * gelu_grad(g.Zp[o]);
            g.C[o] = dz;
            g.Cb[o] = f2bf(dz);
        }
    }
}

template <int EPI>
int launch_gemm(const GemmArgs& g, hipStream_t st)
{
    const uint32_t tiles = ((g.M + 15u) / 16u) * ((g.N + 15u) / 16u);
    const uint32_t blocks = (tiles + 3u) / 4u;
    hipLaunchKernelGGL(gemm_kernel<EPI>, dim3(blocks), dim3(256), 0, st, g);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y,
                               size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        y[i] = f2bf(x[i]);
    }
}

// rows idx[b] of x (rows of `w` floats) into y, for b < B
__global__ void gather_kernel(const float* __restrict__ x, const uint32_t* __restrict__ idx,
                              uint32_t B, uint32_t w, float* __restrict__ y)
{
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < B * w;
         t += gridDim.x * blockDim.x) {
        const uint32_t b = t / w, c = t % w;
        y[t] = x[(size_t)idx[b] * w + c];
    }
}

// Loss of one batch and its gradient with respect to the outputs.
// Y, dY: [B][S+1]; the first S outputs are regressed onto T [B][S] (mean
// squared error), the last is the logit of the label L [B] (binary cross
// entropy, weight wl).  loss = (1/B) sum_b [ (1/S) sum_k (y-t)^2 +
// wl BCE ].  One workgroup; partial[0] receives the loss.
__global__ void __launch_bounds__(256)
loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
            const float* __restrict__ L, uint32_t B, uint32_t S, float wl,
            float* __restrict__ dY, uint16_t* __restrict__ dYb,
            float* __restrict__ loss_out)
{
    __shared__ double red[16];
    const uint32_t W = S + 1;
    double acc = 0.0;
    for (uint32_t t = threadIdx.x; t < B * W; t += blockDim.x) {
        const uint32_t b = t / W, k = t % W;
        const float y = Y[t];
        float g;
        if (k < S) {
            const float d = y - T[(size_t)b * S + k];
            acc += (double)d * (double)d / (double)S;
            g = 2.0f * d / (float)(S * B);
        } else {
            const float lab = L[b];
            // BCE with logits: max(y,0) - y*lab + log(1 + exp(-|y|))
            acc += (double)wl * ((double)fmaxf(y, 0.0f) - (double)y * lab +
                                 log1p(exp(-fabs((double)y))));
            const float sg = 1.0f / (1.0f + expf(-y));
            g = wl * (sg - lab) / (float)B;
        }
        dY[t] = g;
        dYb[t] = f2bf(g);
    }
    acc = vfps_impl::block_sum(acc, red);
    if (threadIdx.x == 0) {
        *loss_out = (float)(acc / (double)B);
    }
}

// db[j] = sum_b dZ[b][j]
__global__ void colsum_kernel(const float* __restrict__ dZ, uint32_t B, uint32_t N,
                              float* __restrict__ db)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < B; b++) s += dZ[(size_t)b * N + j];
    db[j] = s;
}

// Adam (Kingma & Ba): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); BF16 copy of p
__global__ void adam_kernel(float* __restrict__ p, uint16_t* __restrict__ pb,
                            const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / c1;
        const float vh = vi / c2;
        const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
        p[i] = pi;
        pb[i] = f2bf(pi);
    }
}

}  // namespace

struct vfps_mlp {
    vfps_dev* dev = nullptr;
    std::vector<uint32_t> dims;     // n_in, hidden..., n_out
    std::vector<size_t> w_off, b_off;
    size_t nparams = 0;
    uint32_t max_batch = 0;
    float* p = nullptr;             // FP32 master
    uint16_t* pb = nullptr;         // BF16 copy
    float* g = nullptr;             // gradients
    float* m = nullptr;             // Adam moments
    float* v = nullptr;
    uint32_t t = 0;                 // Adam steps taken
    // workspace
    float* x = nullptr;             // [B][n_in]
    uint16_t* xb = nullptr;
    std::vector<float*> z;          // per layer pre-activation [B][out]
    std::vector<uint16_t*> ab;      // per hidden layer GELU output (bf16)
    float* dy = nullptr;            // [B][max width]
    uint16_t* dyb = nullptr;
    float* dz2 = nullptr;
    uint16_t* dzb2 = nullptr;
    float* tgt = nullptr;           // [B][n_out-1]
    float* lab = nullptr;           // [B]
    float* loss = nullptr;          // [1]
    uint32_t nlayers() const { return (uint32_t)dims.size() - 1; }
    uint32_t maxw() const { return *std::max_element(dims.begin(), dims.end()); }
};

namespace {

void mlp_free(vfps_mlp* m)
{
    if (!m) return;
    if (m->dev) (void)hipStreamSynchronize(m->dev->stream);
    for (void* q : {(void*)m->p, (void*)m->pb, (void*)m->g, (void*)m->m,
                    (void*)m->v, (void*)m->x, (void*)m->xb, (void*)m->dy,
                    (void*)m->dyb, (void*)m->dz2, (void*)m->dzb2,
                    (void*)m->tgt, (void*)m->lab, (void*)m->loss}) {
        (void)hipFree(q);
    }
    for (auto* q : m->z) (void)hipFree(q);
    for (auto* q : m->ab) (void)hipFree(q);
    delete m;
}

int forward(vfps_mlp* m, uint32_t B)
{
    hipStream_t st = m->dev->stream;
    const uint32_t L = m->nlayers();
    const uint16_t* in = m->xb;
    for (uint32_t l = 0; l < L; l++) {
        const uint32_t ni = m->dims[l], no = m->dims[l + 1];
        GemmArgs g{};
        g.A = in;
        g.asi = ni;
        g.ask = 1;
        g.B = m->pb + m->w_off[l];
        g.bsk = 1;
        g.bsj = ni;
        g.M = B;
        g.N = no;
        g.K = ni;
        g.C = m->z[l];
        g.ldc = no;
        g.bias = m->p + m->b_off[l];
        int rc;
        if (l + 1 < L) {
            g.Cb = m->ab[l];
            rc = launch_gemm<EPI_BIAS_GELU>(g, st);
            in = m->ab[l];
        } else {
            rc = launch_gemm<EPI_BIAS>(g, st);
        }
        if (rc != VFPS_OK) return rc;
    }
    return VFPS_OK;
}

// gradients of the loss into m->g (after loss_kernel wrote dy/dyb)
int backward(vfps_mlp* m, uint32_t B)
{
    hipStream_t st = m->dev->stream;
    const uint32_t L = m->nlayers();
    float* dz = m->dy;
    uint16_t* dzb = m->dyb;
    float* dz_other = m->dz2;
    uint16_t* dzb_other = m->dzb2;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint32_t ni = m->dims[l], no = m->dims[l + 1];
        const uint16_t* xin = l == 0 ? m->xb : m->ab[l - 1];
        // dW[j][k] = sum_b dZ[b][j] X[b][k]
        GemmArgs g{};
        g.A = dzb;
        g.asi = 1;
        g.ask = no;
        g.B = xin;
        g.bsk = ni;
        g.bsj = 1;
        g.M = no;
        g.N = ni;
        g.K = B;
        g.C = m->g + m->w_off[l];
        g.ldc = ni;
        int rc = launch_gemm<EPI_STORE>(g, st);
        if (rc != VFPS_OK) return rc;
        hipLaunchKernelGGL(colsum_kernel, dim3((no + 255) / 256), dim3(256), 0,
                           st, dz, B, no, m->g + m->b_off[l]);
        VFPS_HIP(hipGetLastError());
        if (l == 0) break;
        // dA[b][k] = sum_j dZ[b][j] W[j][k]; dZprev = dA * gelu'(Zprev)
        GemmArgs h{};
        h.A = dzb;
        h.asi = no;
        h.ask = 1;
        h.B = m->pb + m->w_off[l];
        h.bsk = ni;
        h.bsj = 1;
        h.M = B;
        h.N = ni;
        h.K = no;
        h.C = dz_other;
        h.Cb = dzb_other;
        h.Zp = m->z[l - 1];
        h.ldc = ni;
        rc = launch_gemm<EPI_GELU_BACK>(h, st);
        if (rc != VFPS_OK) return rc;
        // the buffers of the output layer (dy) are free once it is done
        float* tz = dz;
        uint16_t* tzb = dzb;
        dz = dz_other;
        dzb = dzb_other;
        if (tz == m->dy) {
            dz_other = m->dy;
            dzb_other = m->dyb;
        } else {
            dz_other = tz;
            dzb_other = tzb;
        }
    }
    return VFPS_OK;
}

int upload_inputs(vfps_mlp* m, const float* x, uint32_t B, bool host)
{
    hipStream_t st = m->dev->stream;
    const size_t n = (size_t)B * m->dims[0];
    if (host) {
        VFPS_HIP(hipMemcpyAsync(m->x, x, sizeof(float) * n, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3(std::min<size_t>((n + 255) / 256, 1024)),
                       dim3(256), 0, st, m->x, m->xb, n);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

int loss_and_grad(vfps_mlp* m, uint32_t B, float wl)
{
    hipStream_t st = m->dev->stream;
    const uint32_t S = m->dims.back() - 1;
    int rc = forward(m, B);
    if (rc != VFPS_OK) return rc;
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, st,
                       m->z[m->nlayers() - 1], m->tgt, m->lab, B, S, wl, m->dy,
                       m->dyb, m->loss);
    VFPS_HIP(hipGetLastError());
    return backward(m, B);
}

int adam(vfps_mlp* m, float lr, float b1, float b2, float eps)
{
    m->t++;
    const float c1 = 1.0f - std::pow(b1, (float)m->t);
    const float c2 = 1.0f - std::pow(b2, (float)m->t);
    hipLaunchKernelGGL(adam_kernel, dim3(std::min<size_t>((m->nparams + 255) / 256, 1024)),
                       dim3(256), 0, m->dev->stream, m->p, m->pb, m->g, m->m,
                       m->v, m->nparams, lr, b1, b2, eps, c1, c2);
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
}

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int vfps_mlp_create(vfps_dev* dev, const uint32_t* dims, uint32_t ndims,
                    uint32_t max_batch, uint64_t seed, vfps_mlp** out)
{
    VFPS_CHECK_ARG(dev && dims && out && ndims >= 2 && max_batch > 0,
                   "vfps_mlp_create: bad argument");
    VFPS_CHECK_ARG(dims[ndims - 1] >= 2, "vfps_mlp_create: the output needs "
                                         "at least one regression value and "
                                         "the logit");
    for (uint32_t i = 0; i < ndims; i++) {
        VFPS_CHECK_ARG(dims[i] >= 1 && dims[i] <= 4096,
                       "vfps_mlp_create: layer widths must be in [1, 4096]");
    }
    VFPS_TRY_BEGIN
    auto* m = new vfps_mlp();
    std::unique_ptr<vfps_mlp, void (*)(vfps_mlp*)> guard(m, mlp_free);
    m->dev = dev;
    m->dims.assign(dims, dims + ndims);
    m->max_batch = max_batch;
    size_t off = 0;
    for (uint32_t l = 0; l + 1 < ndims; l++) {
        m->w_off.push_back(off);
        off += (size_t)dims[l] * dims[l + 1];
        m->b_off.push_back(off);
        off += dims[l + 1];
    }
    m->nparams = off;
    // He-style uniform initialisation, biases zero
    std::vector<float> p(off, 0.0f);
    uint64_t s = seed;
    for (uint32_t l = 0; l + 1 < ndims; l++) {
        const float bound = std::sqrt(6.0f / (float)dims[l]) * 0.5f;
        for (size_t i = 0; i < (size_t)dims[l] * dims[l + 1]; i++) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            p[m->w_off[l] + i] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    const size_t n = off;
    VFPS_HIP(hipMalloc(&m->p, sizeof(float) * n));
    VFPS_HIP(hipMalloc(&m->pb, sizeof(uint16_t) * n));
    VFPS_HIP(hipMalloc(&m->g, sizeof(float) * n));
    VFPS_HIP(hipMalloc(&m->m, sizeof(float) * n));
    VFPS_HIP(hipMalloc(&m->v, sizeof(float) * n));
    VFPS_HIP(hipMemset(m->g, 0, sizeof(float) * n));
    VFPS_HIP(hipMemset(m->m, 0, sizeof(float) * n));
    VFPS_HIP(hipMemset(m->v, 0, sizeof(float) * n));
    const size_t B = max_batch;
    const uint32_t W = m->maxw();
    VFPS_HIP(hipMalloc(&m->x, sizeof(float) * B * dims[0]));
    VFPS_HIP(hipMalloc(&m->xb, sizeof(uint16_t) * B * dims[0]));
    for (uint32_t l = 0; l + 1 < ndims; l++) {
        float* z = nullptr;
        VFPS_HIP(hipMalloc(&z, sizeof(float) * B * dims[l + 1]));
        m->z.push_back(z);
        if (l + 2 < ndims) {
            uint16_t* a = nullptr;
            VFPS_HIP(hipMalloc(&a, sizeof(uint16_t) * B * dims[l + 1]));
            m->ab.push_back(a);
        }
    }
    VFPS_HIP(hipMalloc(&m->dy, sizeof(float) * B * W));
    VFPS_HIP(hipMalloc(&m->dyb, sizeof(uint16_t) * B * W));
    VFPS_HIP(hipMalloc(&m->dz2, sizeof(float) * B * W));
    VFPS_HIP(hipMalloc(&m->dzb2, sizeof(uint16_t) * B * W));
    VFPS_HIP(hipMalloc(&m->tgt, sizeof(float) * B * (dims[ndims - 1] - 1)));
    VFPS_HIP(hipMalloc(&m->lab, sizeof(float) * B));
    VFPS_HIP(hipMalloc(&m->loss, sizeof(float)));
    *out = guard.release();
    return vfps_mlp_set_params(*out, p.data(), n);
    VFPS_TRY_END
}

int vfps_mlp_destroy(vfps_mlp* m)
{
    mlp_free(m);
    return VFPS_OK;
}

int vfps_mlp_nparams(const vfps_mlp* m, size_t* n)
{
    VFPS_CHECK_ARG(m && n, "vfps_mlp_nparams: NULL argument");
    *n = m->nparams;
    return VFPS_OK;
}

int vfps_mlp_set_params(vfps_mlp* m, const float* params, size_t n)
{
    VFPS_CHECK_ARG(m && params && n == m->nparams,
                   "vfps_mlp_set_params: size mismatch");
    hipStream_t st = m->dev->stream;
    VFPS_HIP(hipMemcpyAsync(m->p, params, sizeof(float) * n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(std::min<size_t>((n + 255) / 256, 1024)),
                       dim3(256), 0, st, m->p, m->pb, n);
    VFPS_HIP(hipGetLastError());
    VFPS_HIP(hipMemsetAsync(m->m, 0, sizeof(float) * n, st));
    VFPS_HIP(hipMemsetAsync(m->v, 0, sizeof(float) * n, st));
    m->t = 0;
    VFPS_HIP(hipStreamSynchronize(st));
    return VFPS_OK;
}

int vfps_mlp_get_params(vfps_mlp* m, float* params, float* grads, size_t n)
{
    VFPS_CHECK_ARG(m && n == m->nparams, "vfps_mlp_get_params: size mismatch");
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    if (params) {
        VFPS_HIP(hipMemcpy(params, m->p, sizeof(float) * n, hipMemcpyDeviceToHost));
    }
    if (grads) {
        VFPS_HIP(hipMemcpy(grads, m->g, sizeof(float) * n, hipMemcpyDeviceToHost));
    }
    return VFPS_OK;
}

int vfps_mlp_forward(vfps_mlp* m, const float* x, uint32_t B, float* y)
{
    VFPS_CHECK_ARG(m && x && y && B >= 1 && B <= m->max_batch,
                   "vfps_mlp_forward: bad argument");
    int rc = upload_inputs(m, x, B, true);
    if (rc != VFPS_OK) return rc;
    if ((rc = forward(m, B)) != VFPS_OK) return rc;
    VFPS_HIP(hipMemcpyAsync(y, m->z[m->nlayers() - 1],
                            sizeof(float) * B * m->dims.back(),
                            hipMemcpyDeviceToHost, m->dev->stream));
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mlp_loss_grad(vfps_mlp* m, const float* x, const float* target,
                       const float* label, uint32_t B, float label_weight,
                       float* loss)
{
    VFPS_CHECK_ARG(m && x && target && label && B >= 1 && B <= m->max_batch,
                   "vfps_mlp_loss_grad: bad argument");
    hipStream_t st = m->dev->stream;
    const uint32_t S = m->dims.back() - 1;
    int rc = upload_inputs(m, x, B, true);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipMemcpyAsync(m->tgt, target, sizeof(float) * B * S,
                            hipMemcpyHostToDevice, st));
    VFPS_HIP(hipMemcpyAsync(m->lab, label, sizeof(float) * B,
                            hipMemcpyHostToDevice, st));
    if ((rc = loss_and_grad(m, B, label_weight)) != VFPS_OK) return rc;
    if (loss) {
        VFPS_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    }
    VFPS_HIP(hipStreamSynchronize(st));
    return VFPS_OK;
}

int vfps_mlp_adam_step(vfps_mlp* m, float lr, float beta1, float beta2,
                       float eps)
{
    VFPS_CHECK_ARG(m && lr >= 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 &&
                       beta2 < 1 && eps > 0,
                   "vfps_mlp_adam_step: bad argument");
    int rc = adam(m, lr, beta1, beta2, eps);
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipStreamSynchronize(m->dev->stream));
    return VFPS_OK;
}

int vfps_mlp_train(vfps_mlp* m, const float* x, const float* target,
                   const float* label, uint32_t N, uint32_t batch,
                   uint32_t epochs, float lr, float label_weight,
                   uint64_t seed, float* epoch_loss)
{
    VFPS_CHECK_ARG(m && x && target && label && N >= 1 && batch >= 1 &&
                       batch <= m->max_batch,
                   "vfps_mlp_train: bad argument");
    VFPS_TRY_BEGIN
    hipStream_t st = m->dev->stream;
    const uint32_t nin = m->dims[0], S = m->dims.back() - 1;
    float *dx = nullptr, *dt = nullptr, *dl = nullptr, *losses = nullptr;
    uint32_t* didx = nullptr;
    const uint32_t nb = (N + batch - 1) / batch;
    int rc = VFPS_OK;
    auto cleanup = [&]() {
        (void)hipFree(dx);
        (void)hipFree(dt);
        (void)hipFree(dl);
        (void)hipFree(didx);
        (void)hipFree(losses);
    };
    if (hipMalloc(&dx, sizeof(float) * N * nin) != hipSuccess ||
        hipMalloc(&dt, sizeof(float) * N * S) != hipSuccess ||
        hipMalloc(&dl, sizeof(float) * N) != hipSuccess ||
        hipMalloc(&didx, sizeof(uint32_t) * N) != hipSuccess ||
        hipMalloc(&losses, sizeof(float) * nb) != hipSuccess) {
        cleanup();
        return vfps_impl::fail(VFPS_ERR_NOMEM, "vfps_mlp_train: device memory");
    }
    if (hipMemcpy(dx, x, sizeof(float) * N * nin, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, target, sizeof(float) * N * S, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dl, label, sizeof(float) * N, hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return vfps_impl::fail(VFPS_ERR_HIP, "vfps_mlp_train: upload failed");
    }
    std::vector<uint32_t> perm(N);
    std::vector<float> lh(nb);
    uint64_t s = seed;
    for (uint32_t e = 0; e < epochs && rc == VFPS_OK; e++) {
        std::iota(perm.begin(), perm.end(), 0u);
        for (uint32_t i = N - 1; i > 0; i--) {
            std::swap(perm[i], perm[splitmix(s) % (i + 1)]);
        }
        hipMemcpyAsync(didx, perm.data(), sizeof(uint32_t) * N,
                       hipMemcpyHostToDevice, st);
        for (uint32_t bi = 0; bi < nb && rc == VFPS_OK; bi++) {
            const uint32_t b0 = bi * batch;
            const uint32_t B = std::min(batch, N - b0);
            hipLaunchKernelGGL(gather_kernel, dim3(64), dim3(256), 0, st, dx,
                               didx + b0, B, nin, m->x);
            hipLaunchKernelGGL(gather_kernel, dim3(64), dim3(256), 0, st, dt,
                               didx + b0, B, S, m->tgt);
            hipLaunchKernelGGL(gather_kernel, dim3(64), dim3(256), 0, st, dl,
                               didx + b0, B, 1u, m->lab);
            rc = upload_inputs(m, nullptr, B, false);
            if (rc == VFPS_OK) rc = loss_and_grad(m, B, label_weight);
            if (rc == VFPS_OK) rc = adam(m, lr, 0.9f, 0.999f, 1e-8f);
            if (rc == VFPS_OK) {
                hipMemcpyAsync(losses + bi, m->loss, sizeof(float),
                               hipMemcpyDeviceToDevice, st);
            }
        }
        if (rc == VFPS_OK && epoch_loss) {
            hipMemcpyAsync(lh.data(), losses, sizeof(float) * nb,
                           hipMemcpyDeviceToHost, st);
            hipStreamSynchronize(st);
            double a = 0;
            for (float v : lh) a += v;
            epoch_loss[e] = (float)(a / nb);
        }
    }
    hipStreamSynchronize(st);
    cleanup();
    if (rc != VFPS_OK) return rc;
    VFPS_HIP(hipGetLastError());
    return VFPS_OK;
    VFPS_TRY_END
}

// C = A B with BF16 operands (rounded from the FP32 inputs) and FP32
// accumulation on the matrix cores: the product the layers use.
// A(i,k) = A[i*asi + k*ask], B(k,j) = B[k*bsk + j*bsj]; C [M][N].
int vfps_gemm_bf16(vfps_dev* dev, uint32_t M, uint32_t N, uint32_t K,
                   const float* A, size_t a_len, int64_t asi, int64_t ask,
                   const float* B, size_t b_len, int64_t bsk, int64_t bsj,
                   float* C)
{
    VFPS_CHECK_ARG(dev && A && B && C && M && N && K, "vfps_gemm_bf16: bad argument");
    VFPS_CHECK_ARG(asi >= 0 && ask >= 0 && bsk >= 0 && bsj >= 0,
                   "vfps_gemm_bf16: strides must be non-negative");
    VFPS_CHECK_ARG((size_t)((M - 1) * asi + (K - 1) * ask) < a_len &&
                       (size_t)((K - 1) * bsk + (N - 1) * bsj) < b_len,
                   "vfps_gemm_bf16: operand too short for the strides");
    VFPS_TRY_BEGIN
    hipStream_t st = dev->stream;
    std::vector<uint16_t> ab(a_len), bb(b_len);
    for (size_t i = 0; i < a_len; i++) ab[i] = f2bf(A[i]);
    for (size_t i = 0; i < b_len; i++) bb[i] = f2bf(B[i]);
    uint16_t *da = nullptr, *db = nullptr;
    float* dc = nullptr;
    VFPS_HIP(hipMalloc(&da, sizeof(uint16_t) * a_len));
    VFPS_HIP(hipMalloc(&db, sizeof(uint16_t) * b_len));
    VFPS_HIP(hipMalloc(&dc, sizeof(float) * M * N));
    hipMemcpy(da, ab.data(), sizeof(uint16_t) * a_len, hipMemcpyHostToDevice);
    hipMemcpy(db, bb.data(), sizeof(uint16_t) * b_len, hipMemcpyHostToDevice);
    GemmArgs g{};
    g.A = da;
    g.B = db;
    g.asi = asi;
    g.ask = ask;
    g.bsk = bsk;
    g.bsj = bsj;
    g.M = M;
    g.N = N;
    g.K = K;
    g.C = dc;
    g.ldc = N;
    int rc = launch_gemm<EPI_STORE>(g, st);
    hipStreamSynchronize(st);
    if (rc == VFPS_OK) {
        hipMemcpy(C, dc, sizeof(float) * M * N, hipMemcpyDeviceToHost);
    }
    hipFree(da);
    hipFree(db);
    hipFree(dc);
    return rc;
    VFPS_TRY_END
}

// BF16 rounding as the kernels do it (host; for the checkers)
int vfps_round_bf16(const float* x, float* y, size_t n)
{
    VFPS_CHECK_ARG(x && y, "vfps_round_bf16: NULL argument");
    for (size_t i = 0; i < n; i++) y[i] = bf2f(f2bf(x[i]));
    return VFPS_OK;
}

}  // extern "C"
