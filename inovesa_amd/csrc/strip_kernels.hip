// Strip-form step kernels of a batch (kick-drift maps, direct wake) and
// their launch.
#include <map>
#include <mutex>

#include "strip_kernel.hpp"

namespace vfps_impl {

namespace {

template <int IT, int IPFP>
const void* strip_fn_cols(uint32_t maxc)
{
    if (maxc <= 1) {
        return reinterpret_cast<const void*>(&strip_kernel<IT, IPFP, 1>);
    }
    if (maxc == 2) {
        return reinterpret_cast<const void*>(&strip_kernel<IT, IPFP, 2>);
    }
    return reinterpret_cast<const void*>(&strip_kernel<IT, IPFP, STEP_MAX_COLS>);
}

template <int IT>
const void* strip_fn_fp(int ipfp, uint32_t maxc)
{
    switch (ipfp) {
    case 0: return strip_fn_cols<IT, 0>(maxc);
    case 3: return strip_fn_cols<IT, 3>(maxc);
    case 4: return strip_fn_cols<IT, 4>(maxc);
    case 5: return strip_fn_cols<IT, 5>(maxc);
    }
    return nullptr;
}

}  // namespace

const void* strip_kernel_for(const StepVariant& v, uint32_t maxc)
{
    if (v.scheme != 0 || !v.direct || maxc > (uint32_t)STEP_MAX_COLS) {
        return nullptr;
    }
    if (v.it == 4) return strip_fn_fp<4>(v.ipfp, maxc);
    if (v.it == 2) return strip_fn_fp<2>(v.ipfp, maxc);
    return nullptr;
}

size_t strip_lds_bytes(uint32_t n, uint32_t T, uint32_t K, uint32_t halo)
{
    return sizeof(float) * (size_t)strip_lds(n, T, K, halo).words;
}

uint32_t strip_ring(uint32_t T, uint32_t halo)
{
    return strip_ring_rows(T, halo);
}

int strip_set_lds_limit(const StepVariant& v, size_t bytes)
{
    static std::mutex mu;
    static std::map<const void*, size_t> set;
    std::lock_guard<std::mutex> lock(mu);
    for (uint32_t maxc : {1u, 2u, (uint32_t)STEP_MAX_COLS}) {
        const void* fn = strip_kernel_for(v, maxc);
        if (!fn) {
            return fail(VFPS_ERR_UNSUPPORTED, "strip step: no kernel for this "
                                              "variant");
        }
        if (bytes > 65536 && bytes > set[fn]) {
            VFPS_HIP(hipFuncSetAttribute(
                fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
            set[fn] = bytes;
        }
    }
    return VFPS_OK;
}

int launch_strip_kernel(const StepVariant& v, const InstDesc* descs_d,
                        const uint32_t* blockmap_d, uint32_t nblocks,
                        size_t lds_bytes, uint32_t threads, uint32_t maxn,
                        uint32_t parity, hipStream_t st)
{
    if (!blockmap_d || nblocks == 0 || nblocks % 8u != 0u) {
        return fail(VFPS_ERR_INVALID, "strip step: a block map of a multiple "
                                      "of 8 entries is needed");
    }
    const void* fn = strip_kernel_for(v, ceil_div(maxn, threads));
    if (!fn) {
        return fail(VFPS_ERR_UNSUPPORTED, "strip step: no kernel for this "
                                          "variant");
    }
    void* args[] = {(void*)&descs_d, (void*)&blockmap_d, (void*)&nblocks,
                    (void*)&parity};
    VFPS_HIP(hipLaunchKernel(fn, dim3(nblocks), dim3(threads), args, lds_bytes,
                             st));
    return VFPS_OK;
}

}  // namespace vfps_impl
