// Kernel tables of the fused step: the definition of step_kernels, to be
// included only by the step_kernels_*.hip files, each of which instantiates
// some (interpolation, wake form, FP stencil) families, so the many kernel
// variants compile in parallel.
#pragma once

#include "step_kernel.hpp"
#include "step_select.hpp"

namespace vfps_impl {

template <int IT, int IPFP, bool DIRECT, int SCHEME, bool CLAMP, int FORM>
static const void* step_fn_cols(uint32_t maxc)
{
    if (maxc <= 1) {
        return reinterpret_cast<const void*>(
            &step_kernel<IT, IPFP, DIRECT, SCHEME, CLAMP, FORM, 1>);
    }
    if (maxc == 2) {
        return reinterpret_cast<const void*>(
            &step_kernel<IT, IPFP, DIRECT, SCHEME, CLAMP, FORM, 2>);
    }
    return reinterpret_cast<const void*>(
        &step_kernel<IT, IPFP, DIRECT, SCHEME, CLAMP, FORM, STEP_MAX_COLS>);
}

// the FFT-wake kernels exist in the one-mesh and per-tile forms (a batch's
// resident-grid form runs the direct wake only)
template <int IT, int IPFP, bool DIRECT, int SCHEME, bool CLAMP>
static const void* step_fn(int form, uint32_t maxc)
{
    if constexpr (!DIRECT) {
        if (form == STEP_SINGLE) {
            return step_fn_cols<IT, IPFP, DIRECT, SCHEME, CLAMP, STEP_SINGLE>(maxc);
        }
        if (form == STEP_PER_TILE) {
            return step_fn_cols<IT, IPFP, DIRECT, SCHEME, CLAMP, STEP_PER_TILE>(maxc);
        }
        return nullptr;
    }
    switch (form) {
    case STEP_SINGLE:
        return step_fn_cols<IT, IPFP, DIRECT, SCHEME, CLAMP, STEP_SINGLE>(maxc);
    case STEP_PER_TILE:
        return step_fn_cols<IT, IPFP, DIRECT, SCHEME, CLAMP, STEP_PER_TILE>(maxc);
    case STEP_PERSISTENT:
        return step_fn_cols<IT, IPFP, DIRECT, SCHEME, CLAMP, STEP_PERSISTENT>(maxc);
    }
    return nullptr;
}

template <int IT, bool DIRECT, int IPFP>
const void* step_kernels(const StepVariant& v, int form, uint32_t maxc)
{
    if (v.scheme == 0) return step_fn<IT, IPFP, DIRECT, 0, false>(form, maxc);
    if (v.clamp) {
        if constexpr (IT == 4) {
            return step_fn<IT, IPFP, DIRECT, 1, true>(form, maxc);
        }
        return nullptr;  // the clamp needs cubic interpolation
    }
    return step_fn<IT, IPFP, DIRECT, 1, false>(form, maxc);
}

}  // namespace vfps_impl
