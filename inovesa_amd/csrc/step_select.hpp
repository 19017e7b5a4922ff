// Kernel lookup of the fused step by family.
#pragma once

#include <cstdint>

#include "step.hpp"

namespace vfps_impl {

// The step kernels of one family, interpolation IT (2 or 4 taps), wake form
// (DIRECT: Toeplitz in the kernel, else the field kernel's offsets) and FP
// stencil IPFP (0: none), in every scheme, launch form (StepForm) and
// column count: the kernel for (v.scheme, v.clamp, form, maxc), or nullptr.
// Defined in step_kernels_*.hip.
template <int IT, bool DIRECT, int IPFP>
const void* step_kernels(const StepVariant& v, int form, uint32_t maxc);

extern template const void* step_kernels<4, true, 0>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, true, 3>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, true, 4>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, true, 5>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, true, 0>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, true, 3>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, true, 4>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, true, 5>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, false, 0>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, false, 3>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, false, 4>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<4, false, 5>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, false, 0>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, false, 3>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, false, 4>(const StepVariant&, int, uint32_t);
extern template const void* step_kernels<2, false, 5>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
