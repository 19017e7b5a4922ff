// Fused step kernels: linear interpolation, direct wake, no FP and five-point FP stencil.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<2, true, 0>(const StepVariant&, int, uint32_t);
template const void* step_kernels<2, true, 5>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
