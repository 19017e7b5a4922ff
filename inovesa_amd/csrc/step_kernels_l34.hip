// Fused step kernels: linear interpolation, direct wake, two-sided and cubic FP stencils.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<2, true, 3>(const StepVariant&, int, uint32_t);
template const void* step_kernels<2, true, 4>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
