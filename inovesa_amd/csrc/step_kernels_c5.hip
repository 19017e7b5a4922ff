// Fused step kernels: cubic interpolation, direct wake, five-point FP stencil.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<4, true, 5>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
