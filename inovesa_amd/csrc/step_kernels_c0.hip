// Fused step kernels: cubic interpolation, direct wake, no FP.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<4, true, 0>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
