// Fused step kernels: FFT-wake kernels (one-mesh and per-tile forms), cubic
// interpolation, four- and five-point Fokker-Planck stencils.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<4, false, 4>(const StepVariant&, int, uint32_t);
template const void* step_kernels<4, false, 5>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
