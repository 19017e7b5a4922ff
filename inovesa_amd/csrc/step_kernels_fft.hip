// Fused step kernels: FFT-wake kernels (one-mesh and per-tile forms), cubic
// interpolation, no or three-point Fokker-Planck stencil.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<4, false, 0>(const StepVariant&, int, uint32_t);
template const void* step_kernels<4, false, 3>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
