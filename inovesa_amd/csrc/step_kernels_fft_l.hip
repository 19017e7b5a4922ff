// Fused step kernels: FFT-wake kernels (one-mesh and per-tile forms), linear
// interpolation, every Fokker-Planck stencil.
#include "step_kernels_impl.hpp"

namespace vfps_impl {

template const void* step_kernels<2, false, 0>(const StepVariant&, int, uint32_t);
template const void* step_kernels<2, false, 3>(const StepVariant&, int, uint32_t);
template const void* step_kernels<2, false, 4>(const StepVariant&, int, uint32_t);
template const void* step_kernels<2, false, 5>(const StepVariant&, int, uint32_t);

}  // namespace vfps_impl
