"""MI355X-native per-step phase-space update of a Vlasov-Fokker-Planck
solver, with the interfaces of Inovesa's PhaseSpace, SourceMaps and
ElectricField.  The compute runs in libvfps.so (HIP, gfx950); see
include/vfps_gpu.h for the C ABI."""
from .vfps import *  # noqa: F401,F403
from .vfps import __all__  # noqa: F401
