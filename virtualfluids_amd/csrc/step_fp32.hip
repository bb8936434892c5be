// step_fp32.hip -- FP32 arithmetic on FP32 or per-lattice-scaled FP16 storage.
#include "kernels.hpp"

namespace vfx {
VFX_INSTANTIATE_LAUNCHERS(float)
VFX_INSTANTIATE_LAUNCHERS(_Float16)
}
