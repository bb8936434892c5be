// step_fp64.hip -- FP64 storage, FP64 arithmetic.  Built with -ffp-contract=off:
// no FMA contraction, so every operation rounds exactly as the reference CPU
// path (g++, SSE2, -ffp-contract=off) does.
#include "kernels.hpp"

namespace vfx {
VFX_INSTANTIATE_LAUNCHERS(double)
}
