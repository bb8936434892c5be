// d3q27.hpp -- D3Q27 lattice constants shared by the gfx950 kernels.
//
// Direction numbering follows VirtualFluids (src/lbm/constants/D3Q27.h:42-72) so
// that a distribution array written here is the array VirtualFluids' GPU path
// would hold: 27 planes of n_nodes values, plane d at d*n_nodes.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vfx {

enum Dir : int {
    d000 = 0, dP00, dM00, d0P0, d0M0, d00P, d00M, dPP0, dMM0, dPM0, dMP0,
    dP0P, dM0M, dP0M, dM0P, d0PP, d0MM, d0PM, d0MP,
    dPPP, dMPP, dPMP, dMMP, dPPM, dMPM, dPMM, dMMM
};

__host__ __device__ constexpr int cx(int d)
{
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
    return t[d];
}
__host__ __device__ constexpr int cy(int d)
{
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    return t[d];
}
__host__ __device__ constexpr int cz(int d)
{
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, 1, -1, -1, -1, -1};
    return t[d];
}
// opposite direction (D3Q27.h:85-110 iXXX)
__host__ __device__ constexpr int inv(int d)
{
    constexpr int t[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 26, 25, 24, 23, 22, 21, 20, 19};
    return t[d];
}
// 0 rest, 1 face, 2 edge, 3 corner
__host__ __device__ constexpr int weight_class(int d)
{
    return (cx(d) != 0) + (cy(d) != 0) + (cz(d) != 0);
}
template <typename T> __host__ __device__ constexpr T weight(int d)
{
    return weight_class(d) == 0 ? T(8. / 27.) : weight_class(d) == 1 ? T(2. / 27.) : weight_class(d) == 2 ? T(1. / 54.) : T(1. / 216.);
}

// Esoteric twist (VirtualFluids KernelUtilities.h:66-125, :252-379).  In a step
// of parity `even`, direction d of node k is read from, and the post-collision
// population of the opposite direction written back to, the same element:
//     plane  slot(d)  = even ? d : inv(d)
//     node   k + off(d),  off(d) = one step along every axis where c_d < 0
// Each element is touched by exactly one node per step (in-place, race free).
__host__ __device__ constexpr int slot(int d, bool even) { return even ? d : inv(d); }
__host__ __device__ constexpr int offx(int d) { return cx(d) < 0 ? 1 : 0; }
__host__ __device__ constexpr int offy(int d) { return cy(d) < 0 ? 1 : 0; }
__host__ __device__ constexpr int offz(int d) { return cz(d) < 0 ? 1 : 0; }
// which of the 8 neighbour indices (bit0 = +x, bit1 = +y, bit2 = +z)
__host__ __device__ constexpr int nbr(int d) { return offx(d) | (offy(d) << 1) | (offz(d) << 2); }

} // namespace vfx
