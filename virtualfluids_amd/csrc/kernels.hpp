// kernels.hpp -- gfx950 kernels of the D3Q27 stepper (templates).
//
// Each translation unit includes this header and instantiates the launchers
// for its storage types; step_fp64.hip is compiled with -ffp-contract=off so
// the FP64 path reproduces the reference CPU arithmetic exactly, step_fp32.hip
// (FP32 and scaled-FP16 storage) lets the compiler contract to FMA.
#pragma once

#include "collision.hpp"
#include "../../include/vfx.h"

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

namespace vfx {

// ------------------------------------------------------------- storage ----
// Population storage S with arithmetic type T.  _Float16 holds
// h_d = (f_d - w_d) * scale / w_d; the per-weight-class factors are kernel
// arguments (ld[c] = w_c / scale, st[c] = scale / w_c).
// Element i of a plane is addressed as plane base (wave-uniform, scalar
// registers) plus a 32-bit byte offset, so the loads and stores use the
// scalar-base + 32-bit-offset form and no 64-bit address arithmetic per lane.
// The host guarantees n_nodes * sizeof(S) < 2^32.
template <typename S> __device__ __forceinline__ const S& elem(const S* p, uint32_t i)
{
    return *(const S*)((const char*)p + i * (uint32_t)sizeof(S));
}
template <typename S> __device__ __forceinline__ S& elem(S* p, uint32_t i)
{
    return *(S*)((char*)p + i * (uint32_t)sizeof(S));
}
template <typename S, typename T> struct Storage;
template <> struct Storage<double, double> {
    static __device__ __forceinline__ double ld(const double* p, uint32_t i, float, float = 0.f) { return elem(p, i); }
    static __device__ __forceinline__ void st(double* p, uint32_t i, double v, float, float = 0.f) { elem(p, i) = v; }
    static __device__ __forceinline__ double enc(double v, float) { return v; }
    static __device__ __forceinline__ double dec(double h, float) { return h; }
};
template <> struct Storage<float, float> {
    static __device__ __forceinline__ float ld(const float* p, uint32_t i, float, float = 0.f) { return elem(p, i); }
    static __device__ __forceinline__ void st(float* p, uint32_t i, float v, float, float = 0.f) { elem(p, i) = v; }
    static __device__ __forceinline__ float enc(float v, float) { return v; }
    static __device__ __forceinline__ float dec(float h, float) { return h; }
};
// sh: offset of the stored deviation (the reference-velocity equilibrium, in
// f - w units); st takes ss = -sh * s, so both are one mixed-precision FMA
template <> struct Storage<_Float16, float> {
    // the stored half of v and the value a load gives back (one expression
    // each, shared by the memory accesses and the fused boundary conditions);
    // explicit FMAs: a contractible multiply-add would be fused in one kernel
    // and rounded twice in another, depending on the surrounding code
    static __device__ __forceinline__ _Float16 enc2(float v, float s, float ss)
    {
        return (_Float16)__builtin_fmaf(v, s, ss);
    }
    static __device__ __forceinline__ float dec2(_Float16 h, float s, float sh) { return __builtin_fmaf((float)h, s, sh); }
    static __device__ __forceinline__ float ld(const _Float16* p, uint32_t i, float s, float sh = 0.f)
    {
        return dec2(elem(p, i), s, sh);
    }
    static __device__ __forceinline__ void st(_Float16* p, uint32_t i, float v, float s, float ss = 0.f)
    {
        elem(p, i) = enc2(v, s, ss);
    }
    static __device__ __forceinline__ _Float16 enc(float v, float s) { return (_Float16)(v * s); }
    static __device__ __forceinline__ float dec(_Float16 h, float s) { return (float)h * s; }
};
// Raw buffer stores of one element type: descriptor = base address and size
// in bytes (dword 3 = 0x00020000, the gfx9-family raw-buffer format), VGPR
// byte offset + SGPR byte offset.  Used where a lane chooses between planes:
// the lane offset stays one 32-bit VGPR and the plane offset an SGPR.
typedef unsigned int vfx_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}
template <typename S> struct Buf;
template <> struct Buf<double> {
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, double v)
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vfx_u32x2, v), r, vo, so, 0);
    }
};
template <> struct Buf<float> {
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, vo, so, 0);
    }
};
template <> struct Buf<_Float16> {
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, _Float16 v)
    {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, vo, so, 0);
    }
};

template <typename S> struct Compute { using T = float; };
template <> struct Compute<double> { using T = double; };

// Device view of a vfx_lattice.
struct Lat {
    void* f;
    uint64_t N;
    uint32_t nx, ny, nz, chunks; // direct addressing; chunks = blocks per x row
    const uint32_t* solid;
    const uint32_t* nbx;
    const uint32_t* nby;
    const uint32_t* nbz;
    const uint32_t* fluid;
    uint32_t nfluid;
    const uint32_t* types; // sparse: typeOfGridNode (nodes not GEO_SOLID/GEO_VOID collide), or NULL
    float ld[4], st[4];
    float sh[27], ss[27]; // FP16 reference-velocity offsets per direction (0 otherwise)
    void* rho;
    void* vx;
    void* vy;
    void* vz;
    int32_t write_macro;
    void* nut;             // turbulent viscosity per node (T), eddy-viscosity models
    const uint32_t* nbinv; // sparse: the node in -x-y-z (AMD model)
    void* bf[3];           // per-node body force (T), ApplyBodyForce
    double gforce[3];      // global force, unscaled
    double force_factor;
    uint32_t xrun_lg;      // 1 + log2 of the XCD tile run of the direct step (0: xcd_run<S>())
    uint32_t tile;         // nodes per workgroup along x of the direct step
    const uint4* rgroups;  // sparse: per 64 work items {base node, y stride, z stride, regular}, or NULL
};

template <typename S> __device__ __forceinline__ S* plane(const Lat& L, int p) { return (S*)L.f + (uint64_t)p * L.N; }

__device__ __forceinline__ bool is_solid(const Lat& L, uint32_t k)
{
    return L.solid != nullptr && ((L.solid[k >> 5] >> (k & 31u)) & 1u);
}

// VirtualFluids node types (src/gpu/core/Calculation/Calculation.h:43-45)
constexpr uint32_t kGeoSolid = 15, kGeoVoid = 16, kGeoFluid = 19;

// sparse layout: the node of work item i (the i-th fluid node, or node i when
// typeOfGridNode selects), false if it does not collide
__device__ __forceinline__ bool sparse_node(const Lat& L, uint64_t i, uint32_t& k)
{
    if (L.types) {
        if (i >= L.N) return false;
        k = (uint32_t)i;
        const uint32_t t = L.types[k];
        return t != kGeoSolid && t != kGeoVoid;
    }
    if (i >= L.nfluid) return false;
    k = L.fluid[i];
    return k < L.N;
}
// work items of a node loop
__host__ __device__ inline uint64_t node_items(const Lat& L, int direct)
{
    return direct || L.types ? L.N : L.nfluid;
}

// The eight nodes a node's populations live on under the esoteric twist:
// kk[b], bit0 = +x, bit1 = +y, bit2 = +z (VirtualFluids ListIndices,
// KernelUtilities.h:221-245).
__device__ __forceinline__ void nbrs_direct_xyz(const Lat& L, uint32_t x, uint32_t y, uint32_t z, uint32_t (&kk)[8])
{
    const uint32_t row = y + L.ny * z;
    const uint32_t xp = (x + 1 == L.nx) ? 0u : x + 1;
    const uint32_t ryp = (y + 1 == L.ny) ? row - y : row + 1;
    const uint32_t rzp = (z + 1 == L.nz) ? y : row + L.ny;
    const uint32_t ryzp = (y + 1 == L.ny) ? rzp - y : rzp + 1;
    kk[0] = row * L.nx + x;
    kk[1] = row * L.nx + xp;
    kk[2] = ryp * L.nx + x;
    kk[3] = ryp * L.nx + xp;
    kk[4] = rzp * L.nx + x;
    kk[5] = rzp * L.nx + xp;
    kk[6] = ryzp * L.nx + x;
    kk[7] = ryzp * L.nx + xp;
}
// interior node (no index wraps): one 24-bit multiply-add chain, then adds
// of wave-uniform strides
__device__ __forceinline__ void nbrs_interior(const Lat& L, uint32_t x, uint32_t y, uint32_t z, uint32_t (&kk)[8])
{
    const uint32_t sy = L.nx, sz = L.nx * L.ny;
    kk[0] = __umul24(__umul24(z, L.ny) + y, L.nx) + x;
    kk[1] = kk[0] + 1;
    kk[2] = kk[0] + sy;
    kk[3] = kk[2] + 1;
    kk[4] = kk[0] + sz;
    kk[5] = kk[4] + 1;
    kk[6] = kk[4] + sy;
    kk[7] = kk[6] + 1;
}
__device__ __forceinline__ void nbrs_direct(const Lat& L, uint32_t k, uint32_t (&kk)[8])
{
    const uint32_t x = k % L.nx;
    const uint32_t r = k / L.nx;
    nbrs_direct_xyz(L, x, r % L.ny, r / L.ny, kk);
}
__device__ __forceinline__ void nbrs_sparse(const Lat& L, uint32_t k, uint32_t (&kk)[8])
{
    kk[0] = k;
    kk[1] = L.nbx[k];
    kk[2] = L.nby[k];
    kk[4] = L.nbz[k];
    kk[3] = L.nby[kk[1]];
    kk[5] = L.nbz[kk[1]];
    kk[6] = L.nbz[kk[2]];
    kk[7] = L.nbz[kk[3]];
}

// A wave-uniform float held in a VGPR.  FP16 loads and stores are one
// mixed-precision FMA with two lattice constants (scale and offset); a VALU
// instruction reads at most one SGPR, so the four per-class scales live in
// VGPRs (four moves per thread instead of one per access).
__device__ __forceinline__ float in_vgpr(float x)
{
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
// OUTER2: the collision hands over 2^c times the populations of weight class
// c (k17_fast<true>), so the store scales carry 2^-c (exact)
template <typename S, bool OUTER2 = false> struct Scales {
    float ld[4], st[4];
    __device__ __forceinline__ explicit Scales(const Lat& L)
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            ld[c] = sizeof(S) == 2 ? in_vgpr(L.ld[c]) : L.ld[c];
            st[c] = sizeof(S) != 2 ? L.st[c] : (OUTER2 && c > 0) ? L.st[c] * (1.f / float(1 << c)) : in_vgpr(L.st[c]);
        }
    }
};

// pre-collision populations of a node in cube order
template <typename S, typename T, bool EVEN>
__device__ __forceinline__ void gather(const Lat& L, const uint32_t (&kk)[8], T (&m)[27])
{
    const Scales<S> sc(L);
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        const int d = dir_at(s);
        m[s] = Storage<S, T>::ld(plane<S>(L, slot(d, EVEN)), kk[nbr(d)], sc.ld[weight_class(d)], L.sh[d]);
    }
}
// post-collision populations: the element direction d was read from receives
// the population of the opposite direction
template <typename S, typename T, bool EVEN, bool OUTER2 = false>
__device__ __forceinline__ void scatter_post(const Lat& L, const uint32_t (&kk)[8], const T (&m)[27])
{
    const Scales<S, OUTER2> sc(L);
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        const int d = dir_at(s);
        const int e = inv(d);
        Storage<S, T>::st(plane<S>(L, slot(d, EVEN)), kk[nbr(d)], m[slot_of(e)], sc.st[weight_class(e)], L.ss[e]);
    }
}
// pre-collision write (initialisation, population upload)
template <typename S, typename T, bool EVEN>
__device__ __forceinline__ void scatter_pre(const Lat& L, const uint32_t (&kk)[8], const T (&m)[27])
{
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        const int d = dir_at(s);
        Storage<S, T>::st(plane<S>(L, slot(d, EVEN)), kk[nbr(d)], m[s], L.st[weight_class(d)], L.ss[d]);
    }
}

// COLL: VFX_K17 / VFX_BGK in the low 4 bits, the eddy-viscosity model
// (TurbModel) of the reference-order K17 above them; FAST selects the
// reduced-operation K17 (FP32 only); OUTER2 (with FAST) leaves class-c
// populations scaled by 2^c for a scaling store.  nut: the node's turbulent
// viscosity in, the model's new value out.
constexpr int kCollBodyForce = 256;
constexpr int coll_op(int coll) { return coll & 15; }
constexpr int coll_turb(int coll) { return (coll >> 4) & 15; }
constexpr bool coll_body_force(int coll) { return (coll & kCollBodyForce) != 0; }

// runCollision with ApplyBodyForce (RunCollision.cuh:88-97): the node's force
// ((forces + bodyForce[k]) / 2) * forceFactor, then bodyForce[k] = 0
template <typename T, int COLL>
__device__ __forceinline__ void node_force(const Lat& L, uint32_t k, const Rates<T>& R, Rates<T>& Rn)
{
    if constexpr (coll_body_force(COLL)) {
        Rn = R;
        T* bx = (T*)L.bf[0];
        T* by = (T*)L.bf[1];
        T* bz = (T*)L.bf[2];
        const T ff = (T)L.force_factor;
        Rn.fx = ((T)L.gforce[0] + bx[k]) * T(0.5) * ff;
        Rn.fy = ((T)L.gforce[1] + by[k]) * T(0.5) * ff;
        Rn.fz = ((T)L.gforce[2] + bz[k]) * T(0.5) * ff;
        bx[k] = T(0);
        by[k] = T(0);
        bz[k] = T(0);
    }
}
template <typename T, int COLL, bool FAST, bool OUTER2 = false>
__device__ __forceinline__ void collide(T (&m)[27], const Rates<T>& R, T (&mv)[4], T& nut)
{
    if constexpr (coll_op(COLL) == VFX_K17) {
        if constexpr (FAST && sizeof(T) == 4 && coll_turb(COLL) == kTurbNone) k17_fast<OUTER2>(m, R, mv);
        else k17<T, coll_turb(COLL)>(m, R, mv, nut);
    } else {
        bgk<T>(m, R.omega, mv);
    }
}
// runCollision's handling of the turbulent viscosity (RunCollision.cuh:116-141):
// read before the collision, stored after it; the AMD model needs the
// velocities of every node
template <typename T, int COLL> __device__ __forceinline__ T load_nut(const Lat& L, uint32_t k)
{
    if constexpr (coll_turb(COLL) != kTurbNone) return ((const T*)L.nut)[k];
    return T(0);
}
template <typename T, int COLL> __device__ __forceinline__ void store_nut(const Lat& L, uint32_t k, T nut)
{
    if constexpr (coll_turb(COLL) == kTurbSmagorinsky || coll_turb(COLL) == kTurbQR) ((T*)L.nut)[k] = nut;
}

template <typename T> __device__ __forceinline__ void write_macro(const Lat& L, uint32_t k, const T (&mv)[4])
{
    ((T*)L.rho)[k] = mv[0];
    ((T*)L.vx)[k] = mv[1];
    ((T*)L.vy)[k] = mv[2];
    ((T*)L.vz)[k] = mv[3];
}

// ------------------------------------------------------ collide + stream ----
// The command processor hands the workgroups of a launch to the 8 XCDs round
// robin (workgroup b to XCD b mod 8), and every XCD has its own L2.  Blocks
// next to each other in the lattice share cache lines at their ends (the
// +1-shifted reads of the esoteric twist, rows that do not start on a line), so
// the logical block of hardware workgroup b is chosen to give every XCD runs of
// C consecutive logical blocks: groups of 8C logical blocks, XCD x taking
// blocks x*C .. x*C+C-1 of each group (a tail shorter than 8C keeps the plain
// order).  Shared lines inside a run meet in one L2.  Runs much longer than
// this spread the XCDs over distant parts of each plane at once, which costs
// more than it saves for the 2-byte form (measured on MI355X: 16 blocks for
// 2- and 4-byte populations, 256 for 8-byte ones).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G, uint32_t C)
{
    const uint32_t span = kXcds * C, full = G / span * span;
    return b >= full ? b : (b / span) * span + (b % kXcds) * C + (b / kXcds) % C;
}
template <typename S> constexpr uint32_t xcd_run() { return sizeof(S) == 8 ? 256u : 16u; }
// the same for a run of 2^lg tiles (shifts only)
__device__ __forceinline__ uint32_t xcd_block_pow2(uint32_t b, uint32_t G, uint32_t lg)
{
    const uint32_t sh = lg + 3u, full = (G >> sh) << sh;
    return b >= full ? b : ((b >> sh) << sh) + ((b & (kXcds - 1u)) << lg) + ((b >> 3) & ((1u << lg) - 1u));
}

// Direct (structured) addressing: one block covers `blockDim.x` consecutive
// nodes of one x row, so every row index below is wave-uniform (scalar ALU).
// Block `block` of the lattice's `blocks`.
template <typename S, int COLL, bool EVEN, bool FAST>
__device__ __forceinline__ void step_direct_body(const Lat& L, const Rates<typename Compute<S>::T>& R, uint32_t block,
                                                 uint32_t blocks)
{
    using T = typename Compute<S>::T;
    const uint32_t blk = L.xrun_lg ? xcd_block_pow2(block, blocks, L.xrun_lg - 1u) : xcd_block(block, blocks, xcd_run<S>());
    const uint32_t row = blk / L.chunks;
    const uint32_t x = (blk - row * L.chunks) * blockDim.x + threadIdx.x;
    if (x >= L.nx) return;
    const uint32_t y = row % L.ny, z = row / L.ny;
    uint32_t kk[8];
    nbrs_direct_xyz(L, x, y, z, kk);
    if (is_solid(L, kk[0])) return;
    T m[27], mv[4];
    constexpr bool O2 = FAST && coll_op(COLL) == VFX_K17 && coll_turb(COLL) == kTurbNone && sizeof(S) == 2;
    T nut = load_nut<T, COLL>(L, kk[0]);
    Rates<T> Rn;
    node_force<T, COLL>(L, kk[0], R, Rn);
    const Rates<T>& Ru = coll_body_force(COLL) ? Rn : R;
    gather<S, T, EVEN>(L, kk, m);
    collide<T, COLL, FAST, O2>(m, Ru, mv, nut);
    scatter_post<S, T, EVEN, O2>(L, kk, m);
    if (L.write_macro) write_macro<T>(L, kk[0], mv);
    store_nut<T, COLL>(L, kk[0], nut);
}
template <typename S, int COLL, bool EVEN, bool FAST>
__global__ void __launch_bounds__(256) k_step_direct(Lat L, Rates<typename Compute<S>::T> R)
{
    step_direct_body<S, COLL, EVEN, FAST>(L, R, blockIdx.x, gridDim.x);
}

// The eight nodes of sparse node k (as nbrs_sparse) with the second- and
// third-level neighbour reads shared across the wavefront.  VirtualFluids
// numbers a grid x-fastest, so the +x neighbour of a lane's node is mostly the
// next lane's node: where it is, nby/nbz of the +x neighbour are the next
// lane's first-level reads and nbz of the +x+y neighbour its +y neighbour's
// +z read (cross-lane shuffles instead of two dependent rounds of loads);
// the last lane reads the next node's arrays itself, and a lane whose +x
// neighbour is not the next node falls back to the loads.  `node` is the
// lane's node (0, the dummy node, for a lane without one), `next` the node of
// work item i + 1.  Same indices as nbrs_sparse.
__device__ __forceinline__ void nbrs_sparse_wave(const Lat& L, uint32_t node, uint32_t next, uint32_t (&kk)[8])
{
    const bool last = (threadIdx.x & 63u) == 63u;
    kk[0] = node;
    kk[1] = L.nbx[node];
    kk[2] = L.nby[node];
    kk[4] = L.nbz[node];
    uint32_t kn = __shfl_down(node, 1, 64), yn = __shfl_down(kk[2], 1, 64), zn = __shfl_down(kk[4], 1, 64);
    if (last) {
        kn = next;
        yn = L.nby[next];
        zn = L.nbz[next];
    }
    const bool hit = kk[1] == kn;
    kk[3] = yn;
    kk[5] = zn;
    if (!hit) {
        kk[3] = L.nby[kk[1]];
        kk[5] = L.nbz[kk[1]];
    }
    kk[6] = L.nbz[kk[2]];
    // the +x+y node's +z neighbour: the next lane's +y node's +z read
    const uint32_t z6n = __shfl_down(kk[6], 1, 64);
    kk[7] = z6n;
    if (!hit || last) kk[7] = L.nbz[kk[3]];
}

// Regular wavefronts of the sparse layout.  Where a grid is numbered x
// fastest (as VirtualFluids' generator numbers it), the 64 work items of a
// wavefront are mostly 64 consecutive nodes whose eight twist nodes are
// node + dx (j & 1) + sy ((j >> 1) & 1) + sz (j >> 2) with dx = 1 and one y
// and one z stride; at the end of a periodic row the last lane's +x neighbour
// is the row's first node, dx = 1 - sy (flag bit 1).  k_sparse_groups marks
// those wavefronts (every colliding lane checked against its index arrays,
// every formed index in bounds); the step forms their indices arithmetically
// and reads no index array.  Same indices, so the same populations, as the
// loads.
__device__ __forceinline__ void regular_nbrs(uint32_t k, uint32_t dx, uint32_t sy, uint32_t sz, uint32_t (&kk)[8])
{
#pragma unroll
    for (int j = 0; j < 8; ++j) kk[j] = k + dx * (j & 1) + sy * ((j >> 1) & 1) + sz * (j >> 2);
}
template <int = 0> __global__ void __launch_bounds__(256) k_sparse_groups(Lat L, uint4* table, uint32_t ngroups)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t k = 0;
    const bool active = sparse_node(L, i, k);
    const bool has = L.types ? i < L.N : active;
    const uint32_t node = L.types ? (i < L.N ? (uint32_t)i : 0u) : (active ? k : 0u);
    const uint32_t k0 = __shfl(node, 0, 64);
    uint32_t sy = 0, sz = 0;
    if (lane == 0 && has) {
        sy = L.nby[node] - node;
        sz = L.nbz[node] - node;
    }
    sy = __shfl(sy, 0, 64);
    sz = __shfl(sz, 0, 64);
    bool ok = has && node == k0 + lane, okw = ok; // okw: the last lane wrapping in x
    if (ok && active) {
        uint32_t kk[8], r[8], rw[8];
        nbrs_sparse(L, node, kk);
        regular_nbrs(node, 1u, sy, sz, r);
        regular_nbrs(node, lane == 63u ? 1u - sy : 1u, sy, sz, rw);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            ok = ok && kk[j] == r[j];
            okw = okw && kk[j] == rw[j];
        }
    }
    const bool inside = (uint64_t)k0 + 64u + (uint64_t)sy + (uint64_t)sz < L.N && sy < L.N && sz < L.N;
    const bool all = __ballot(ok) == ~0ull, allw = __ballot(okw) == ~0ull;
    const uint64_t w = i >> 6;
    if (lane == 0 && w < ngroups) table[w] = make_uint4(k0, sy, sz, !inside ? 0u : all ? 1u : allw ? 3u : 0u);
}

// VirtualFluids sparse addressing over the fluid-node list
// (RunCollision.cuh:74-143 semantics).
template <typename S, int COLL, bool EVEN, bool FAST>
__global__ void __launch_bounds__(256) k_step_sparse(Lat L, Rates<typename Compute<S>::T> R)
{
    using T = typename Compute<S>::T;
    const uint32_t blk = L.xrun_lg ? xcd_block_pow2(blockIdx.x, gridDim.x, L.xrun_lg - 1u)
                                   : xcd_block(blockIdx.x, gridDim.x, xcd_run<S>());
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    uint32_t kk[8], k;
    bool active;
    uint4 g = make_uint4(0u, 0u, 0u, 0u);
    if (L.rgroups) g = L.rgroups[__builtin_amdgcn_readfirstlane(i >> 6)];
    if (g.w & 1u) {
        // a regular wavefront: its nodes and their twist nodes by arithmetic
        const uint32_t lane = threadIdx.x & 63u;
        k = g.x + lane;
        active = true;
        if (L.types) {
            const uint32_t t = L.types[k];
            active = t != kGeoSolid && t != kGeoVoid;
        }
        regular_nbrs(k, (g.w & 2u) && lane == 63u ? 1u - g.y : 1u, g.y, g.z, kk);
    } else {
        active = sparse_node(L, i, k);
        // the node of work item i + 1 (the last lane of a wavefront reads its arrays)
        uint32_t next = 0;
        if ((threadIdx.x & 63u) == 63u) {
            if (L.types) next = i + 1 < L.N ? i + 1 : 0u;
            else if (i + 1 < L.nfluid) next = L.fluid[i + 1] < L.N ? L.fluid[i + 1] : 0u;
        }
        // every lane takes part in the shuffles: lanes without a node use the dummy node
        const uint32_t node = L.types ? (i < L.N ? i : 0u) : (active ? k : 0u);
        nbrs_sparse_wave(L, node, next, kk);
    }
    if (!active) return;
    T m[27], mv[4];
    constexpr bool O2 = FAST && coll_op(COLL) == VFX_K17 && coll_turb(COLL) == kTurbNone && sizeof(S) == 2;
    T nut = load_nut<T, COLL>(L, kk[0]);
    Rates<T> Rn;
    node_force<T, COLL>(L, kk[0], R, Rn);
    const Rates<T>& Ru = coll_body_force(COLL) ? Rn : R;
    gather<S, T, EVEN>(L, kk, m);
    collide<T, COLL, FAST, O2>(m, Ru, mv, nut);
    scatter_post<S, T, EVEN, O2>(L, kk, m);
    if (L.write_macro) write_macro<T>(L, kk[0], mv);
    store_nut<T, COLL>(L, kk[0], nut);
}

// ------------------------------------------------------------ node loops ----
// Visits every node that collides: direct -> every non-solid node of the box,
// sparse -> the fluid-node list.
struct NodeIter {
    __device__ static bool get(const Lat& L, uint64_t i, uint32_t (&kk)[8], int direct)
    {
        if (direct) {
            if (i >= L.N) return false;
            const uint32_t k = (uint32_t)i;
            if (is_solid(L, k)) return false;
            nbrs_direct(L, k, kk);
        } else {
            uint32_t k;
            if (!sparse_node(L, i, k)) return false;
            nbrs_sparse(L, k, kk);
        }
        return true;
    }
};

// shifted compressible equilibrium (InitNavierStokesCompressible_Device.cu:170-210)
template <typename T>
__device__ __forceinline__ void equilibrium(T (&m)[27], T drho, T vx1, T vx2, T vx3)
{
    const T cu_sq = T(1.5) * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
    m[slot_of(d000)] = T(8. / 27.) * (drho - cu_sq * (T(1) + drho));
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        T cu;
        if (cx(d) != 0) {
            cu = cx(d) > 0 ? vx1 : -vx1;
            if (cy(d) != 0) cu = cy(d) > 0 ? cu + vx2 : cu - vx2;
            if (cz(d) != 0) cu = cz(d) > 0 ? cu + vx3 : cu - vx3;
        } else if (cy(d) != 0) {
            cu = cy(d) > 0 ? vx2 : -vx2;
            if (cz(d) != 0) cu = cz(d) > 0 ? cu + vx3 : cu - vx3;
        } else {
            cu = cz(d) > 0 ? vx3 : -vx3;
        }
        m[slot_of(d)] = weight<T>(d) * (drho + (T(1) + drho) * (T(3) * cu + T(4.5) * cu * cu - cu_sq));
    }
}

template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_init(Lat L, int direct, const double* drho, const double* ux, const double* uy,
                                              const double* uz)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kk[8];
    if (!NodeIter::get(L, i, kk, direct)) return;
    const uint32_t k = kk[0];
    T m[27];
    equilibrium<T>(m, (T)drho[k], (T)ux[k], (T)uy[k], (T)uz[k]);
    scatter_pre<S, T, EVEN>(L, kk, m);
}

template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_read(Lat L, int direct, double* out)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kk[8];
    if (!NodeIter::get(L, i, kk, direct)) return;
    T m[27];
    gather<S, T, EVEN>(L, kk, m);
#pragma unroll
    for (int s = 0; s < 27; ++s) out[(uint64_t)kk[0] * 27 + dir_at(s)] = (double)m[s];
}

template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_write(Lat L, int direct, const double* in)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kk[8];
    if (!NodeIter::get(L, i, kk, direct)) return;
    T m[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) m[s] = (T)in[(uint64_t)kk[0] * 27 + dir_at(s)];
    scatter_pre<S, T, EVEN>(L, kk, m);
}

// FP64 density deviation and momentum of one node (pyramid sums of
// MacroscopicQuantities.h:55-85 on the populations widened to double)
template <typename T>
__device__ __forceinline__ void node_moments(const T (&m)[27], double& r, double& jx, double& jy, double& jz)
{
    double f[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) f[dir_at(s)] = (double)m[s];
    const double corners = ((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]));
    const double edges = (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P]))) +
                         ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]));
    const double faces = ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0])) + (f[d00M] + f[d00P]);
    r = ((corners + edges) + f[d000]) + faces;
    jx = (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
         (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
         (f[dP00] - f[dM00]);
    jy = (((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
         (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
         (f[d0P0] - f[d0M0]);
    jz = (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
         (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
         (f[d00P] - f[d00M]);
}

// macroscopic values of the pre-collision state (FP64 output)
template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_macro(Lat L, int direct, double* drho, double* vx, double* vy, double* vz)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kk[8];
    if (!NodeIter::get(L, i, kk, direct)) return;
    T m[27];
    gather<S, T, EVEN>(L, kk, m);
    double r, jx, jy, jz;
    node_moments<T>(m, r, jx, jy, jz);
    const double oor = 1.0 / (r + 1.0);
    const uint32_t k = kk[0];
    drho[k] = r;
    vx[k] = jx * oor;
    vy[k] = jy * oor;
    vz[k] = jz * oor;
}

// calculateMacroscopicQuantitiesCompressible_device
// (PostProcessor/MacroscopicQuantities.cu:194-237): every node's outputs are
// zeroed, then the fluid nodes get vf::lbm::getDensity, the compressible
// velocities (numerator / (rho + 1)) and getPressure (MacroscopicQuantities.h:
// 55-64, 100-115, 139-149) in the lattice's arithmetic type, in the
// reference's operation order.  LIST = 0 walks every node (direct: not solid;
// sparse with typeOfGridNode: isValidFluidNode, KernelUtilities.h:216-219, types
// 19, 5, 6, 7); LIST = 1 computes the nodes of the fluid-node list after a
// LIST = 0 pass has zeroed everything.
template <typename T> __device__ __forceinline__ void vf_macroscopic(const T (&f)[27], T& rho, T& vx, T& vy, T& vz,
                                                                     T& press)
{
    rho = ((((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]))) +
           (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P])) +
            ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]))) +
           f[d000]) +
          ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0]) + (f[d00M] + f[d00P]));
    const T jx = ((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
                  (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
                  (f[dP00] - f[dM00]));
    const T jy = ((((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
                  (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
                  (f[d0P0] - f[d0M0]));
    const T jz = ((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
                  (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
                  (f[d00P] - f[d00M]));
    vx = jx / (rho + T(1));
    vy = jy / (rho + T(1));
    vz = jz / (rho + T(1));
    press = (f[dP00] + f[dM00] + f[d0P0] + f[d0M0] + f[d00P] + f[d00M] +
             T(2) * (f[dPP0] + f[dMM0] + f[dPM0] + f[dMP0] + f[dP0P] + f[dM0M] + f[dP0M] + f[dM0P] + f[d0PP] +
                     f[d0MM] + f[d0PM] + f[d0MP]) +
             T(3) * (f[dPPP] + f[dMMP] + f[dPMP] + f[dMPP] + f[dPPM] + f[dMMM] + f[dPMM] + f[dMPM]) - rho -
             (vx * vx + vy * vy + vz * vz) * (T(1) + rho)) *
                T(0.5) +
            rho;
}

template <typename S, bool EVEN, int LIST>
__global__ void __launch_bounds__(256) k_macro_vf(Lat L, int direct, typename Compute<S>::T* rho,
                                                  typename Compute<S>::T* vx, typename Compute<S>::T* vy,
                                                  typename Compute<S>::T* vz, typename Compute<S>::T* press)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k;
    bool fluid;
    if constexpr (LIST) {
        if (i >= L.nfluid) return;
        k = L.fluid[i];
        if (k >= L.N) return;
        fluid = true;
    } else {
        if (i >= L.N) return;
        k = (uint32_t)i;
        press[k] = T(0);
        rho[k] = T(0);
        vx[k] = T(0);
        vy[k] = T(0);
        vz[k] = T(0);
        if (direct) {
            fluid = !is_solid(L, k);
        } else if (L.types) {
            const uint32_t t = L.types[k];
            fluid = t == 19u || t == 5u || t == 6u || t == 7u;
        } else {
            fluid = false;
        }
    }
    if (!fluid) return;
    uint32_t kk[8];
    if (direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    T m[27];
    gather<S, T, EVEN>(L, kk, m);
    T f[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) f[dir_at(s)] = m[s];
    T r, ux, uy, uz, p;
    vf_macroscopic<T>(f, r, ux, uy, uz, p);
    rho[k] = r;
    vx[k] = ux;
    vy[k] = uy;
    vz[k] = uz;
    press[k] = p;
}

// Time-averaged fields (calculateMeanCompressible_device,
// MacroscopicQuantities.cu:390-536): adds the node's density deviation and
// velocity to running sums (sum order of that kernel), and to the pressure sum
// the pressure expression evaluated with the new density and velocity sums in
// place of the node's own values, as the reference kernel does.  Every node
// of the all-node form zeroes its sums first, so only fluid nodes accumulate.
template <typename T> __device__ __forceinline__ void vf_mean_add(const T (&f)[27], T& RHO, T& VX, T& VY, T& VZ, T& PRESS)
{
#pragma clang fp contract(off)
    const T drho = ((((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]))) +
                    (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P])) +
                     ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]))) +
                    ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0]) + (f[d00M] + f[d00P]))) +
                   f[d000];
    const T rho = T(1) + drho;
    const T r = drho + RHO;
    const T x = (((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
                  (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
                  (f[dP00] - f[dM00])) /
                 rho) +
                VX;
    const T y = (((((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
                  (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
                  (f[d0P0] - f[d0M0])) /
                 rho) +
                VY;
    const T z = (((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
                  (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
                  (f[d00P] - f[d00M])) /
                 rho) +
                VZ;
    PRESS = (f[dP00] + f[dM00] + f[d0P0] + f[d0M0] + f[d00P] + f[d00M] +
             T(2) * (f[dPP0] + f[dMM0] + f[dPM0] + f[dMP0] + f[dP0P] + f[dM0M] + f[dP0M] + f[dM0P] + f[d0PP] +
                     f[d0MM] + f[d0PM] + f[d0MP]) +
             T(3) * (f[dPPP] + f[dMMP] + f[dPMP] + f[dMPP] + f[dPPM] + f[dMMM] + f[dPMM] + f[dMPM]) - r -
             (x * x + y * y + z * z) * (T(1) + r)) *
                T(0.5) +
            r + PRESS;
    RHO = r;
    VX = x;
    VY = y;
    VZ = z;
}

template <typename S, bool EVEN, int LIST>
__global__ void __launch_bounds__(256) k_mean_vf(Lat L, int direct, typename Compute<S>::T* vx,
                                                 typename Compute<S>::T* vy, typename Compute<S>::T* vz,
                                                 typename Compute<S>::T* rho, typename Compute<S>::T* press)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k;
    bool fluid;
    if constexpr (LIST) {
        if (i >= L.nfluid) return;
        k = L.fluid[i];
        if (k >= L.N) return;
        fluid = true;
    } else {
        if (i >= L.N) return;
        k = (uint32_t)i;
        fluid = direct ? !is_solid(L, k) : (L.types && L.types[k] == kGeoFluid);
    }
    T P = press[k], R = rho[k], X = vx[k], Y = vy[k], Z = vz[k];
    if (!LIST) {
        press[k] = T(0);
        rho[k] = T(0);
        vx[k] = T(0);
        vy[k] = T(0);
        vz[k] = T(0);
    }
    if (!fluid) return;
    uint32_t kk[8];
    if (direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    T m[27];
    gather<S, T, EVEN>(L, kk, m);
    T f[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) f[dir_at(s)] = m[s];
    vf_mean_add<T>(f, R, X, Y, Z, P);
    rho[k] = R;
    vx[k] = X;
    vy[k] = Y;
    vz[k] = Z;
    press[k] = P;
}

template <typename T> __global__ void __launch_bounds__(256) k_zero5(T* a, T* b, T* c, T* d, T* e, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = T(0);
    b[i] = T(0);
    c[i] = T(0);
    d[i] = T(0);
    e[i] = T(0);
}

// ---------------------------------------------------------- reductions ----
// Deterministic two-pass FP64 sums: every block sums a fixed chunk of
// kRedBlock * kRedPerThread nodes (per-thread sequential, then a wave64
// butterfly and a fixed-order sum of the four waves); a single block then
// adds the per-block partials in index order.
constexpr int kRedBlock = 256;
constexpr int kRedPerThread = 8;
constexpr int kRedChunk = kRedBlock * kRedPerThread;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int NV> __device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double part[kRedBlock / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        v[c] = wave_sum(v[c]);
        if (lane == 0) part[wave][c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double t = part[0][c];
            for (int w = 1; w < kRedBlock / 64; ++w) t += part[w][c];
            v[c] = t;
        }
    }
}

// per node: drho, jx, jy, jz (mass deviation and momentum of the pre-collision state)
template <typename S, bool EVEN>
__global__ void __launch_bounds__(kRedBlock) k_reduce_mass(Lat L, int direct, double* partial)
{
    using T = typename Compute<S>::T;
    double acc[4] = {0, 0, 0, 0};
    for (int j = 0; j < kRedPerThread; ++j) {
        const uint64_t i = (uint64_t)blockIdx.x * kRedChunk + (uint64_t)j * kRedBlock + threadIdx.x;
        uint32_t kk[8];
        if (!NodeIter::get(L, i, kk, direct)) continue;
        T m[27];
        gather<S, T, EVEN>(L, kk, m);
        double r, jx, jy, jz;
        node_moments<T>(m, r, jx, jy, jz);
        acc[0] += r;
        acc[1] += jx;
        acc[2] += jy;
        acc[3] += jz;
    }
    block_sum<4>(acc);
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; ++c) partial[(uint64_t)blockIdx.x * 4 + c] = acc[c];
}

template <int NV> __global__ void __launch_bounds__(kRedBlock) k_reduce_finish(const double* partial, uint32_t nblocks, double* out)
{
    double acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += kRedBlock)
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] += partial[(uint64_t)b * NV + c];
    block_sum<NV>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = acc[c];
}

// Structured-box neighbour along one axis (periodic index wrap), -1 if solid.
__device__ __forceinline__ int64_t fluid_nbr(const Lat& L, uint32_t x, uint32_t y, uint32_t z, int ax, int step)
{
    int64_t c[3] = {x, y, z};
    const int64_t n[3] = {L.nx, L.ny, L.nz};
    c[ax] = ((c[ax] + step) % n[ax] + n[ax]) % n[ax];
    const uint32_t k = (uint32_t)(c[0] + n[0] * (c[1] + n[1] * c[2]));
    return is_solid(L, k) ? -1 : (int64_t)k;
}

// Kinetic energy 1/2 rho |u|^2 summed over fluid nodes whose six axis
// neighbours are fluid, and their count (KineticEnergyAnalyzer.cu:137-154).
// Enstrophy 1/2 rho |curl u|^2 on the same nodes with central differences of
// order 8, reduced to 6/4/2 where the stencil meets a wall
// (EnstrophyAnalyzer.cu:134-336).
template <bool ENSTROPHY>
__global__ void __launch_bounds__(kRedBlock) k_reduce_energy(Lat L, const double* rho, const double* vx, const double* vy,
                                                             const double* vz, double* partial)
{
    double acc[2] = {0, 0};
    for (int j = 0; j < kRedPerThread; ++j) {
        const uint64_t i = (uint64_t)blockIdx.x * kRedChunk + (uint64_t)j * kRedBlock + threadIdx.x;
        if (i >= L.N) continue;
        const uint32_t k = (uint32_t)i;
        if (is_solid(L, k)) continue;
        const uint32_t x = k % L.nx, r = k / L.nx, y = r % L.ny, z = r / L.ny;
        bool ok = true;
        for (int ax = 0; ax < 3 && ok; ++ax)
            if (fluid_nbr(L, x, y, z, ax, 1) < 0 || fluid_nbr(L, x, y, z, ax, -1) < 0) ok = false;
        if (!ok) continue;
        if (!ENSTROPHY) {
            acc[0] += 0.5 * (vx[k] * vx[k] + vy[k] * vy[k] + vz[k] * vz[k]) * (rho[k] + 1.0);
            acc[1] += 1.0;
            continue;
        }
        // derivative of field v along axis ax at node k
        auto deriv = [&](const double* v, int ax) {
            int order = 8;
            int64_t p[5], m[5];
            for (int s = 1; s <= 4; ++s) {
                p[s] = fluid_nbr(L, x, y, z, ax, s);
                m[s] = fluid_nbr(L, x, y, z, ax, -s);
            }
            if (p[2] < 0 || m[2] < 0) order = 2;
            if ((p[3] < 0 || m[3] < 0) && order > 4) order = 4;
            if ((p[4] < 0 || m[4] < 0) && order > 6) order = 6;
            auto d = [&](int s) { return v[p[s]] - v[m[s]]; };
            if (order == 2) return d(1) / 2.0;
            if (order == 4) return ((8.0 * v[p[1]] - 8.0 * v[m[1]]) - d(2)) / 12.0;
            if (order == 6) return ((5.0 * 9.0) * d(1) - 9.0 * d(2) + d(3)) / (6.0 * 10.0);
            return ((28.0 * 8.0) * d(1) - (7.0 * 8.0) * d(2) + (8.0 * 4.0 * (1.0 / 3.0)) * d(3) - d(4)) / (7.0 * 10.0 * 4.0);
        };
        const double dyvz = deriv(vz, 1), dzvy = deriv(vy, 2);
        const double dzvx = deriv(vx, 2), dxvz = deriv(vz, 0);
        const double dxvy = deriv(vy, 0), dyvx = deriv(vx, 1);
        const double wx = dyvz - dzvy, wy = dzvx - dxvz, wz = dxvy - dyvx;
        acc[0] += 0.5 * (rho[k] + 1.0) * (wx * wx + wy * wy + wz * wz);
        acc[1] += 1.0;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        partial[(uint64_t)blockIdx.x * 2] = acc[0];
        partial[(uint64_t)blockIdx.x * 2 + 1] = acc[1];
    }
}

// Running FP64 sums for time averages (CalcTurbulenceIntensity,
// TurbulenceIntensity.cu:45-85): planes 0-2 sum u, planes 3-8 sum
// u_x u_x, u_y u_y, u_z u_z, u_x u_y, u_x u_z, u_y u_z.
template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_stats(Lat L, int direct, double* sums)
{
    using T = typename Compute<S>::T;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kk[8];
    if (!NodeIter::get(L, i, kk, direct)) return;
    T m[27];
    gather<S, T, EVEN>(L, kk, m);
    double r, jx, jy, jz;
    node_moments<T>(m, r, jx, jy, jz);
    const double oor = 1.0 / (r + 1.0);
    const double ux = jx * oor, uy = jy * oor, uz = jz * oor;
    const uint64_t k = kk[0], N = L.N;
    sums[k] += ux;
    sums[N + k] += uy;
    sums[2 * N + k] += uz;
    sums[3 * N + k] += ux * ux;
    sums[4 * N + k] += uy * uy;
    sums[5 * N + k] += uz * uz;
    sums[6 * N + k] += ux * uy;
    sums[7 * N + k] += ux * uz;
    sums[8 * N + k] += uy * uz;
}

// -------------------------------------------------- boundary conditions ----
// Every floating-point table of a boundary condition is of the lattice's
// arithmetic type T (double for FP64 storage, float otherwise), the `real` of
// VirtualFluids' QforBoundaryConditions; bc_at<T> reads element i.
struct BCDev {
    uint32_t n;
    const uint32_t* k;
    const void* q;         // q27[0]: 27 * n subgrid distances, SoA
    const void* vx;        // wall velocity per boundary node (NULL = 0)
    const void* vy;
    const void* vz;
    int32_t cpu_float;
    int32_t direct;
    double omega;          // interpolated kinds
    const void* rho;       // pressure: density deviation per boundary node
    const uint32_t* kn;    // pressure / outflow: inward neighbour of each boundary node
    int32_t normal;        // pressure / outflow: outward direction of the boundary (d000..dMMM)
    double corr;           // outflow with pressure correction: density correction factor
    // stress (wall-model) kinds, arrays of the lattice's arithmetic type
    const void* wm_normal[3];     // wall normal per boundary node
    void* wm_el[3];               // filtered velocity at the exchange location (state)
    void* wm_wall[3];             // filtered velocity at the wall node (state)
    const int32_t* wm_offset;     // sampling offset (nodes from the wall)
    const void* wm_z0;            // roughness length
    void* wm_ustar;               // monitors (NULL: none)
    void* wm_force[3];
    // precursor kinds: plane stencil per boundary node and the two plane
    // arrays (lattice's arithmetic type), blended last*(1 - tr) + current*tr
    uint32_t n_pre;
    const uint32_t* pn[4];        // precursor nodes 0PP, 0PM, 0MP, 0MM
    const void* pw[4];            // their inverse-distance weights
    const void* plast;
    const void* pcur;
    double tratio;
    double vratio;                // velocity ratio (velocity kind)
    double pu[3];                 // velocity kind: scalar velocity added to vx/vy/vz
};

template <typename T> __device__ __forceinline__ T bc_at(const void* a, uint64_t i) { return ((const T*)a)[i]; }
template <typename T> __device__ __forceinline__ T bc_q(const BCDev& B, int d, uint32_t i)
{
    return bc_at<T>(B.q, (uint64_t)d * B.n + i);
}

// The plane value at boundary node i: the four precursor nodes around it
// weighted by inverse distance, or the 0PP node alone when its weight is
// >= 1e6 (PrecursorNonReflectiveCompressible.cu:82-110,
// PrecursorDistributions.cu:95-152).  The weighted sum runs in the
// reference's order and is scaled by 1 / (sum of weights).
template <typename T> struct PlaneStencil {
    uint32_t n[4];
    T w[4];
    T inv;
    bool one;
    bool ok; // every precursor node the stencil reads lies inside the planes
};

template <typename T> __device__ __forceinline__ PlaneStencil<T> plane_stencil(const BCDev& B, uint32_t i)
{
    PlaneStencil<T> s;
    s.n[0] = B.pn[0][i];
    s.w[0] = ((const T*)B.pw[0])[i];
    s.one = !(s.w[0] < T(1e6));
    s.inv = T(0);
    s.ok = s.n[0] < B.n_pre;
    if (!s.one) {
        for (int c = 1; c < 4; ++c) {
            s.n[c] = B.pn[c][i];
            s.w[c] = ((const T*)B.pw[c])[i];
            s.ok = s.ok && s.n[c] < B.n_pre;
        }
        s.inv = T(1) / (s.w[0] + s.w[1] + s.w[2] + s.w[3]);
    }
    return s;
}

template <typename T> __device__ __forceinline__ T plane_value(const PlaneStencil<T>& s, const T* a)
{
    if (s.one) return a[s.n[0]];
    return (a[s.n[0]] * s.w[0] + a[s.n[1]] * s.w[1] + a[s.n[2]] * s.w[2] + a[s.n[3]] * s.w[3]) * s.inv;
}

// last*(1 - tr) + current*tr of plane component c (PrecursorDistributions.cu:163-171)
template <typename T> __device__ __forceinline__ T plane_blend(const BCDev& B, const PlaneStencil<T>& s, int c)
{
    const T tr = (T)B.tratio;
    const size_t off = (size_t)c * B.n_pre;
    return plane_value(s, (const T*)B.plast + off) * (T(1) - tr) + plane_value(s, (const T*)B.pcur + off) * tr;
}

// c_d . (ux, uy, uz) spelled as VirtualFluids writes it: first non-zero
// component with its sign, then the others added or subtracted in x, y, z order
template <typename T> __device__ __forceinline__ T proj(int d, T ux, T uy, T uz)
{
    T v;
    if (cx(d) != 0) {
        v = cx(d) > 0 ? ux : -ux;
        if (cy(d) != 0) v = cy(d) > 0 ? v + uy : v - uy;
        if (cz(d) != 0) v = cz(d) > 0 ? v + uz : v - uz;
    } else if (cy(d) != 0) {
        v = cy(d) > 0 ? uy : -uy;
        if (cz(d) != 0) v = cz(d) > 0 ? v + uz : v - uz;
    } else {
        v = cz(d) > 0 ? uz : -uz;
    }
    return v;
}

// 6 w_d (c_d . u) with the component order VirtualFluids writes it
// (VelocityBounceBack.cu:159-184); cpu_float reproduces BoundaryConditions.h:195-262
// (float wall velocity, float component sum, double scale, float result).
template <typename T>
__device__ __forceinline__ T wall_term(int d, double ux, double uy, double uz, int cpu_float)
{
    const int n = weight_class(d);
    const double coef = n == 1 ? 4. / 9. : n == 2 ? 1. / 9. : 1. / 36.;
    if (cpu_float) {
        const float u[3] = {(float)ux, (float)uy, (float)uz};
        const int c[3] = {cx(d), cy(d), cz(d)};
        float s = 0.f;
        bool first = true;
        for (int a = 0; a < 3; ++a) {
            if (c[a] == 0) continue;
            if (first) { s = c[a] > 0 ? u[a] : -u[a]; first = false; }
            else s = c[a] > 0 ? s + u[a] : s - u[a];
        }
        return (T)(double)(float)(coef * (double)s);
    }
    const T u[3] = {(T)ux, (T)uy, (T)uz};
    const int c[3] = {cx(d), cy(d), cz(d)};
    T s = T(0);
    bool first = true;
    for (int a = 0; a < 3; ++a) {
        if (c[a] == 0) continue;
        if (first) { s = c[a] > 0 ? u[a] : -u[a]; first = false; }
        else s = c[a] > 0 ? s + u[a] : s - u[a];
    }
    return (T)coef * s;
}

// Moving-wall halfway bounce-back (VelocityBounceBack_Device): for every wall
// link d of a boundary node, the population leaving towards the wall comes back
// as the opposite direction, minus the wall term.  Runs between the collision
// of parity EVEN and the parity flip.
template <typename S, bool EVEN>
__device__ __forceinline__ void bc_velocity_bb_entry(const Lat& L, const BCDev& B, uint32_t i)
{
    using T = typename Compute<S>::T;
    if (i >= B.n) return;
    uint32_t kk[8];
    const uint32_t k = B.k[i];
    if (k >= L.N) return; // malformed list entry: never touch memory outside the lattice
    if (B.direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    const double ux = B.vx ? (double)bc_at<T>(B.vx, i) : 0.0, uy = B.vy ? (double)bc_at<T>(B.vy, i) : 0.0,
                 uz = B.vz ? (double)bc_at<T>(B.vz, i) : 0.0;
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const T q = bc_q<T>(B, d, i);
        if (!(q >= T(0) && q <= T(1))) continue;
        const int e = inv(d);
        // post-collision population of direction d sits where e was read from
        const T fpost = Storage<S, T>::ld(plane<S>(L, slot(e, EVEN)), kk[nbr(e)], L.ld[weight_class(d)], L.sh[d]);
        const T v = fpost - wall_term<T>(d, ux, uy, uz, B.cpu_float);
        // next step (parity !EVEN) reads direction e of node k from here
        Storage<S, T>::st(plane<S>(L, slot(e, !EVEN)), kk[nbr(e)], v, L.st[weight_class(e)], L.ss[e]);
    }
}

template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_bc_velocity_bb(Lat L, BCDev B)
{
    bc_velocity_bb_entry<S, EVEN>(L, B, blockIdx.x * blockDim.x + threadIdx.x);
}

// Interpolated bounce-back family (NoSlipInterpolatedCompressible.cu:44-378,
// VelocityInterpolatedCompressible.cu:44-…, SlipCompressible.cu:45-…, formulas
// KernelUtilities.h:171-197): from the post-collision populations of the
// boundary node, density and velocity (with the sum order those kernels use),
// the symmetric equilibrium part of each wall link, and the Bouzidi-type
// interpolation with the subgrid distance q.  KIND: 1 no-slip, 2 velocity, 3 slip.
//
// The node part of the interpolated boundaries: density, velocity, rates and
// the boundary velocity of the node's post-collision populations p.
template <typename T> struct InterpNode {
    T drho, vx1, vx2, vx3, cu_sq, omega, inv1mo, VX, VY, VZ;
};
template <typename T, int KIND>
__device__ __forceinline__ InterpNode<T> interp_node(const Lat& L, const BCDev& B, uint32_t i, uint32_t k,
                                                     const T (&p)[27], int kd)
{
    // no contraction into FMAs here: which product the compiler would fuse
    // depends on the code around the inlined body (a shared subexpression
    // changes the choice), and the separate boundary kernel and the fused
    // collide-stream kernel must give the same bits
#pragma clang fp contract(off)
    InterpNode<T> n;
    const T drho = p[dPMP] + p[dMPP] + p[dPPP] + p[dMMP] + p[dPMM] + p[dMPM] + p[dPPM] + p[dMMM] + p[d0PM] + p[d0MP] +
                   p[d0PP] + p[d0MM] + p[dP0M] + p[dM0P] + p[dP0P] + p[dM0M] + p[dPM0] + p[dMP0] + p[dPP0] + p[dMM0] +
                   p[d00P] + p[d00M] + p[d0P0] + p[d0M0] + p[dP00] + p[dM00] + p[d000];
    n.drho = drho;
    // division by 1 + drho: exact in FP64 (reference order), the hardware
    // reciprocal in FP32
    const T ovr = sizeof(T) == 4 ? T(__builtin_amdgcn_rcpf(float(T(1) + drho))) : T(0);
    auto per_rho = [&](T x) -> T {
        if constexpr (sizeof(T) == 4) return x * ovr;
        else return x / (T(1) + drho);
    };
    n.vx1 = per_rho((((p[dPMP] - p[dMPM]) - (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                     ((p[dP0M] - p[dM0P]) + (p[dP0P] - p[dM0M])) + ((p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                     (p[dP00] - p[dM00])));
    n.vx2 = per_rho(((-(p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                     ((p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + (-(p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                     (p[d0P0] - p[d0M0])));
    n.vx3 = per_rho((((p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) + (p[dMMP] - p[dPPM])) +
                     (-(p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + ((p[dP0P] - p[dM0M]) - (p[dP0M] - p[dM0P])) +
                     (p[d00P] - p[d00M])));
    n.cu_sq = T(1.5) * (n.vx1 * n.vx1 + n.vx2 * n.vx2 + n.vx3 * n.vx3) * (T(1) + drho);
    T omega = (T)B.omega;
    // SlipTurbulentViscosityCompressible.cu:130-134: the node's rate with its turbulent viscosity
    if constexpr (KIND == VFX_BC_SLIP_TURBULENT_VISCOSITY)
        omega = omega / (T(1) + T(3) * omega * fmax(T(0), ((const T*)L.nut)[k]));
    n.omega = omega;
    n.inv1mo = sizeof(T) == 4 ? T(1) / (T(1) - omega) : T(0);
    n.VX = n.VY = n.VZ = T(0);
    return n;
}
// the wall velocity of the velocity kinds
template <typename T> __device__ __forceinline__ void interp_wall_velocity(InterpNode<T>& n, const BCDev& B, uint32_t i, int kd)
{
    if (kd == VFX_BC_VELOCITY_INTERPOLATED || kd == VFX_BC_VELOCITY_PRESSURE_INTERPOLATED ||
        kd == VFX_BC_PRECURSOR_VELOCITY) {
        n.VX = B.vx ? bc_at<T>(B.vx, i) : T(0);
        n.VY = B.vy ? bc_at<T>(B.vy, i) : T(0);
        n.VZ = B.vz ? bc_at<T>(B.vz, i) : T(0);
    }
}

// c_d . u with the direction's components given at run time (the order and
// signs of proj())
template <typename T> __device__ __forceinline__ T proj_rt(int ex, int ey, int ez, T ux, T uy, T uz)
{
    T v;
    if (ex != 0) {
        v = ex > 0 ? ux : -ux;
        if (ey != 0) v = ey > 0 ? v + uy : v - uy;
        if (ez != 0) v = ez > 0 ? v + uz : v - uz;
    } else if (ey != 0) {
        v = ey > 0 ? uy : -uy;
        if (ez != 0) v = ez > 0 ? v + uz : v - uz;
    } else {
        v = ez > 0 ? uz : -uz;
    }
    return v;
}

// The population re-entering along a wall link of fraction q (0 <= q <= 1):
// f, fi the node's post-collision populations of the link and its opposite,
// w the link weight, vlb = c . u of the node, vbc = c . u of the wall.
template <typename T>
__device__ __forceinline__ T interp_link_value(const InterpNode<T>& n, int kd, T w, T vlb, T vbc, T q, T f, T fi)
{
#pragma clang fp contract(off)
    const T feq = w * (n.drho + T(4.5) * vlb * vlb * (T(1) + n.drho) - n.cu_sq);
    // FP64 keeps the reference's divisions (bit parity with the CPU
    // restatement); the FP32 arithmetic of the reduced forms uses one
    // hardware reciprocal per link and 1/(1 - omega) once per node
    T a, r1q;
    if constexpr (sizeof(T) == 4) {
        r1q = __builtin_amdgcn_rcpf(T(1) + q);
        a = (T(1) - q) * r1q * (f - fi + (f + fi - T(2) * feq * n.omega) * n.inv1mo) * T(0.5);
    } else {
        r1q = T(0);
        a = (T(1) - q) / (T(1) + q) * (f - fi + (f + fi - T(2) * feq * n.omega) / (T(1) - n.omega)) * T(0.5);
    }
    T v;
    if (kd == VFX_BC_NOSLIP_INTERPOLATED) {
        if constexpr (sizeof(T) == 4) v = a + q * (f + fi) * r1q;
        else v = a + (q * (f + fi)) / (T(1) + q);
    } else {
        if constexpr (sizeof(T) == 4) v = a + (q * (f + fi) - T(6) * w * vbc) * r1q;
        else v = a + (q * (f + fi) - T(6) * w * vbc) / (T(1) + q);
        // getInterpolatedDistributionForVeloWithPressureBC (KernelUtilities.h:207-213):
        // the entering population carries the bulk pressure
        if (kd == VFX_BC_VELOCITY_PRESSURE_INTERPOLATED || kd == VFX_BC_PRECURSOR_VELOCITY) v = v - w * n.drho;
    }
    return v;
}

// The links of boundary entry i at node k from its post-collision
// populations p (indexed by direction): computes and stores the population
// re-entering along every wall link.  KIND >= 0 fixes the kind at compile
// time; KIND = -1 takes it from `kind` at run time (kinds 1, 2, 3, 5).
template <typename S, typename T, bool EVEN, int KIND>
__device__ __forceinline__ void interp_links(const Lat& L, const BCDev& B, uint32_t i, uint32_t k,
                                             const uint32_t (&kk)[8], const T (&p)[27], int kind = KIND)
{
#pragma clang fp contract(off)
    const int kd = KIND >= 0 ? KIND : kind;
    InterpNode<T> n = interp_node<T, KIND>(L, B, i, k, p, kd);
    // all link fractions first: loaded inside the link loop, each q waited for
    // the previous link's store as well (one memory counter covers both) and
    // the 26 links ran as 26 round trips
    T qs[27];
#pragma unroll
    for (int d = 1; d < 27; ++d) qs[d] = bc_q<T>(B, d, i);
    interp_wall_velocity<T>(n, B, i, kd);
    if constexpr (KIND == VFX_BC_PRECURSOR_VELOCITY) {
        // PrecursorNonReflectiveCompressible.cu:112-114: (velocityX + (1 - tr) last + tr current) / velocityRatio
        const PlaneStencil<T> s = plane_stencil<T>(B, i);
        if (!s.ok) return; // malformed plane index: never read outside the planes
        n.VX += (T)B.pu[0];
        n.VY += (T)B.pu[1];
        n.VZ += (T)B.pu[2];
        const T tr = (T)B.tratio, vr = (T)B.vratio;
        const T* last = (const T*)B.plast;
        const T* cur = (const T*)B.pcur;
        const size_t np = B.n_pre;
        n.VX = (n.VX + (T(1) - tr) * plane_value(s, last) + tr * plane_value(s, cur)) / vr;
        n.VY = (n.VY + (T(1) - tr) * plane_value(s, last + np) + tr * plane_value(s, cur + np)) / vr;
        n.VZ = (n.VZ + (T(1) - tr) * plane_value(s, last + 2 * np) + tr * plane_value(s, cur + 2 * np)) / vr;
    }
    bool fx = false, fy = false, fz = false; // slip: axes normal to a face link
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        const T q = qs[d];
        if (!(q >= T(0) && q <= T(1))) continue;
        const T w = weight<T>(d);
        const T vlb = proj<T>(d, n.vx1, n.vx2, n.vx3);
        T f = p[d], fi = p[inv(d)];
        if constexpr (KIND == VFX_BC_PRECURSOR_VELOCITY) {
            // kept from PrecursorNonReflectiveCompressible.cu:383 and :392: the
            // d0PM link takes f_0PP for its opposite, the d0MP link f_0PP for its own
            if (d == d0PM) fi = p[d0PP];
            if (d == d0MP) f = p[d0PP];
        }
        T vbc = T(0);
        if (kd == VFX_BC_SLIP || kd == VFX_BC_SLIP_TURBULENT_VISCOSITY) {
            if (weight_class(d) == 1) {
                if (cx(d) != 0) fx = true;
                if (cy(d) != 0) fy = true;
                if (cz(d) != 0) fz = true;
            }
            const T sx = fx ? T(0) : n.vx1, sy = fy ? T(0) : n.vx2, sz = fz ? T(0) : n.vx3;
            vbc = proj<T>(d, sx, sy, sz);
        } else if (kd != VFX_BC_NOSLIP_INTERPOLATED) {
            vbc = proj<T>(d, n.VX, n.VY, n.VZ);
        }
        const T v = interp_link_value<T>(n, kd, w, vlb, vbc, q, f, fi);
        const int e = inv(d);
        // the precursor inlet writes with the parity's own pointers
        // (PrecursorNonReflectiveCompressible.cu:142): over the outgoing
        // population it read, which the node reads back as e two steps later
        constexpr bool WRITE_EVEN = KIND == VFX_BC_PRECURSOR_VELOCITY ? EVEN : !EVEN;
        Storage<S, T>::st(plane<S>(L, slot(e, WRITE_EVEN)), kk[nbr(e)], v, L.st[weight_class(e)], L.ss[e]);
    }
}

template <typename S, bool EVEN, int KIND>
__device__ __forceinline__ void bc_interp_entry(const Lat& L, const BCDev& B, uint32_t i, int kind = KIND)
{
    using T = typename Compute<S>::T;
    if (i >= B.n) return;
    uint32_t kk[8];
    const uint32_t k = B.k[i];
    if (k >= L.N) return;
    if (B.direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    T p[27]; // post-collision population of direction d
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int e = inv(d);
        p[d] = Storage<S, T>::ld(plane<S>(L, slot(e, EVEN)), kk[nbr(e)], L.ld[weight_class(d)], L.sh[d]);
    }
    interp_links<S, T, EVEN, KIND>(L, B, i, k, kk, p, kind);
}

template <typename S, bool EVEN, int KIND>
__global__ void __launch_bounds__(256) k_bc_interp(Lat L, BCDev B)
{
    bc_interp_entry<S, EVEN, KIND>(L, B, blockIdx.x * blockDim.x + threadIdx.x);
}

// Precursor distributions (PrecursorDistributions.cu:44-185): post-collision,
// parity EVEN.  The nine populations with a positive x component entering
// boundary node k are the blended plane values f0..f8, written for the next
// step like the other post-collision boundary conditions.
template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_bc_precursor_dist(Lat L, BCDev B)
{
    using T = typename Compute<S>::T;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t k = B.k[i];
    if (k >= L.N) return;
    uint32_t kk[8];
    if (B.direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    const PlaneStencil<T> s = plane_stencil<T>(B, i);
    if (!s.ok) return; // malformed plane index: never read outside the planes
    constexpr int dirs[9] = {dP00, dPP0, dPM0, dP0P, dP0M, dPPP, dPMP, dPPM, dPMM};
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const int x = dirs[c];
        const T v = plane_blend(B, s, c);
        Storage<S, T>::st(plane<S>(L, slot(x, !EVEN)), kk[nbr(x)], v, L.st[weight_class(x)], L.ss[x]);
    }
}

// Non-equilibrium pressure boundary (PressureNonEquilibriumCompressible.cu:45-…):
// pre-collision, parity EVEN of the step about to run.  The populations of the
// inward neighbour keep their non-equilibrium part and get the equilibrium of
// the prescribed density; they replace the boundary node's populations of every
// direction not pointing out of the domain (c_d . normal <= 0).
template <typename T> __device__ __forceinline__ T feq_shifted(int d, T drho, T vx1, T vx2, T vx3, T cusq)
{
#pragma clang fp contract(off)
    if (d == d000) return T(8. / 27.) * (drho - (drho + T(1)) * cusq);
    const T cu = proj<T>(d, vx1, vx2, vx3);
    return weight<T>(d) * (drho + (drho + T(1)) * (T(3) * cu + T(4.5) * cu * cu - cusq));
}

// PressureNonEquilibriumCompressible.cu: the inward neighbour's populations f
// (by direction) with their equilibrium replaced by the one of the prescribed
// density rho_bc, in place.
template <typename T> __device__ __forceinline__ void pressure_values(T (&f)[27], T rho_bc)
{
    // no FMA contraction: the same bits wherever the body is inlined
#pragma clang fp contract(off)
    // getCompressibleMacroscopicValues (MacroscopicQuantities.h:117-127)
    const T corners = ((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]));
    const T edges = (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P]))) +
                    ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]));
    const T faces = ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0])) + (f[d00M] + f[d00P]);
    T drho1 = ((corners + edges) + f[d000]) + faces;
    T vx1 = (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
            (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
            (f[dP00] - f[dM00]);
    T vx2 = (((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
            (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
            (f[d0P0] - f[d0M0]);
    T vx3 = (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
            (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
            (f[d00P] - f[d00M]);
    const T oor = T(1) / (drho1 + T(1));
    vx1 *= oor;
    vx2 *= oor;
    vx3 *= oor;
    const T cusq = T(1.5) * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
#pragma unroll
    for (int d = 0; d < 27; ++d) f[d] -= feq_shifted<T>(d, drho1, vx1, vx2, vx3, cusq);
    drho1 = rho_bc;
#pragma unroll
    for (int d = 0; d < 27; ++d) f[d] += feq_shifted<T>(d, drho1, vx1, vx2, vx3, cusq);
}

template <typename S, bool EVEN>
__device__ __forceinline__ void bc_pressure_entry(const Lat& L, const BCDev& B, uint32_t i)
{
    using T = typename Compute<S>::T;
    if (i >= B.n) return;
    const uint32_t k = B.k[i], kn = B.kn[i];
    if (k >= L.N || kn >= L.N) return;
    uint32_t kk[8], kkn[8];
    if (B.direct) {
        nbrs_direct(L, k, kk);
        nbrs_direct(L, kn, kkn);
    } else {
        nbrs_sparse(L, k, kk);
        nbrs_sparse(L, kn, kkn);
    }
    T m[27];
    gather<S, T, EVEN>(L, kkn, m);
    T f[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) f[dir_at(s)] = m[s];
    pressure_values<T>(f, bc_at<T>(B.rho, i));
    const int nx = cx(B.normal), ny = cy(B.normal), nz = cz(B.normal);
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        if (cx(d) * nx + cy(d) * ny + cz(d) * nz > 0) continue;
        Storage<S, T>::st(plane<S>(L, slot(d, EVEN)), kk[nbr(d)], f[d], L.st[weight_class(d)], L.ss[d]);
    }
}

template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) k_bc_pressure(Lat L, BCDev B)
{
    bc_pressure_entry<S, EVEN>(L, B, blockIdx.x * blockDim.x + threadIdx.x);
}

// Several boundary groups, possibly of different lattices of one storage
// type, in one launch: block b runs the job whose block range holds b (jobs
// in increasing first_block), at that job's parity (bit j of even_mask).
// Kinds 0 (velocity bounce-back), 1, 2, 3, 5 (interpolated) and 4 (pressure),
// each entry exactly as its own kernel computes it.  The groups of one launch
// must not touch each other's elements (the post-collision groups of a step
// write wall nodes' elements only; a pressure group runs with the groups of
// its own phase).
constexpr int kMaxBCJobs = 32;
struct BCJob {
    Lat L;
    BCDev B;
    int32_t kind;
    uint32_t first_block;
};
template <typename S>
__global__ void __launch_bounds__(256) k_bc_set(const BCJob* __restrict__ jobs, uint32_t njobs, uint32_t even_mask)
{
    uint32_t j = 0;
    for (uint32_t t = 1; t < njobs; ++t)
        if (jobs[t].first_block <= blockIdx.x) j = t;
    const BCJob& J = jobs[j];
    const uint32_t i = (blockIdx.x - J.first_block) * blockDim.x + threadIdx.x;
    const bool even = (even_mask >> j) & 1u;
    switch (J.kind) {
    case VFX_BC_VELOCITY_BOUNCE_BACK:
        if (even) bc_velocity_bb_entry<S, true>(J.L, J.B, i);
        else bc_velocity_bb_entry<S, false>(J.L, J.B, i);
        break;
    case VFX_BC_PRESSURE_NEQ:
        if (even) bc_pressure_entry<S, true>(J.L, J.B, i);
        else bc_pressure_entry<S, false>(J.L, J.B, i);
        break;
    default:
        if (even) bc_interp_entry<S, true, -1>(J.L, J.B, i, J.kind);
        else bc_interp_entry<S, false, -1>(J.L, J.B, i, J.kind);
        break;
    }
}


// Wall-model (stress) boundaries (BoundaryConditions/Stress/*.cu with the
// instantaneous wall model iMEM, iMEM.cuh:43-141): post-collision, parity EVEN.
// MODE 0 StressCompressible (interpolated no-slip with pressure at the node's
// turbulent-viscosity rate, wall term / (1 + q)), 1 StressBounceBackCompressible,
// 2 StressBounceBackPressureCompressible.  The links are visited in the
// kernels' order (E W N S T B NE SW SE NW TE BW BE TW TN BS BN TS TNE BSW BNE
// TSW TSE BNW BSE TNW), which fixes the wall-momentum sums.  Kept from the
// reference: the wall model takes q of the B (00M) link, forceFactor 1 + q(B)
// in MODE 0, the NE link's pressure term in MODE 0 weighs 2/27, and the
// launcher hands velocityY in for the z velocity (Stress.cu:57-59), which the
// launch below reproduces by passing the y array twice.
__host__ __device__ constexpr int stress_link(int j)
{
    constexpr int order[26] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 26, 23, 22, 21, 24, 25, 20};
    return order[j];
}

template <typename T>
__device__ __forceinline__ void wall_model(const BCDev& B, uint32_t i, uint32_t kN, const T* vx, const T* vy, const T* vz,
                                           T vxw, T vyw, T vzw, T rho, T q, T forceFactor, T wmx, T wmy, T wmz,
                                           T& Ux, T& Uy, T& Uz)
{
    const T* nrm[3] = {(const T*)B.wm_normal[0], (const T*)B.wm_normal[1], (const T*)B.wm_normal[2]};
    T* el[3] = {(T*)B.wm_el[0], (T*)B.wm_el[1], (T*)B.wm_el[2]};
    T* wm[3] = {(T*)B.wm_wall[0], (T*)B.wm_wall[1], (T*)B.wm_wall[2]};
    const T eps = T(0.001f);
    const T nx = nrm[0][i], ny = nrm[1][i], nz = nrm[2][i];
    T ex = eps * vx[kN] + (T(1.0) - eps) * el[0][i];
    T ey = eps * vy[kN] + (T(1.0) - eps) * el[1][i];
    T ez = eps * vz[kN] + (T(1.0) - eps) * el[2][i];
    el[0][i] = ex;
    el[1][i] = ey;
    el[2][i] = ez;
    T mx = eps * vxw + (T(1.0) - eps) * wm[0][i];
    T my = eps * vyw + (T(1.0) - eps) * wm[1][i];
    T mz = eps * vzw + (T(1.0) - eps) * wm[2][i];
    wm[0][i] = mx;
    wm[1][i] = my;
    wm[2][i] = mz;
    const T dn_el = ex * nx + ey * ny + ez * nz;
    ex -= dn_el * nx;
    ey -= dn_el * ny;
    ez -= dn_el * nz;
    const T mag_el = sqrt(ex * ex + ey * ey + ez * ez);
    const T dn_wm = mx * nx + my * ny + mz * nz;
    mx -= dn_wm * nx;
    my -= dn_wm * ny;
    mz -= dn_wm * nz;
    const T mag_wm = sqrt(mx * mx + my * my + mz * mz);
    const T dn_w = vxw * nx + vyw * ny + vzw * nz;
    const T wx = vxw - dn_w * nx, wy = vyw - dn_w * ny, wz = vzw - dn_w * nz;
    const T z = (T)B.wm_offset[i] + q;
    const T kappa = T(0.4);
    const T u_star = mag_el * kappa / (log(z / ((const T*)B.wm_z0)[i]));
    if (B.wm_ustar) ((T*)B.wm_ustar)[i] = u_star;
    const T tau_w = u_star * u_star;
    const T A = T(1.0);
    const T Fwx = (tau_w * A) * (wx / mag_wm);
    const T Fwy = (tau_w * A) * (wy / mag_wm);
    const T Fwz = (tau_w * A) * (wz / mag_wm);
    const T mdn = wmx * nx + wmy * ny + wmz * nz;
    const T Fx = Fwx - (wmx - mdn * nx) / rho;
    const T Fy = Fwy - (wmy - mdn * ny) / rho;
    const T Fz = Fwz - (wmz - mdn * nz) / rho;
    const T clip = T(2.0);
    const T cx_ = clip * ex, cy_ = clip * ey, cz_ = clip * ez;
    auto clamp = [&](T c, T F) -> T {
        const T v = T(-3.0) * F * forceFactor;
        return c > -c ? fmin(c, fmax(-c, v)) : fmax(c, fmin(-c, v));
    };
    Ux = clamp(cx_, Fx);
    Uy = clamp(cy_, Fy);
    Uz = clamp(cz_, Fz);
}

template <typename S, bool EVEN, int MODE>
__global__ void __launch_bounds__(256) k_bc_stress(Lat L, BCDev B)
{
    using T = typename Compute<S>::T;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t k = B.k[i], kN = B.kn[i];
    if (k >= L.N || kN >= L.N) return;
    uint32_t kk[8];
    if (B.direct) nbrs_direct(L, k, kk);
    else nbrs_sparse(L, k, kk);
    T p[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int e = inv(d);
        p[d] = Storage<S, T>::ld(plane<S>(L, slot(e, EVEN)), kk[nbr(e)], L.ld[weight_class(d)], L.sh[d]);
    }
    const T drho = p[dPMP] + p[dMPP] + p[dPPP] + p[dMMP] + p[dPMM] + p[dMPM] + p[dPPM] + p[dMMM] + p[d0PM] + p[d0MP] +
                   p[d0PP] + p[d0MM] + p[dP0M] + p[dM0P] + p[dP0P] + p[dM0M] + p[dPM0] + p[dMP0] + p[dPP0] + p[dMM0] +
                   p[d00P] + p[d00M] + p[d0P0] + p[d0M0] + p[dP00] + p[dM00] + p[d000];
    const T vx1 = (((p[dPMP] - p[dMPM]) - (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                   ((p[dP0M] - p[dM0P]) + (p[dP0P] - p[dM0M])) + ((p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                   (p[dP00] - p[dM00])) /
                  (T(1) + drho);
    const T vx2 = ((-(p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                   ((p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + (-(p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                   (p[d0P0] - p[d0M0])) /
                  (T(1) + drho);
    const T vx3 = (((p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) + (p[dMMP] - p[dPPM])) +
                   (-(p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + ((p[dP0P] - p[dM0M]) - (p[dP0M] - p[dM0P])) +
                   (p[d00P] - p[d00M])) /
                  (T(1) + drho);
    const T cu_sq = T(1.5) * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3) * (T(1) + drho);
    T om = (T)B.omega;
    if constexpr (MODE == 0) {
        const T nut = L.nut ? ((const T*)L.nut)[k] : T(0);
        om = om / (T(1) + T(3) * om * fmax(T(0), nut));
    }
    T qs[27];
#pragma unroll
    for (int d = 1; d < 27; ++d) qs[d] = bc_q<T>(B, d, i);
    T fin[27] = {};
    T wmx = T(0), wmy = T(0), wmz = T(0);
#pragma unroll
    for (int j = 0; j < 26; ++j) {
        const int d = stress_link(j);
        const T q = qs[d];
        if (!(q >= T(0) && q <= T(1))) continue;
        const T f = p[d], fi = p[inv(d)];
        T v;
        if constexpr (MODE == 0) {
            const T w = weight<T>(d);
            const T vlb = proj<T>(d, vx1, vx2, vx3);
            const T feq = w * (drho + T(4.5) * vlb * vlb * (T(1) + drho) - cu_sq);
            const T wp = d == dPP0 ? T(2. / 27.) : w; // the NE link's pressure weight
            v = (T(1) - q) / (T(1) + q) * (f - fi + (f + fi - T(2) * feq * om) / (T(1) - om)) * T(0.5) +
                (q * (f + fi)) / (T(1) + q) - wp * drho;
        } else if constexpr (MODE == 1) {
            v = f;
        } else {
            v = f - weight<T>(d) * drho;
        }
        fin[inv(d)] = v;
        const T s = f + v;
        if (cx(d) > 0) wmx += s;
        if (cx(d) < 0) wmx -= s;
        if (cy(d) > 0) wmy += s;
        if (cy(d) < 0) wmy -= s;
        if (cz(d) > 0) wmz += s;
        if (cz(d) < 0) wmz -= s;
    }
    const T qB = qs[d00M];
    T Ux, Uy, Uz;
    wall_model<T>(B, i, kN, (const T*)L.vx, (const T*)L.vy, (const T*)L.vy, vx1, vx2, vx3, T(1) + drho, qB,
                  MODE == 0 ? T(1.0) + qB : T(1.0), wmx, wmy, wmz, Ux, Uy, Uz);
#pragma unroll
    for (int j = 0; j < 26; ++j) {
        const int d = stress_link(j);
        const T q = qs[d];
        if (!(q >= T(0) && q <= T(1))) continue;
        T t = T(6) * weight<T>(d) * proj<T>(d, Ux, Uy, Uz);
        if constexpr (MODE == 0) t = t / (T(1) + q);
        const int e = inv(d);
        Storage<S, T>::st(plane<S>(L, slot(e, !EVEN)), kk[nbr(e)], fin[e] - t, L.st[weight_class(e)], L.ss[e]);
        if (cx(d) > 0) wmx += -t;
        if (cx(d) < 0) wmx -= -t;
        if (cy(d) > 0) wmy += -t;
        if (cy(d) < 0) wmy -= -t;
        if (cz(d) > 0) wmz += -t;
        if (cz(d) < 0) wmz -= -t;
    }
    if (B.wm_force[0]) {
        ((T*)B.wm_force[0])[i] = wmx;
        ((T*)B.wm_force[1])[i] = wmy;
        ((T*)B.wm_force[2])[i] = wmz;
    }
}

// Non-reflecting outflow (OutflowNonReflecting.cu:44-280; with CORR
// OutflowNonReflectingPressureCorrection.cu:46-245): pre-collision, parity EVEN.
// From the node's populations f and its inward neighbour's fn, every direction
// d entering through the boundary face (c_d . normal < 0) gets
// fn[d] cs + (1 - cs) f[d] [- w_d corr drho(f)], cs = 1 / sqrtf(3), written
// into the element the next collision of the other parity reads as direction
// d of the node: slot (d, !EVEN) at the node's own neighbour index.  That
// element belongs to the node across the boundary, which does not collide, so
// the collision of parity EVEN leaves it alone.  The corrected kernel builds
// its "neighbour" indices from the boundary node, so for directions with a
// negative component fn[d] is the node's own value; kept as is.  A normal that
// is not a face direction writes nothing (the kernels' default branch).
template <typename S, bool EVEN, bool CORR>
__global__ void __launch_bounds__(256) k_bc_outflow(Lat L, BCDev B)
{
    using T = typename Compute<S>::T;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t k = B.k[i], kn = B.kn[i];
    if (k >= L.N || kn >= L.N) return;
    const int nx = cx(B.normal), ny = cy(B.normal), nz = cz(B.normal);
    if (nx * nx + ny * ny + nz * nz != 1) return;
    uint32_t kk[8], kkn[8];
    if (B.direct) {
        nbrs_direct(L, k, kk);
        nbrs_direct(L, kn, kkn);
    } else {
        nbrs_sparse(L, k, kk);
        nbrs_sparse(L, kn, kkn);
    }
    T m[27], mn[27];
    gather<S, T, EVEN>(L, kk, m);
    gather<S, T, EVEN>(L, kkn, mn);
    T f[27], fn[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        const int d = dir_at(s);
        f[d] = m[s];
        fn[d] = (CORR && nbr(d) != 0) ? m[s] : mn[s];
    }
    T rc = T(0);
    if constexpr (CORR) {
        // vf::lbm::getDensity (MacroscopicQuantities.h:55-64)
        const T corners = ((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]));
        const T edges = (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P]))) +
                        ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]));
        const T faces = ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0])) + (f[d00M] + f[d00P]);
        rc = (T)B.corr * (((corners + edges) + f[d000]) + faces);
    }
    const T cs = T(1) / (T)sqrtf(3.f);
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        if (cx(d) * nx + cy(d) * ny + cz(d) * nz >= 0) continue;
        T v = fn[d] * cs + (T(1) - cs) * f[d];
        if constexpr (CORR) v = v - weight<T>(d) * rc;
        Storage<S, T>::st(plane<S>(L, slot(d, !EVEN)), kk[nbr(d)], v, L.st[weight_class(d)], L.ss[d]);
    }
}

// ---------------------------------------------- fused boundary conditions ----
// Boundary conditions applied inside the collide-stream kernel of a structured
// (direct) lattice, so that a boundary node's 27 populations are not read a
// second time by a kernel of its own.  On a face normal to x every element a
// boundary node touches sits in its own cache line (x is the fastest index),
// so a separate boundary kernel moves ~64 times the bytes it uses.
//  * post-collision interpolated kinds (1 no-slip, 2 velocity, 3 slip, 5
//    velocity with pressure): the thread that collided node k computes the
//    re-entering populations from its post-collision values and stores them,
//    as k_bc_interp would right after the kernel (the element written belongs
//    to the wall node, which no thread of the launch touches);
//  * the non-equilibrium pressure boundary on a face normal to +-x
//    (PressureNonEquilibriumCompressible, pre-collision): node k's incoming
//    populations are replaced at the start of the next step from its inward
//    neighbour k -+ 1, whose pre-collision populations the neighbouring lane
//    of the same wavefront has just gathered (cross-lane shuffle), instead of
//    by a kernel after the parity flip.  The values are the ones that kernel
//    would store and the collision read back (the storage round trip is
//    applied), so fused and separate runs agree bit for bit.
// Node k belongs to a group iff bit k of `mask` is set; its entry is
// base[k/32] + the number of set bits below it in its word (entries sorted by
// node).
struct FusedGroup {
    const uint32_t* mask;
    const uint32_t* base;
    BCDev B;
    int32_t kind;
};

__device__ __forceinline__ bool in_group(const FusedGroup& G, uint32_t k)
{
    return G.mask && ((G.mask[k >> 5] >> (k & 31u)) & 1u);
}

__device__ __forceinline__ bool fused_entry(const FusedGroup& G, uint32_t k, uint32_t& i)
{
    const uint32_t w = G.mask[k >> 5];
    const uint32_t b = k & 31u;
    if (!((w >> b) & 1u)) return false;
    i = G.base[k >> 5] + (uint32_t)__builtin_popcount(w & ((1u << b) - 1u));
    return i < G.B.n && G.B.k[i] == k; // malformed tables select nothing
}

// what a store of v as direction d followed by a load gives back (the value
// the separate boundary kernels exchange through memory)
template <typename S, typename T> __device__ __forceinline__ T round_trip(T v, float st, float ss, float ld, float sh)
{
    if constexpr (sizeof(S) == 2) return Storage<S, T>::dec2(Storage<S, T>::enc2(v, st, ss), ld, sh);
    else return v;
}

template <typename S, typename T, bool EVEN>
__device__ __forceinline__ void fused_post(const Lat& L, const FusedGroup& G, uint32_t k)
{
    uint32_t i;
    if (!fused_entry(G, k, i)) return;
    // The neighbour indices are recomputed from an index opaque to the
    // compiler: otherwise it keeps them, or the 64-bit element addresses of
    // the scatter, for the loads below and spills them on every node.
    asm volatile("" : "+v"(k));
    uint32_t kb[8];
    nbrs_direct(L, k, kb);
    // the node's post-collision populations, read back from the elements this
    // thread has just stored them to (no other node writes those elements in
    // the step): keeping the 27 collision outputs in registers for this rarely
    // taken path would spill them on every node
    T p[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        const int e = inv(d);
        p[d] = Storage<S, T>::ld(plane<S>(L, slot(e, EVEN)), kb[nbr(e)], L.ld[weight_class(d)], L.sh[d]);
    }
    interp_links<S, T, EVEN, -1>(L, G.B, i, kb[0], kb, p, G.kind);
}

// Link-parallel form of fused_post for a wavefront with few boundary nodes:
// the wavefront takes its boundary nodes one after another, lane x (< 27)
// loads the node's post-collision population of direction x, the lanes share
// them, and lane d (1..26) computes and stores the re-entering population of
// link d.  The node part and every link value are the expressions of
// interp_links (kinds 1, 2, 5; slip marks its face axes in link order and
// stays per node), so the results are the same bits.
constexpr uint32_t kLinkParallelNodes = 8; // boundary nodes per wavefront up to which links go across lanes
struct LaneDir {
    int x;              // direction of this lane (0 for lanes >= 27)
    int ix;             // the opposite direction
    int ex, ey, ez;     // components of x
    uint32_t rd_plane;  // plane of the node's post-collision population x
    uint32_t wr_plane;  // plane of the re-entering population of link x
    uint32_t nbsel;     // which of the node's 8 neighbour indices holds both
    float ld, sh, st, ss;
};
template <bool EVEN> __device__ __forceinline__ LaneDir lane_dir(const Lat& L, int lane)
{
    LaneDir r;
    const int x = lane < 27 ? lane : 0;
    r.x = x;
    int ex = 0, ey = 0, ez = 0, ix = 0;
    float sh = 0.f, ss = 0.f;
    // constant-index selects: the kernel argument arrays stay in place
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        if (x == d) {
            ex = cx(d);
            ey = cy(d);
            ez = cz(d);
            ix = inv(d);
            sh = L.sh[d];
            ss = L.ss[inv(d)];
        }
    }
    r.ix = ix;
    r.ex = ex;
    r.ey = ey;
    r.ez = ez;
    // population x is read where direction inv(x) is read (slot(inv(x), EVEN));
    // link x re-enters as inv(x), written at slot(inv(x), !EVEN); both at the
    // neighbour that is one step along every axis where c_x > 0
    r.rd_plane = (uint32_t)(EVEN ? ix : x);
    r.wr_plane = (uint32_t)(EVEN ? x : ix);
    r.nbsel = (uint32_t)((ex > 0) | ((ey > 0) << 1) | ((ez > 0) << 2));
    const int c = (ex != 0) + (ey != 0) + (ez != 0);
    r.ld = c == 0 ? L.ld[0] : c == 1 ? L.ld[1] : c == 2 ? L.ld[2] : L.ld[3];
    r.st = c == 0 ? L.st[0] : c == 1 ? L.st[1] : c == 2 ? L.st[2] : L.st[3];
    r.sh = sh;
    r.ss = ss;
    return r;
}
template <typename T> __device__ __forceinline__ T weight_rt(int ex, int ey, int ez)
{
    const int c = (ex != 0) + (ey != 0) + (ez != 0);
    return c == 0 ? T(8. / 27.) : c == 1 ? T(2. / 27.) : c == 2 ? T(1. / 54.) : T(1. / 216.);
}
template <typename S, typename T, bool EVEN>
__device__ __forceinline__ void fused_post_links(const Lat& L, const FusedGroup& G, uint32_t kb, int lane,
                                                 const LaneDir& D)
{
    uint32_t i;
    if (!fused_entry(G, kb, i)) return; // the same for every lane
    uint32_t nb[8];
    nbrs_direct(L, kb, nb);
    uint32_t idx = nb[0];
#pragma unroll
    for (int j = 1; j < 8; ++j)
        if (D.nbsel == (uint32_t)j) idx = nb[j];
    const T px = Storage<S, T>::ld(plane<S>(L, (int)D.rd_plane), idx, D.ld, D.sh);
    T p[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) p[d] = __shfl(px, d, 64);
    const int kd = G.kind;
    InterpNode<T> n = interp_node<T, -1>(L, G.B, i, kb, p, kd);
    T q = T(-1);
    if (lane >= 1 && lane < 27) q = bc_q<T>(G.B, lane, i);
    interp_wall_velocity<T>(n, G.B, i, kd);
    const T fi = __shfl(px, D.ix, 64);
    if (!(lane >= 1 && lane < 27) || !(q >= T(0) && q <= T(1))) return;
    const T w = weight_rt<T>(D.ex, D.ey, D.ez);
    const T vlb = proj_rt<T>(D.ex, D.ey, D.ez, n.vx1, n.vx2, n.vx3);
    const T vbc = kd == VFX_BC_NOSLIP_INTERPOLATED ? T(0) : proj_rt<T>(D.ex, D.ey, D.ez, n.VX, n.VY, n.VZ);
    const T v = interp_link_value<T>(n, kd, w, vlb, vbc, q, px, fi);
    Storage<S, T>::st(plane<S>(L, (int)D.wr_plane), idx, v, D.st, D.ss);
}

// K17 (default operation order of the storage form) with up to two fused
// post-collision groups and one fused pressure group; apply_pre = 0 on the
// first step of a run (no pressure boundary ran before it).
// The boundary paths need more registers than the collision; the occupancy of
// the plain kernel (7 waves per SIMD at <= 72 VGPRs for 2- and 4-byte storage)
// is kept by capping the register budget, so that those rarely taken paths
// spill instead.
template <typename S> constexpr int fused_waves() { return sizeof(S) == 8 ? 3 : 7; }
template <typename S, bool EVEN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fused_waves<S>())))
    k_step_fused(Lat L, Rates<typename Compute<S>::T> R, FusedGroup g0, FusedGroup g1, FusedGroup gp,
                 const uint32_t* any, int32_t apply_pre)
{
    using T = typename Compute<S>::T;
    constexpr bool FAST = sizeof(S) != 8;
    constexpr bool O2 = sizeof(S) == 2;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x, xcd_run<S>());
    const uint32_t row = blk / L.chunks;
    const uint32_t x = (blk - row * L.chunks) * blockDim.x + threadIdx.x;
    const uint32_t y = row % L.ny, z = row / L.ny;
    uint32_t kk[8];
    nbrs_direct_xyz(L, x < L.nx ? x : L.nx - 1, y, z, kk);
    const bool active = x < L.nx && !is_solid(L, kk[0]);
    // union of the groups' node masks, loaded with the populations: nodes of
    // no group (most of them) take no further round trip to memory
    const bool member = any == nullptr || ((any[kk[0] >> 5] >> (kk[0] & 31u)) & 1u);
    T m[27], mv[4], nut = T(0);
#pragma unroll
    for (int s = 0; s < 27; ++s) m[s] = T(0);
    if (active) gather<S, T, EVEN>(L, kk, m);
    if (gp.mask && apply_pre && __ballot(active && member)) {
        uint32_t ip = 0;
        const bool pl = active && member && fused_entry(gp, kk[0], ip);
        if (__ballot(pl)) {
            // inward neighbour k - 1 (outward normal +x) or k + 1 (-x): the lane next to this one
            const int lane = (int)(threadIdx.x & 63u);
            const int src = gp.B.normal == dP00 ? lane - 1 : lane + 1;
            T f[27];
#pragma unroll
            for (int s = 0; s < 27; ++s) f[dir_at(s)] = __shfl(m[s], src & 63, 64);
            const uint32_t kn = gp.B.normal == dP00 ? kk[0] - 1u : kk[0] + 1u;
            const bool row = gp.B.normal == dP00 ? x >= 1u : x + 1u < L.nx; // neighbour in this row
            if (pl && row && src >= 0 && src < 64 && gp.B.kn[ip] == kn) {
                pressure_values<T>(f, bc_at<T>(gp.B.rho, ip));
                const int ncx = cx(gp.B.normal);
#pragma unroll
                for (int s = 0; s < 27; ++s) {
                    const int d = dir_at(s);
                    if (cx(d) * ncx > 0) continue;
                    const int c = weight_class(d);
                    m[s] = round_trip<S, T>(f[d], L.st[c], L.ss[d], L.ld[c], L.sh[d]);
                }
            }
        }
    }
    // lanes without a node stay for the link-parallel boundary path below
    if (active) {
        collide<T, VFX_K17, FAST, O2>(m, R, mv, nut);
        scatter_post<S, T, EVEN, O2>(L, kk, m);
    }
    // the boundary code reads the stores above back from memory
    asm volatile("" ::: "memory");
    // only wavefronts with nodes of a group go on (whatever the boundary code
    // needs is set up behind this branch, not on every node)
    const uint32_t k = kk[0];
    const bool bc = active && member && (in_group(g0, k) || in_group(g1, k));
    const uint64_t any_bc = __ballot(bc);
    if (!any_bc) return;
    const int lane = (int)(threadIdx.x & 63u);
    if (__popcll(any_bc) <= kLinkParallelNodes) {
        // the stores of other lanes are read back
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        const LaneDir D = lane_dir<EVEN>(L, lane);
#pragma unroll 1
        for (int j = 0; j < 2; ++j) {
            const FusedGroup& G = j == 0 ? g0 : g1;
            if (!G.mask) continue;
            const bool mine = bc && in_group(G, k);
            uint64_t bal = __ballot(mine);
            if (G.kind == VFX_BC_SLIP) {
                if (mine) fused_post<S, T, EVEN>(L, G, k);
                continue;
            }
            while (bal) {
                const int b = __builtin_ctzll(bal);
                bal &= bal - 1;
                fused_post_links<S, T, EVEN>(L, G, (uint32_t)__shfl((int)k, b, 64), lane, D);
            }
        }
        return;
    }
    if (!bc) return;
    // one inlined copy of the boundary code serves both groups
#pragma unroll 1
    for (int j = 0; j < 2; ++j) {
        const FusedGroup& G = j == 0 ? g0 : g1;
        if (G.mask) fused_post<S, T, EVEN>(L, G, k);
    }
}

// The coarse step and the first fine step of a nest (or any two structured
// lattices of one storage type) in one launch: the plain step's body
// (k_step_direct) per level, each level's parity a template argument.  Level
// j owns the blocks [first_j, first_j + blocks_j), first_1 a multiple of 8 so
// that the XCD-aware tile order holds per level (padding blocks exit), blocks
// of 256 nodes along x.
template <typename S> struct LevelPair {
    Lat L0, L1;
    Rates<typename Compute<S>::T> R0, R1;
    uint32_t n, blocks0, first1, blocks1;
};
template <typename S, bool E0, bool E1>
__global__ void __launch_bounds__(256) k_step_levels(LevelPair<S> P)
{
    constexpr bool FAST = sizeof(S) != 8;
    if (P.n > 1 && blockIdx.x >= P.first1) {
        if (blockIdx.x - P.first1 < P.blocks1)
            step_direct_body<S, VFX_K17, E1, FAST>(P.L1, P.R1, blockIdx.x - P.first1, P.blocks1);
    } else if (blockIdx.x < P.blocks0) {
        step_direct_body<S, VFX_K17, E0, FAST>(P.L0, P.R0, blockIdx.x, P.blocks0);
    }
}

// ------------------------------------------------------------- batched ----
// Many independent box domains (an ensemble of cavities) advanced by one launch.
// Every domain is a box of nx*ny*nz nodes whose outer layer is wall; only the
// (nx-2)(ny-2)(nz-2) interior nodes collide.  Wall links are closed by the
// moving-wall halfway bounce-back of VelocityBounceBack_Device
// (VelocityBounceBack.cu:159-184) fused into the scatter: a population leaving
// towards a wall node is written, minus the wall term, straight into the element
// the node reads the opposite direction from in the next step.  Links into the
// top plane (z = nz-1) move with the domain's lid velocity.
//
// Both destinations of a population (its streaming slot, or the bounce-back
// slot) sit at the same element of two different planes.  Every population is
// streamed unconditionally (a wall link's streaming slot belongs to the wall
// node and is never read); only the bounce-back copy of a wall link is a
// lane-masked store.  The wall layer therefore holds stale values after a
// step, which nothing reads.

// Exact unsigned division by a run-time invariant d (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1):
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// q = (t + ((n - t) >> 1)) >> (l - 1) with t = mulhi(m, n); d = 1 is passed
// through.
struct FastDiv {
    uint32_t d, m, l;
};
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{d, 0, 0};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.l = l;
    f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    return f;
}
// a value the caller knows to be uniform across the wavefront, as a scalar
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    return (uint64_t)uni((uint32_t)v) | ((uint64_t)uni((uint32_t)(v >> 32)) << 32);
}

__device__ __forceinline__ uint32_t fastdiv(uint32_t n, const FastDiv& f)
{
    if (f.d <= 1) return n;
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> 1)) >> (f.l - 1);
}

template <typename T> struct DomainDev {
    uint64_t f_off;             // element offset of the domain's 27 planes in the pool
    uint32_t N, nx, ny, nz;     // box nodes / dims (wall layer included)
    uint32_t nfx, nfy, nfluid;  // interior dims and count
    uint32_t pad;
    FastDiv div_x, div_y;       // division by nfx, nfy
    T lid_term[27];             // 6 w_e (c_e . u_lid) per direction (used where c_e,z > 0)
    Rates<T> R;
};

struct BatchDev {
    void* pool;
    const void* domains;        // DomainDev<T>[n]
    const uint2* tiles;         // (domain, first interior node) per tile
    uint32_t ntiles;
    uint32_t xcd_run;           // tiles per XCD run (xcd_block); 1 = plain order
    float ld[4], st[4];
};

// tiles and domains are read-only for the whole launch (restrict), so their
// workgroup-uniform loads become scalar loads into SGPRs and every plane
// address stays a scalar base plus a 32-bit lane offset
template <typename S, bool EVEN, bool FAST>
__global__ void __launch_bounds__(256)
    k_step_batched(BatchDev B, const uint2* __restrict__ tiles, const DomainDev<typename Compute<S>::T>* __restrict__ doms)
{
    using T = typename Compute<S>::T;
    // workgroup b starts at tile xcd_block(b): runs of xcd_run consecutive
    // tiles (which share cache lines at their ends) on one XCD, or plain order
    const uint32_t t0 = B.xcd_run > 1 ? xcd_block(blockIdx.x, gridDim.x, B.xcd_run) : blockIdx.x;
    for (uint32_t t = t0; t < B.ntiles; t += gridDim.x) {
        const uint2 tile = tiles[t];
        const DomainDev<T>& D = doms[tile.x];
        const uint32_t idx = tile.y + threadIdx.x;
        const uint32_t nfx = D.nfx, nfy = D.nfy;
        if (idx >= D.nfluid) continue;
        const uint32_t r = fastdiv(idx, D.div_x);
        const uint32_t zi = fastdiv(r, D.div_y);
        const uint32_t x = idx - r * nfx + 1, y = r - zi * nfy + 1, z = zi + 1;
        Lat L;
        L.f = (S*)B.pool + D.f_off;
        L.N = D.N;
        L.nx = D.nx;
        L.ny = D.ny;
        L.nz = D.nz;
        uint32_t kk[8];
        nbrs_interior(L, x, y, z, kk);
        T m[27], mv[4], nut = T(0);
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int d = dir_at(s);
            m[s] = Storage<S, T>::ld(plane<S>(L, slot(d, EVEN)), kk[nbr(d)], B.ld[weight_class(d)]);
        }
        constexpr bool O2 = FAST && sizeof(S) == 2; // class-c populations come out as 2^c f
        collide<T, VFX_K17, FAST, O2>(m, D.R, mv, nut);
        // which faces of the box the node touches (lane masks: scalar logic)
        const bool xl = x == 1, xh = x == L.nx - 2, yl = y == 1, yh = y == L.ny - 2, zl = z == 1, zh = z == L.nz - 2;
        // one buffer descriptor over the domain's planes: a lane's byte offset
        // is one VGPR per neighbour node, the plane's byte offset an SGPR
        const uint32_t pb = L.N * (uint32_t)sizeof(S);
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(L.f, 27u * pb);
        uint32_t vo[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) vo[b] = kk[b] * (uint32_t)sizeof(S);
        // links into the lid plane carry the moving-wall term (only waves
        // with a lane next to the lid do the arithmetic; top = 0 leaves the
        // value unchanged exactly)
        if (__ballot(zh)) {
            const T top = zh ? T(1) : T(0);
#pragma unroll
            for (int s = 0; s < 27; ++s)
                if (cz(dir_at(s)) > 0)
                    m[s] = m[s] - (top * T(O2 ? 1 << weight_class(dir_at(s)) : 1)) * D.lid_term[dir_at(s)];
        }
        T stc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) stc[c] = (O2 && c > 0) ? B.st[c] * (1.f / float(1 << c)) : B.st[c];
        S sv[27];
#pragma unroll
        for (int s = 0; s < 27; ++s) sv[s] = Storage<S, T>::enc(m[s], stc[weight_class(dir_at(s))]);
        // One store per population.  Streaming: population e goes to the slot
        // direction d = inv(e) was read from.  On a wall link that slot is the
        // one the wall node would read e from -- the wall layer never collides,
        // nothing reads it -- so the lane writes the bounce-back instead: the
        // slot node k reads direction d from next step, in the other parity's
        // plane at the same element (plane offset moved into the lane's offset).
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int e = dir_at(s);
            const int d = inv(e);
            const bool wall = (cx(e) < 0 && xl) || (cx(e) > 0 && xh) || (cy(e) < 0 && yl) || (cy(e) > 0 && yh) ||
                              (cz(e) < 0 && zl) || (cz(e) > 0 && zh);
            // the lower of the two planes in the scalar offset, the lane adds a
            // non-negative distance to its plane
            const int se = slot(d, EVEN), so = slot(d, !EVEN), mn = se < so ? se : so;
            const uint32_t off = vo[nbr(d)] + (uint32_t)((wall ? so : se) - mn) * pb;
            Buf<S>::st(rs, off, (uint32_t)mn * pb, sv[s]);
        }
    }
}

// Initialise every domain's interior to the rest state f - w = 0 (both parities
// read the same zeros, walls included).
template <typename S>
__global__ void __launch_bounds__(256) k_batch_zero(void* pool, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) ((S*)pool)[i] = S(0);
}

} // namespace vfx

#include "refinement.hpp"

namespace vfx {

// ---------------------------------------------- VirtualFluids array export ----
// The distribution array exactly as VirtualFluids holds it (distributions.f[0]:
// 27 planes of n_nodes shifted populations at the current parity) in a chosen
// floating type: the array RestartObject::serialize/deserialize write and read
// (src/gpu/core/Restart/RestartObject.cpp:19-48).  Element j of plane p holds
// direction d = slot^-1(p) (p itself at even parity, inv(p) at odd), which
// fixes the FP16 scale and reference offset; FP32/FP64 storage is copied.
// Export writes NaN as 0, as serialize does.
template <typename S, bool EVEN, typename O>
__global__ void __launch_bounds__(256) k_export(Lat L, O* out)
{
    using T = typename Compute<S>::T;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n = 27ull * L.N;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int p = (int)(i / L.N);
        const uint32_t j = (uint32_t)(i - (uint64_t)p * L.N);
        const int d = EVEN ? p : inv(p);
        T v = Storage<S, T>::ld(plane<S>(L, p), j, L.ld[weight_class(d)], L.sh[d]);
        if (v != v) v = T(0);
        out[i] = (O)v;
    }
}
template <typename S, bool EVEN, typename O>
__global__ void __launch_bounds__(256) k_import(Lat L, const O* in)
{
    using T = typename Compute<S>::T;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n = 27ull * L.N;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int p = (int)(i / L.N);
        const uint32_t j = (uint32_t)(i - (uint64_t)p * L.N);
        const int d = EVEN ? p : inv(p);
        Storage<S, T>::st(plane<S>(L, p), j, (T)in[i], L.st[weight_class(d)], L.ss[d]);
    }
}

// ------------------------------------------------------- AMD viscosity ----
// Anisotropic minimum-dissipation eddy viscosity of every fluid node from the
// velocity fields the collision wrote (calcAMD,
// src/gpu/core/PostProcessor/TurbulentViscosityKernels.cu:62-113): central
// differences where both neighbours along an axis are fluid, one-sided where
// one is, zero where none is; nu_t = max(0, -C * numerator) / |grad u|^2
// (0 where the gradient vanishes).  Sparse layout: the -x/-y/-z neighbours
// come from the -x-y-z node (neighborInverse) and two + steps, fluid means
// typeOfGridNode == GEO_FLUID; direct layout: periodic index wrap, fluid =
// not solid.  Expression order as the reference's (FP64 reproduces it).
template <typename T>
__device__ __forceinline__ void amd_derivatives(uint32_t k, uint32_t km, uint32_t kp, bool fm, bool fp, const T* vx,
                                                const T* vy, const T* vz, T& dvx, T& dvy, T& dvz)
{
    const T div = (fm && fp) ? T(0.5) : T(1);
    dvx = ((fp ? vx[kp] : vx[k]) - (fm ? vx[km] : vx[k])) * div;
    dvy = ((fp ? vy[kp] : vy[k]) - (fm ? vy[km] : vy[k])) * div;
    dvz = ((fp ? vz[kp] : vz[k]) - (fm ? vz[km] : vz[k])) * div;
}

template <typename T> __global__ void __launch_bounds__(256) k_turb_amd(Lat L, int direct, T C)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.N) return;
    const uint32_t k = (uint32_t)i;
    uint32_t kp[3], km[3];
    bool fp[3], fm[3];
    if (direct) {
        if (is_solid(L, k)) return;
        const uint32_t x = k % L.nx, r = k / L.nx, y = r % L.ny, z = r / L.ny;
        const uint32_t sy = L.nx, sz = L.nx * L.ny;
        kp[0] = x + 1 == L.nx ? k - x : k + 1;
        km[0] = x == 0 ? k + L.nx - 1 : k - 1;
        kp[1] = y + 1 == L.ny ? k - y * sy : k + sy;
        km[1] = y == 0 ? k + (L.ny - 1) * sy : k - sy;
        kp[2] = z + 1 == L.nz ? k - z * sz : k + sz;
        km[2] = z == 0 ? k + (L.nz - 1) * sz : k - sz;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            fp[a] = !is_solid(L, kp[a]);
            fm[a] = !is_solid(L, km[a]);
        }
    } else {
        if (L.types[k] != kGeoFluid) return;
        kp[0] = L.nbx[k];
        kp[1] = L.nby[k];
        kp[2] = L.nbz[k];
        const uint32_t w = L.nbinv[k];
        km[0] = L.nbz[L.nby[w]];
        km[1] = L.nbz[L.nbx[w]];
        km[2] = L.nby[L.nbx[w]];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            fp[a] = L.types[kp[a]] == kGeoFluid;
            fm[a] = L.types[km[a]] == kGeoFluid;
        }
    }
    const T* vx = (const T*)L.vx;
    const T* vy = (const T*)L.vy;
    const T* vz = (const T*)L.vz;
    T dvxdx, dvxdy, dvxdz, dvydx, dvydy, dvydz, dvzdx, dvzdy, dvzdz;
    amd_derivatives<T>(k, km[0], kp[0], fm[0], fp[0], vx, vy, vz, dvxdx, dvydx, dvzdx);
    amd_derivatives<T>(k, km[1], kp[1], fm[1], fp[1], vx, vy, vz, dvxdy, dvydy, dvzdy);
    amd_derivatives<T>(k, km[2], kp[2], fm[2], fp[2], vx, vy, vz, dvxdz, dvydz, dvzdz);
    const T den = dvxdx * dvxdx + dvydx * dvydx + dvzdx * dvzdx + dvxdy * dvxdy + dvydy * dvydy + dvzdy * dvzdy +
                  dvxdz * dvxdz + dvydz * dvydz + dvzdz * dvzdz;
    const T num = (dvxdx * dvxdx + dvxdy * dvxdy + dvxdz * dvxdz) * dvxdx +
                  (dvydx * dvydx + dvydy * dvydy + dvydz * dvydz) * dvydy +
                  (dvzdx * dvzdx + dvzdy * dvzdy + dvzdz * dvzdz) * dvzdz +
                  (dvxdx * dvydx + dvxdy * dvydy + dvxdz * dvydz) * (dvxdy + dvydx) +
                  (dvxdx * dvzdx + dvxdy * dvzdy + dvxdz * dvzdz) * (dvxdz + dvzdx) +
                  (dvydx * dvzdx + dvydy * dvzdy + dvydz * dvzdz) * (dvydz + dvzdy);
    ((T*)L.nut)[k] = den != T(0) ? fmax(T(0), -C * num) / den : T(0);
}

// ------------------------------------------------------------ launchers ----
constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Occupancy cap of the structured step: LDS requested at launch and not
// used, so that at most 160 KiB / cap workgroups share a CU.  The FP32
// reduced-operation step (58 VGPRs, 8 waves per SIMD by registers) runs
// fastest with at most 6 workgroups per CU (MI355X, 512x256x256, tuned launch
// shapes: 1.33 ms per step against 1.40 uncapped; the tuner then picks
// 128-node tiles, ~3 waves per SIMD); the sparse FP32 step with the same
// cap 1.497 ms against 1.530.  The FP16 step is fastest at the 7
// waves its registers allow (6 or 5 waves per SIMD measured slower), FP64 and
// the reference-order kernels are limited by their registers; other
// collision templates keep the uncapped launch.
template <typename S, int COLL, bool FAST> constexpr unsigned step_lds_cap()
{
    return sizeof(S) == 4 && FAST && coll_op(COLL) == VFX_K17 && coll_turb(COLL) == kTurbNone &&
                   !coll_body_force(COLL)
               ? 24576u
               : 0u;
}

template <typename S, int COLL, bool EVEN, bool FAST>
hipError_t launch_step_t(const Lat& L, int direct, const Rates<typename Compute<S>::T>& R, hipStream_t st)
{
    if (direct) {
        const unsigned grid = L.chunks * L.ny * L.nz;
        const unsigned block = (unsigned)((L.nx + L.chunks - 1) / L.chunks);
        const unsigned b = block <= 64 ? 64 : block <= 128 ? 128 : 256;
        hipLaunchKernelGGL((k_step_direct<S, COLL, EVEN, FAST>), dim3(grid), dim3(b), (step_lds_cap<S, COLL, FAST>()), st, L,
                           R);
    } else {
        const uint64_t n = node_items(L, 0);
        if (n == 0) return hipSuccess;
        hipLaunchKernelGGL((k_step_sparse<S, COLL, EVEN, FAST>), dim3(grid_for(n, kBlock)), dim3(kBlock),
                           (step_lds_cap<S, COLL, FAST>()), st, L, R);
    }
    return hipGetLastError();
}

template <typename S, int COLL>
hipError_t launch_step_p(const Lat& L, int direct, bool even, const Rates<typename Compute<S>::T>& R, hipStream_t st)
{
    return even ? launch_step_t<S, COLL, true, false>(L, direct, R, st)
                : launch_step_t<S, COLL, false, false>(L, direct, R, st);
}

// variant: 0 = default (FP64: reference operation order; FP32/FP16: reduced form), 1 = reference order;
// turb: TurbModel of the K17 collision (reference operation order in every variant)
template <typename S> hipError_t launch_step(const Lat& L, int direct, int coll, bool even, int variant, double omega,
                                             const double* lim, const double* force, int turb, double sgs,
                                             hipStream_t st)
{
    using T = typename Compute<S>::T;
    const Rates<T> R = make_rates<T>(omega, lim, force, sgs);
    const bool bf = (coll & kCollBodyForce) != 0; // ApplyBodyForce (K17 only)
    coll &= 15;
    const bool fast = sizeof(T) == 4 && variant == 0 && coll == VFX_K17;
    constexpr int K = VFX_K17, BF = kCollBodyForce;
    if (coll == VFX_K17 && turb != kTurbNone) {
        switch (turb) {
        case kTurbSmagorinsky:
            return bf ? launch_step_p<S, K | (kTurbSmagorinsky << 4) | BF>(L, direct, even, R, st)
                      : launch_step_p<S, K | (kTurbSmagorinsky << 4)>(L, direct, even, R, st);
        case kTurbAMD:
            return bf ? launch_step_p<S, K | (kTurbAMD << 4) | BF>(L, direct, even, R, st)
                      : launch_step_p<S, K | (kTurbAMD << 4)>(L, direct, even, R, st);
        default:
            return bf ? launch_step_p<S, K | (kTurbQR << 4) | BF>(L, direct, even, R, st)
                      : launch_step_p<S, K | (kTurbQR << 4)>(L, direct, even, R, st);
        }
    }
    if (coll == VFX_K17 && bf) {
        if (fast) return even ? launch_step_t<S, K | BF, true, true>(L, direct, R, st)
                              : launch_step_t<S, K | BF, false, true>(L, direct, R, st);
        return even ? launch_step_t<S, K | BF, true, false>(L, direct, R, st)
                    : launch_step_t<S, K | BF, false, false>(L, direct, R, st);
    }
    if (coll == VFX_K17) {
        if (fast) return even ? launch_step_t<S, VFX_K17, true, true>(L, direct, R, st)
                              : launch_step_t<S, VFX_K17, false, true>(L, direct, R, st);
        return even ? launch_step_t<S, VFX_K17, true, false>(L, direct, R, st)
                    : launch_step_t<S, VFX_K17, false, false>(L, direct, R, st);
    }
    return even ? launch_step_t<S, VFX_BGK, true, false>(L, direct, R, st)
                : launch_step_t<S, VFX_BGK, false, false>(L, direct, R, st);
}

// K17 collide-and-stream of a structured lattice with fused boundary groups
// (k_step_fused); same grid as the direct k_step_direct launch
template <typename S>
hipError_t launch_step_fused(const Lat& L, bool even, double omega, const double* lim, const double* force,
                             const FusedGroup& g0, const FusedGroup& g1, const FusedGroup& gp, const uint32_t* any,
                             int apply_pre, hipStream_t st)
{
    using T = typename Compute<S>::T;
    const Rates<T> R = make_rates<T>(omega, lim, force, 0.0);
    const unsigned grid = L.chunks * L.ny * L.nz;
    const unsigned block = (unsigned)((L.nx + L.chunks - 1) / L.chunks);
    const unsigned b = block <= 64 ? 64 : block <= 128 ? 128 : 256;
    if (even) hipLaunchKernelGGL((k_step_fused<S, true>), dim3(grid), dim3(b), 0, st, L, R, g0, g1, gp, any, apply_pre);
    else hipLaunchKernelGGL((k_step_fused<S, false>), dim3(grid), dim3(b), 0, st, L, R, g0, g1, gp, any, apply_pre);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_bc_set(const BCJob* jobs, uint32_t njobs, uint32_t blocks, uint32_t even_mask, hipStream_t st)
{
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((k_bc_set<S>), dim3(blocks), dim3(kBlock), 0, st, jobs, njobs, even_mask);
    return hipGetLastError();
}

// The levels of one launch: per level its lattice, rates inputs and parity.
struct LevelInput {
    Lat L;
    double omega;
    const double* lim;
    const double* force;
    bool even;
};
template <typename S> hipError_t launch_step_levels(const LevelInput* in, int n, hipStream_t st)
{
    using T = typename Compute<S>::T;
    constexpr bool FAST = sizeof(S) != 8;
    if (n == 1) {
        // one level: the plain step itself, in the lattice's own launch shape
        const Rates<T> R = make_rates<T>(in[0].omega, in[0].lim, in[0].force, 0.0);
        return in[0].even ? launch_step_t<S, VFX_K17, true, FAST>(in[0].L, 1, R, st)
                          : launch_step_t<S, VFX_K17, false, FAST>(in[0].L, 1, R, st);
    }
    static_assert(sizeof(LevelPair<S>) <= 4000, "kernel arguments");
    LevelPair<S> P;
    std::memset(&P, 0, sizeof(P));
    P.n = (uint32_t)n;
    Lat* L[2] = {&P.L0, &P.L1};
    Rates<T>* R[2] = {&P.R0, &P.R1};
    uint32_t* blocks[2] = {&P.blocks0, &P.blocks1};
    // one workgroup width for both levels: their own where they agree, else 256
    const uint32_t tile = in[0].L.tile == in[1].L.tile && in[0].L.tile ? in[0].L.tile : 256u;
    uint32_t first = 0;
    for (int j = 0; j < n; ++j) {
        *L[j] = in[j].L;
        L[j]->chunks = (L[j]->nx + tile - 1u) / tile; // blocks of `tile` nodes in this launch
        *R[j] = make_rates<T>(in[j].omega, in[j].lim, in[j].force, 0.0);
        *blocks[j] = L[j]->chunks * L[j]->ny * L[j]->nz;
        if (j == 1) P.first1 = first;
        first += (*blocks[j] + kXcds - 1) / kXcds * kXcds;
    }
    if (first == 0) return hipSuccess;
    const bool e0 = in[0].even, e1 = n > 1 ? in[1].even : e0;
    const dim3 g(first), b(tile);
    switch ((e0 ? 2 : 0) + (e1 ? 1 : 0)) {
    case 3: hipLaunchKernelGGL((k_step_levels<S, true, true>), g, b, 0, st, P); break;
    case 2: hipLaunchKernelGGL((k_step_levels<S, true, false>), g, b, 0, st, P); break;
    case 1: hipLaunchKernelGGL((k_step_levels<S, false, true>), g, b, 0, st, P); break;
    default: hipLaunchKernelGGL((k_step_levels<S, false, false>), g, b, 0, st, P); break;
    }
    return hipGetLastError();
}

// one entry per wavefront of the step's grid (4 per 256 work items)
inline uint64_t sparse_groups(const Lat& L) { return (node_items(L, 0) + kBlock - 1) / kBlock * (kBlock / 64); }
inline hipError_t launch_sparse_groups(const Lat& L, uint4* table, uint32_t ngroups, hipStream_t st)
{
    const uint64_t items = (uint64_t)ngroups * 64u;
    hipLaunchKernelGGL((k_sparse_groups<0>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, st, L, table, ngroups);
    return hipGetLastError();
}

template <typename S> hipError_t launch_turb_amd(const Lat& L, int direct, double sgs, hipStream_t st)
{
    using T = typename Compute<S>::T;
    hipLaunchKernelGGL((k_turb_amd<T>), dim3(grid_for(L.N, kBlock)), dim3(kBlock), 0, st, L, direct, (T)sgs);
    return hipGetLastError();
}

template <typename S> hipError_t launch_init(const Lat& L, int direct, bool even, const double* drho, const double* ux,
                                             const double* uy, const double* uz, hipStream_t st)
{
    const uint64_t n = node_items(L, direct);
    if (n == 0) return hipSuccess;
    if (even) hipLaunchKernelGGL((k_init<S, true>), dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, L, direct, drho, ux, uy, uz);
    else hipLaunchKernelGGL((k_init<S, false>), dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, L, direct, drho, ux, uy, uz);
    return hipGetLastError();
}

template <typename S> hipError_t launch_rw(const Lat& L, int direct, bool even, double* out, const double* in, hipStream_t st)
{
    const uint64_t n = node_items(L, direct);
    if (n == 0) return hipSuccess;
    const dim3 g(grid_for(n, kBlock)), b(kBlock);
    if (out) {
        if (even) hipLaunchKernelGGL((k_read<S, true>), g, b, 0, st, L, direct, out);
        else hipLaunchKernelGGL((k_read<S, false>), g, b, 0, st, L, direct, out);
    } else {
        if (even) hipLaunchKernelGGL((k_write<S, true>), g, b, 0, st, L, direct, in);
        else hipLaunchKernelGGL((k_write<S, false>), g, b, 0, st, L, direct, in);
    }
    return hipGetLastError();
}

// out_fp64: the VirtualFluids array in double (else float); exactly one of
// out / in is non-null
template <typename S>
hipError_t launch_vf_array(const Lat& L, bool even, bool fp64, void* out, const void* in, hipStream_t st)
{
    const uint64_t n = 27ull * L.N;
    const unsigned g = (unsigned)std::min<uint64_t>((n + kBlock - 1) / kBlock, 65536), b = kBlock;
    if (out) {
        if (fp64) {
            if (even) hipLaunchKernelGGL((k_export<S, true, double>), g, b, 0, st, L, (double*)out);
            else hipLaunchKernelGGL((k_export<S, false, double>), g, b, 0, st, L, (double*)out);
        } else {
            if (even) hipLaunchKernelGGL((k_export<S, true, float>), g, b, 0, st, L, (float*)out);
            else hipLaunchKernelGGL((k_export<S, false, float>), g, b, 0, st, L, (float*)out);
        }
    } else {
        if (fp64) {
            if (even) hipLaunchKernelGGL((k_import<S, true, double>), g, b, 0, st, L, (const double*)in);
            else hipLaunchKernelGGL((k_import<S, false, double>), g, b, 0, st, L, (const double*)in);
        } else {
            if (even) hipLaunchKernelGGL((k_import<S, true, float>), g, b, 0, st, L, (const float*)in);
            else hipLaunchKernelGGL((k_import<S, false, float>), g, b, 0, st, L, (const float*)in);
        }
    }
    return hipGetLastError();
}

template <typename S> hipError_t launch_macro(const Lat& L, int direct, bool even, double* r, double* vx, double* vy,
                                              double* vz, hipStream_t st)
{
    const uint64_t n = node_items(L, direct);
    if (n == 0) return hipSuccess;
    const dim3 g(grid_for(n, kBlock)), b(kBlock);
    if (even) hipLaunchKernelGGL((k_macro<S, true>), g, b, 0, st, L, direct, r, vx, vy, vz);
    else hipLaunchKernelGGL((k_macro<S, false>), g, b, 0, st, L, direct, r, vx, vy, vz);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_macro_vf(const Lat& L, int direct, bool even, void* rho, void* vx, void* vy, void* vz, void* press,
                           hipStream_t st)
{
    using T = typename Compute<S>::T;
    T *r = (T*)rho, *x = (T*)vx, *y = (T*)vy, *z = (T*)vz, *p = (T*)press;
    const dim3 b(kBlock);
    if (even) hipLaunchKernelGGL((k_macro_vf<S, true, 0>), dim3(grid_for(L.N, kBlock)), b, 0, st, L, direct, r, x, y, z, p);
    else hipLaunchKernelGGL((k_macro_vf<S, false, 0>), dim3(grid_for(L.N, kBlock)), b, 0, st, L, direct, r, x, y, z, p);
    if (!direct && !L.types && L.nfluid) {
        if (even) hipLaunchKernelGGL((k_macro_vf<S, true, 1>), dim3(grid_for(L.nfluid, kBlock)), b, 0, st, L, direct, r, x, y, z, p);
        else hipLaunchKernelGGL((k_macro_vf<S, false, 1>), dim3(grid_for(L.nfluid, kBlock)), b, 0, st, L, direct, r, x, y, z, p);
    }
    return hipGetLastError();
}

template <typename S>
hipError_t launch_mean_vf(const Lat& L, int direct, bool even, void* vx, void* vy, void* vz, void* rho, void* press,
                          hipStream_t st)
{
    using T = typename Compute<S>::T;
    T *x = (T*)vx, *y = (T*)vy, *z = (T*)vz, *r = (T*)rho, *p = (T*)press;
    const dim3 b(kBlock);
    if (direct || L.types) {
        if (even) hipLaunchKernelGGL((k_mean_vf<S, true, 0>), dim3(grid_for(L.N, kBlock)), b, 0, st, L, direct, x, y, z, r, p);
        else hipLaunchKernelGGL((k_mean_vf<S, false, 0>), dim3(grid_for(L.N, kBlock)), b, 0, st, L, direct, x, y, z, r, p);
    } else if (L.nfluid) {
        if (even) hipLaunchKernelGGL((k_mean_vf<S, true, 1>), dim3(grid_for(L.nfluid, kBlock)), b, 0, st, L, direct, x, y, z, r, p);
        else hipLaunchKernelGGL((k_mean_vf<S, false, 1>), dim3(grid_for(L.nfluid, kBlock)), b, 0, st, L, direct, x, y, z, r, p);
    }
    return hipGetLastError();
}

template <typename T> hipError_t launch_zero5(uint64_t n, void* a, void* b, void* c, void* d, void* e, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_zero5<T>), dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, (T*)a, (T*)b, (T*)c, (T*)d, (T*)e, n);
    return hipGetLastError();
}

template <typename S>
hipError_t launch_batched(const BatchDev& B, bool even, int variant, unsigned block, unsigned grid, hipStream_t st)
{
    using T = typename Compute<S>::T;
    const bool fast = sizeof(T) == 4 && variant == 0;
    const dim3 g(grid), b(block);
    if (fast) {
        if (even) hipLaunchKernelGGL((k_step_batched<S, true, true>), g, b, 0, st, B, B.tiles, (const DomainDev<T>*)B.domains);
        else hipLaunchKernelGGL((k_step_batched<S, false, true>), g, b, 0, st, B, B.tiles, (const DomainDev<T>*)B.domains);
    } else {
        if (even) hipLaunchKernelGGL((k_step_batched<S, true, false>), g, b, 0, st, B, B.tiles, (const DomainDev<T>*)B.domains);
        else hipLaunchKernelGGL((k_step_batched<S, false, false>), g, b, 0, st, B, B.tiles, (const DomainDev<T>*)B.domains);
    }
    return hipGetLastError();
}

template <typename S> hipError_t launch_batch_zero(void* pool, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint64_t g = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_batch_zero<S>), dim3((unsigned)(g < 65536 ? g : 65536)), dim3(kBlock), 0, st, pool, n);
    return hipGetLastError();
}

inline uint32_t reduce_blocks(uint64_t n) { return (uint32_t)((n + kRedChunk - 1) / kRedChunk); }

template <typename S> hipError_t launch_reduce_mass(const Lat& L, int direct, bool even, double* ws, double* out, hipStream_t st)
{
    const uint64_t n = node_items(L, direct);
    const uint32_t nb = reduce_blocks(n ? n : 1);
    if (even) hipLaunchKernelGGL((k_reduce_mass<S, true>), dim3(nb), dim3(kRedBlock), 0, st, L, direct, ws);
    else hipLaunchKernelGGL((k_reduce_mass<S, false>), dim3(nb), dim3(kRedBlock), 0, st, L, direct, ws);
    hipLaunchKernelGGL((k_reduce_finish<4>), dim3(1), dim3(kRedBlock), 0, st, ws, nb, out);
    return hipGetLastError();
}

template <typename S> hipError_t launch_stats(const Lat& L, int direct, bool even, double* sums, hipStream_t st)
{
    const uint64_t n = node_items(L, direct);
    if (n == 0) return hipSuccess;
    const dim3 g(grid_for(n, kBlock)), b(kBlock);
    if (even) hipLaunchKernelGGL((k_stats<S, true>), g, b, 0, st, L, direct, sums);
    else hipLaunchKernelGGL((k_stats<S, false>), g, b, 0, st, L, direct, sums);
    return hipGetLastError();
}

template <typename S> hipError_t launch_bc(const Lat& L, const BCDev& B, int kind, bool even, hipStream_t st)
{
    if (B.n == 0) return hipSuccess;
    const dim3 g(grid_for(B.n, kBlock)), b(kBlock);
    switch (kind) {
    case VFX_BC_VELOCITY_BOUNCE_BACK:
        if (even) hipLaunchKernelGGL((k_bc_velocity_bb<S, true>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_velocity_bb<S, false>), g, b, 0, st, L, B);
        break;
    case VFX_BC_NOSLIP_INTERPOLATED:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_NOSLIP_INTERPOLATED>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_NOSLIP_INTERPOLATED>), g, b, 0, st, L, B);
        break;
    case VFX_BC_VELOCITY_INTERPOLATED:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_VELOCITY_INTERPOLATED>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_VELOCITY_INTERPOLATED>), g, b, 0, st, L, B);
        break;
    case VFX_BC_SLIP:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_SLIP>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_SLIP>), g, b, 0, st, L, B);
        break;
    case VFX_BC_PRESSURE_NEQ:
        if (even) hipLaunchKernelGGL((k_bc_pressure<S, true>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_pressure<S, false>), g, b, 0, st, L, B);
        break;
    case VFX_BC_VELOCITY_PRESSURE_INTERPOLATED:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_VELOCITY_PRESSURE_INTERPOLATED>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_VELOCITY_PRESSURE_INTERPOLATED>), g, b, 0, st, L, B);
        break;
    case VFX_BC_SLIP_TURBULENT_VISCOSITY:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_SLIP_TURBULENT_VISCOSITY>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_SLIP_TURBULENT_VISCOSITY>), g, b, 0, st, L, B);
        break;
    case VFX_BC_STRESS_INTERPOLATED:
        if (even) hipLaunchKernelGGL((k_bc_stress<S, true, 0>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_stress<S, false, 0>), g, b, 0, st, L, B);
        break;
    case VFX_BC_STRESS_BOUNCE_BACK:
        if (even) hipLaunchKernelGGL((k_bc_stress<S, true, 1>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_stress<S, false, 1>), g, b, 0, st, L, B);
        break;
    case VFX_BC_STRESS_BOUNCE_BACK_PRESSURE:
        if (even) hipLaunchKernelGGL((k_bc_stress<S, true, 2>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_stress<S, false, 2>), g, b, 0, st, L, B);
        break;
    case VFX_BC_PRECURSOR_VELOCITY:
        if (even) hipLaunchKernelGGL((k_bc_interp<S, true, VFX_BC_PRECURSOR_VELOCITY>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_interp<S, false, VFX_BC_PRECURSOR_VELOCITY>), g, b, 0, st, L, B);
        break;
    case VFX_BC_PRECURSOR_DISTRIBUTIONS:
        if (even) hipLaunchKernelGGL((k_bc_precursor_dist<S, true>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_precursor_dist<S, false>), g, b, 0, st, L, B);
        break;
    case VFX_BC_OUTFLOW_NON_REFLECTING:
        if (even) hipLaunchKernelGGL((k_bc_outflow<S, true, false>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_outflow<S, false, false>), g, b, 0, st, L, B);
        break;
    case VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION:
        if (even) hipLaunchKernelGGL((k_bc_outflow<S, true, true>), g, b, 0, st, L, B);
        else hipLaunchKernelGGL((k_bc_outflow<S, false, true>), g, b, 0, st, L, B);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#define VFX_INSTANTIATE_LAUNCHERS(S)                                                                                   \
    template hipError_t launch_step<S>(const Lat&, int, int, bool, int, double, const double*, const double*, int,      \
                                       double, hipStream_t);                                                           \
    template hipError_t launch_step_fused<S>(const Lat&, bool, double, const double*, const double*, const FusedGroup&, \
                                             const FusedGroup&, const FusedGroup&, const uint32_t*, int, hipStream_t); \
    template hipError_t launch_turb_amd<S>(const Lat&, int, double, hipStream_t);                                      \
    template hipError_t launch_bc_set<S>(const BCJob*, uint32_t, uint32_t, uint32_t, hipStream_t);                    \
    template hipError_t launch_step_levels<S>(const LevelInput*, int, hipStream_t);                                    \
    template hipError_t launch_init<S>(const Lat&, int, bool, const double*, const double*, const double*,             \
                                       const double*, hipStream_t);                                                    \
    template hipError_t launch_rw<S>(const Lat&, int, bool, double*, const double*, hipStream_t);                      \
    template hipError_t launch_vf_array<S>(const Lat&, bool, bool, void*, const void*, hipStream_t);                  \
    template hipError_t launch_macro<S>(const Lat&, int, bool, double*, double*, double*, double*, hipStream_t);       \
    template hipError_t launch_macro_vf<S>(const Lat&, int, bool, void*, void*, void*, void*, void*, hipStream_t);     \
    template hipError_t launch_mean_vf<S>(const Lat&, int, bool, void*, void*, void*, void*, void*, hipStream_t);      \
    template hipError_t launch_bc<S>(const Lat&, const BCDev&, int, bool, hipStream_t);                            \
    template hipError_t launch_reduce_mass<S>(const Lat&, int, bool, double*, double*, hipStream_t);              \
    template hipError_t launch_stats<S>(const Lat&, int, bool, double*, hipStream_t);                             \
    template hipError_t launch_batched<S>(const BatchDev&, bool, int, unsigned, unsigned, hipStream_t);           \
    template hipError_t launch_batch_zero<S>(void*, uint64_t, hipStream_t);                                        \
    template hipError_t launch_scale<S>(const Lat&, const Lat&, const ICellsDev&, bool, bool, bool, double, double,     \
                                        hipStream_t);                                                                  \
    template hipError_t launch_scale_both<S>(const Lat&, const Lat&, const ICellsDev&, const ICellsDev&, bool, bool,     \
                                             double, double, hipStream_t);

#define VFX_EXTERN_LAUNCHERS(S)                                                                                        \
    extern template hipError_t launch_step<S>(const Lat&, int, int, bool, int, double, const double*, const double*,   \
                                              int, double, hipStream_t);                                               \
    extern template hipError_t launch_step_fused<S>(const Lat&, bool, double, const double*, const double*,            \
                                                    const FusedGroup&, const FusedGroup&, const FusedGroup&,           \
                                                    const uint32_t*, int, hipStream_t);                                \
    extern template hipError_t launch_turb_amd<S>(const Lat&, int, double, hipStream_t);                               \
    extern template hipError_t launch_bc_set<S>(const BCJob*, uint32_t, uint32_t, uint32_t, hipStream_t);             \
    extern template hipError_t launch_step_levels<S>(const LevelInput*, int, hipStream_t);                             \
    extern template hipError_t launch_init<S>(const Lat&, int, bool, const double*, const double*, const double*,      \
                                              const double*, hipStream_t);                                             \
    extern template hipError_t launch_rw<S>(const Lat&, int, bool, double*, const double*, hipStream_t);               \
    extern template hipError_t launch_vf_array<S>(const Lat&, bool, bool, void*, const void*, hipStream_t);           \
    extern template hipError_t launch_macro<S>(const Lat&, int, bool, double*, double*, double*, double*, hipStream_t); \
    extern template hipError_t launch_macro_vf<S>(const Lat&, int, bool, void*, void*, void*, void*, void*, hipStream_t); \
    extern template hipError_t launch_mean_vf<S>(const Lat&, int, bool, void*, void*, void*, void*, void*, hipStream_t); \
    extern template hipError_t launch_bc<S>(const Lat&, const BCDev&, int, bool, hipStream_t);                     \
    extern template hipError_t launch_reduce_mass<S>(const Lat&, int, bool, double*, double*, hipStream_t);       \
    extern template hipError_t launch_stats<S>(const Lat&, int, bool, double*, hipStream_t);                      \
    extern template hipError_t launch_batched<S>(const BatchDev&, bool, int, unsigned, unsigned, hipStream_t);    \
    extern template hipError_t launch_batch_zero<S>(void*, uint64_t, hipStream_t);                                 \
    extern template hipError_t launch_scale<S>(const Lat&, const Lat&, const ICellsDev&, bool, bool, bool, double,     \
                                               double, hipStream_t);                                                   \
    extern template hipError_t launch_scale_both<S>(const Lat&, const Lat&, const ICellsDev&, const ICellsDev&, bool,   \
                                                    bool, double, double, hipStream_t);

} // namespace vfx
