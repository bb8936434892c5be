// refinement.hpp -- interface interpolation between a coarse level and its 2:1
// fine level (included by kernels.hpp).
//
// Reference behaviour: scaleCoarseToFineCompressible_Device
// (src/gpu/core/GridScaling/scaleCoarseToFineCompressible.cu:305-529) and
// scaleFineToCoarseCompressible_Device (scaleFineToCoarseCompressible.cu:76-124)
// with vf::lbm::MomentsOnSourceNodeSet / interpolateCF / interpolateFC
// (src/lbm/interpolation/InterpolationCoefficients.h:74-432,
// src/lbm/refinement/InterpolationCF.h:57-306, InterpolationFC.h:56-263).
//
// Work mapping: eight lanes per interface cell, one per source node (corner c,
// bit 0 = +x, bit 1 = +y, bit 2 = +z).  Each lane gathers its corner's 27
// pre-collision populations and reduces them to the nine source moments
// (drho, u, five non-equilibrium second moments); the eight moment sets meet in
// LDS and every lane forms the cell's interpolation polynomial.  Coarse to
// fine: lane c then writes fine node c of the cell (at (+-1/4, +-1/4, +-1/4)
// cell units).  Fine to coarse: the eight lanes evaluate the coarse node at the
// cell centre together and each writes a share of its 27 populations.  So every
// lane issues 27 loads and <= 27 stores, 32 cells per 256-thread workgroup.
//
// The polynomial (per velocity component v, x/y/z in cell units):
//   v = v000 + v100 x + v010 y + v001 z + v200 x^2 + v020 y^2 + v002 z^2
//     + v110 xy + v101 xz + v011 yz + v111 xyz,   drho trilinear.
// With corner sums S q = sum_c q_c and signed sums D_x q = sum_c s_x q_c,
// D_xy q = sum_c s_x s_y q_c, ... (s = +-1 the corner's side) the reference's
// coefficients are (a, b, c for vx, vy, vz; d for drho):
//   v100 = D_x v / 4, v110 = D_xy v / 2, v111 = D_xyz v (likewise y, z, d)
//   a200 = (2 D_xy vy + 2 D_xz vz + D_x kxxMyy + D_x kxxMzz) / 16
//   a020 = (D_y kxy - 2 D_xy vy) / 8,  a002 = (D_z kxz - 2 D_xz vz) / 8
//   b200 = (D_x kxy - 2 D_xy vx) / 8,  b002 = (D_z kyz - 2 D_yz vz) / 8
//   b020 = (2 D_xy vx + 2 D_yz vz - 2 D_y kxxMyy + D_y kxxMzz) / 16
//   c200 = (D_x kxz - 2 D_xz vx) / 8,  c020 = (D_y kyz - 2 D_yz vy) / 8
//   c002 = (2 D_xz vx + 2 D_yz vy - 2 D_z kxxMzz + D_z kxxMyy) / 16
//   v000 = S v / 8 - (v200 + v020 + v002) / 4,  d000 = S drho / 8
// followed by the reference's shift to the cell offset (refinement into a
// wall) and its pressure curvature term.
#pragma once

namespace vfx {

struct ICellsDev {
    const uint32_t* coarse; // coarse-to-fine: coarse corner-MMM node; fine-to-coarse: destination coarse node
    const uint32_t* fine;   // fine corner-MMM node of the cell
    const void* ox;         // offsets of the cells, arithmetic type T (may be null: no offset)
    const void* oy;
    const void* oz;
    uint32_t n;
    int32_t direct_c, direct_f;
};

constexpr int kCellLanes = 8;
constexpr int kCellsPerBlock = 256 / kCellLanes;

// the node at corner c of the cell whose corner MMM is `base`
__device__ __forceinline__ uint32_t cell_corner(const Lat& L, int direct, uint32_t base, int c)
{
    if (direct) {
        uint32_t x = base % L.nx;
        const uint32_t r = base / L.nx;
        uint32_t y = r % L.ny, z = r / L.ny;
        if (c & 1) x = (x + 1 == L.nx) ? 0u : x + 1;
        if (c & 2) y = (y + 1 == L.ny) ? 0u : y + 1;
        if (c & 4) z = (z + 1 == L.nz) ? 0u : z + 1;
        return x + L.nx * (y + L.ny * z);
    }
    uint32_t k = base;
    if (c & 1) k = L.nbx[k];
    if (c & 2) k = L.nby[k];
    if (c & 4) k = L.nbz[k];
    return k;
}

enum { kDrho, kVx, kVy, kVz, kKxy, kKyz, kKxz, kKxxMyy, kKxxMzz, kMoments };

// MomentsOnSourceNode::calculate (InterpolationCoefficients.h:89-128)
template <typename T> __device__ __forceinline__ void source_moments(const T (&m)[27], T omega, T (&s)[kMoments])
{
    T drho, jx, jy, jz;
    pyramid_sums(m, drho, jx, jy, jz);
    const T r = T(1) / (drho + T(1));
    const T vx = jx * r, vy = jy * r, vz = jz * r;
    T pxy = 0, pyz = 0, pxz = 0, pxxMyy = 0, pxxMzz = 0;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int d = dir_at(q);
        const int x = cx(d), y = cy(d), z = cz(d);
        if (x * y) pxy += T(x * y) * m[q];
        if (y * z) pyz += T(y * z) * m[q];
        if (x * z) pxz += T(x * z) * m[q];
        if (x * x - y * y) pxxMyy += T(x * x - y * y) * m[q];
        if (x * x - z * z) pxxMzz += T(x * x - z * z) * m[q];
    }
    s[kDrho] = drho;
    s[kVx] = vx;
    s[kVy] = vy;
    s[kVz] = vz;
    s[kKxy] = T(-3) * omega * (pxy * r - vx * vy);
    s[kKyz] = T(-3) * omega * (pyz * r - vy * vz);
    s[kKxz] = T(-3) * omega * (pxz * r - vx * vz);
    s[kKxxMyy] = T(-1.5) * omega * (pxxMyy * r - (vx * vx - vy * vy));
    s[kKxxMzz] = T(-1.5) * omega * (pxxMzz * r - (vx * vx - vz * vz));
}

enum { P000, P100, P010, P001, P200, P020, P002, P110, P101, P011, P111, NP };

template <typename T> struct Coeffs {
    T a[NP], b[NP], c[NP], d[NP];
    T laplace_rho;
};

template <typename T> struct CornerSums {
    T S, Dx, Dy, Dz, Dxy, Dxz, Dyz, Dxyz;
};

// sums over the 8 corners of moment `q` held in LDS as src[c * kMoments + q]
template <typename T> __device__ __forceinline__ CornerSums<T> corner_sums(const T* src, int q)
{
    CornerSums<T> r{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const T v = src[c * kMoments + q];
        const T vx = (c & 1) ? v : -v;  // s_x v
        const T vxy = (c & 2) ? vx : -vx;
        const T vy = (c & 2) ? v : -v;
        const T vyz = (c & 4) ? vy : -vy;
        r.S += v;
        r.Dx += vx;
        r.Dy += vy;
        r.Dz += (c & 4) ? v : -v;
        r.Dxy += vxy;
        r.Dxz += (c & 4) ? vx : -vx;
        r.Dyz += vyz;
        r.Dxyz += (c & 4) ? vxy : -vxy;
    }
    return r;
}

template <typename T> __device__ __forceinline__ void linear_part(const CornerSums<T>& s, T* p)
{
    p[P100] = s.Dx * T(0.25);
    p[P010] = s.Dy * T(0.25);
    p[P001] = s.Dz * T(0.25);
    p[P110] = s.Dxy * T(0.5);
    p[P101] = s.Dxz * T(0.5);
    p[P011] = s.Dyz * T(0.5);
    p[P111] = s.Dxyz;
}

// MomentsOnSourceNodeSet::calculateCoefficients (:174-428) including the shift
// to the cell offset (:389-424), which, as in the reference, keeps the second
// order and xyz coefficients and ignores the xyz contribution to the shifted
// first-order ones
template <typename T>
__device__ __forceinline__ void cell_coefficients(const T* src, T xoff, T yoff, T zoff, Coeffs<T>& k)
{
    const CornerSums<T> vx = corner_sums(src, kVx), vy = corner_sums(src, kVy), vz = corner_sums(src, kVz);
    const CornerSums<T> rho = corner_sums(src, kDrho);
    const CornerSums<T> kxy = corner_sums(src, kKxy), kyz = corner_sums(src, kKyz), kxz = corner_sums(src, kKxz);
    const CornerSums<T> kMyy = corner_sums(src, kKxxMyy), kMzz = corner_sums(src, kKxxMzz);
    T* a = k.a;
    T* b = k.b;
    T* c = k.c;
    T* d = k.d;
    linear_part(vx, a);
    linear_part(vy, b);
    linear_part(vz, c);
    linear_part(rho, d);
    const T e16 = T(1. / 16.), e8 = T(0.125);
    a[P200] = (T(2) * vy.Dxy + T(2) * vz.Dxz + kMyy.Dx + kMzz.Dx) * e16;
    a[P020] = (kxy.Dy - T(2) * vy.Dxy) * e8;
    a[P002] = (kxz.Dz - T(2) * vz.Dxz) * e8;
    b[P200] = (kxy.Dx - T(2) * vx.Dxy) * e8;
    b[P020] = (T(2) * vx.Dxy + T(2) * vz.Dyz - T(2) * kMyy.Dy + kMzz.Dy) * e16;
    b[P002] = (kyz.Dz - T(2) * vz.Dyz) * e8;
    c[P200] = (kxz.Dx - T(2) * vx.Dxz) * e8;
    c[P020] = (kyz.Dy - T(2) * vy.Dyz) * e8;
    c[P002] = (T(2) * vx.Dxz + T(2) * vy.Dyz - T(2) * kMzz.Dz + kMyy.Dz) * e16;
    a[P000] = vx.S * e8 - (a[P200] + a[P020] + a[P002]) * T(0.25);
    b[P000] = vy.S * e8 - (b[P200] + b[P020] + b[P002]) * T(0.25);
    c[P000] = vz.S * e8 - (c[P200] + c[P020] + c[P002]) * T(0.25);
    d[P000] = rho.S * e8;
    d[P200] = d[P020] = d[P002] = T(0);
    k.laplace_rho = (xoff != T(0) || yoff != T(0) || zoff != T(0))
                        ? T(0)
                        : T(-3) * (a[P100] * a[P100] + b[P010] * b[P010] + c[P001] * c[P001]) -
                              T(6) * (b[P100] * a[P010] + c[P100] * a[P001] + c[P010] * b[P001]);
    T* polys[4] = {a, b, c, d};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T* p = polys[i];
        p[P000] += xoff * p[P100] + yoff * p[P010] + zoff * p[P001] + xoff * xoff * p[P200] +
                   yoff * yoff * p[P020] + zoff * zoff * p[P002] + xoff * yoff * p[P110] + xoff * zoff * p[P101] +
                   yoff * zoff * p[P011];
        p[P100] += T(2) * xoff * p[P200] + yoff * p[P110] + zoff * p[P101];
        p[P010] += T(2) * yoff * p[P020] + xoff * p[P110] + zoff * p[P011];
        p[P001] += T(2) * zoff * p[P002] + xoff * p[P101] + yoff * p[P011];
    }
}

template <typename T> __device__ __forceinline__ T poly_value(const T* p, T x, T y, T z)
{
    return p[P000] + x * p[P100] + y * p[P010] + z * p[P001] + x * x * p[P200] + y * y * p[P020] +
           z * z * p[P002] + x * y * p[P110] + x * z * p[P101] + y * z * p[P011] + x * y * z * p[P111];
}
template <typename T> __device__ __forceinline__ void poly_gradient(const T* p, T x, T y, T z, T (&g)[3])
{
    g[0] = p[P100] + T(2) * x * p[P200] + y * p[P110] + z * p[P101] + y * z * p[P111];
    g[1] = p[P010] + T(2) * y * p[P020] + x * p[P110] + z * p[P011] + x * z * p[P111];
    g[2] = p[P001] + T(2) * z * p[P002] + x * p[P101] + y * p[P011] + x * y * p[P111];
}

// destination populations from drho, u and the velocity gradient (gx = grad
// vx, ...): second-order central moments carry the destination level's
// non-equilibrium stress -(eps/omega)(1+drho) S_ab, third order is zero,
// fourth and sixth order take drho/9 and drho/27 (InterpolationCF.h:180-236)
template <typename T>
__device__ __forceinline__ void destination(T drho, T ux, T uy, T uz, const T (&gx)[3], const T (&gy)[3],
                                            const T (&gz)[3], T eps, T omega, T (&m)[27])
{
    const T s = eps / omega * (T(1) + drho);
    const T c1o3 = T(1. / 3.), c2o3 = T(2. / 3.);
    const T mxxMyy = -c2o3 * (gx[0] - gy[1]) * s;
    const T mxxMzz = -c2o3 * (gx[0] - gz[2]) * s;
#pragma unroll
    for (int q = 0; q < 27; ++q) m[q] = T(0);
    m[cube(0, 0, 0)] = drho;
    m[cube(0, 1, 1)] = -c1o3 * (gy[2] + gz[1]) * s;
    m[cube(1, 0, 1)] = -c1o3 * (gx[2] + gz[0]) * s;
    m[cube(1, 1, 0)] = -c1o3 * (gx[1] + gy[0]) * s;
    m[cube(2, 0, 0)] = c1o3 * (mxxMyy + mxxMzz + drho);
    m[cube(0, 2, 0)] = c1o3 * (T(-2) * mxxMyy + mxxMzz + drho);
    m[cube(0, 0, 2)] = c1o3 * (mxxMyy - T(2) * mxxMzz + drho);
    m[cube(0, 2, 2)] = m[cube(2, 0, 2)] = m[cube(2, 2, 0)] = drho * T(1. / 9.);
    m[cube(2, 2, 2)] = drho * T(1. / 27.);
    backward_sweeps(m, ux, uy, uz);
}

// hasTurbulentViscosity: a level with an eddy-viscosity model relaxes node k at
// omega / (1 + 3 omega nu_t[k]) (calculateOmegaWithturbulentViscosity), in the
// source moments of each corner and at each destination node
// (scaleCoarseToFineCompressible.cu:94, 369; scaleFineToCoarseCompressible.cu:64)
template <typename T> __device__ __forceinline__ T node_omega(const Lat& L, uint32_t k, T omega)
{
    return L.nut ? omega / (T(1) + T(3) * omega * ((const T*)L.nut)[k]) : omega;
}

// moments of this lane's source node into LDS, polynomial of the cell
template <typename S, bool EVEN_SRC>
__device__ __forceinline__ void load_cell(const Lat& L, int direct, uint32_t base, int lane, bool active,
                                          typename Compute<S>::T omega, typename Compute<S>::T* sh)
{
    using T = typename Compute<S>::T;
    if (active) {
        const uint32_t k = cell_corner(L, direct, base, lane);
        uint32_t kk[8];
        if (direct) nbrs_direct(L, k, kk);
        else nbrs_sparse(L, k, kk);
        T m[27];
        gather<S, T, EVEN_SRC>(L, kk, m);
        T s[kMoments];
        source_moments(m, node_omega<T>(L, k, omega), s);
#pragma unroll
        for (int q = 0; q < kMoments; ++q) sh[lane * kMoments + q] = s[q];
    }
}

// Coarse to fine: coarse moments at the coarse parity, fine nodes written
// pre-collision at the fine parity (scaleCoarseToFineCompressible.cu:305-529;
// eps = 1/2, fine omega).
template <typename S, bool EVEN_C, bool EVEN_F>
__device__ __forceinline__ void scale_cf_block(const Lat& C, const Lat& F, const ICellsDev& cells, double omega_c,
                                               double omega_f, uint32_t block, typename Compute<S>::T* sh)
{
    using T = typename Compute<S>::T;
    const int lane = threadIdx.x % kCellLanes;
    const int slot_in_block = threadIdx.x / kCellLanes;
    const uint32_t cell = block * kCellsPerBlock + slot_in_block;
    const bool active = cell < cells.n && cells.coarse[cell] < C.N && cells.fine[cell] < F.N;
    T* src = sh + slot_in_block * kCellLanes * kMoments;
    load_cell<S, EVEN_C>(C, cells.direct_c, active ? cells.coarse[cell] : 0u, lane, active, (T)omega_c, src);
    __syncthreads();
    if (!active) return;
    Coeffs<T> k;
    const T xo = cells.ox ? ((const T*)cells.ox)[cell] : T(0);
    const T yo = cells.oy ? ((const T*)cells.oy)[cell] : T(0);
    const T zo = cells.oz ? ((const T*)cells.oz)[cell] : T(0);
    cell_coefficients(src, xo, yo, zo, k);
    const T x = (lane & 1) ? T(0.25) : T(-0.25);
    const T y = (lane & 2) ? T(0.25) : T(-0.25);
    const T z = (lane & 4) ? T(0.25) : T(-0.25);
    // InterpolationCF.h:168-169: pressure with the curvature term 3 x^2 laplace
    const T drho = poly_value(k.d, x, y, z) + T(3) * x * x * k.laplace_rho;
    T gx[3], gy[3], gz[3];
    poly_gradient(k.a, x, y, z, gx);
    poly_gradient(k.b, x, y, z, gy);
    poly_gradient(k.c, x, y, z, gz);
    T m[27];
    const uint32_t kf = cell_corner(F, cells.direct_f, cells.fine[cell], lane);
    destination(drho, poly_value(k.a, x, y, z), poly_value(k.b, x, y, z), poly_value(k.c, x, y, z), gx, gy, gz,
                T(0.5), node_omega<T>(F, kf, (T)omega_f), m);
    uint32_t kk[8];
    if (cells.direct_f) nbrs_direct(F, kf, kk);
    else nbrs_sparse(F, kf, kk);
    scatter_pre<S, T, EVEN_F>(F, kk, m);
}

template <typename S, bool EVEN_C, bool EVEN_F>
__global__ void __launch_bounds__(256) k_scale_cf(Lat C, Lat F, ICellsDev cells, double omega_c, double omega_f)
{
    __shared__ typename Compute<S>::T sh[kCellsPerBlock * kCellLanes * kMoments];
    scale_cf_block<S, EVEN_C, EVEN_F>(C, F, cells, omega_c, omega_f, blockIdx.x, sh);
}

// Fine to coarse: fine moments at the fine parity, the coarse node written
// pre-collision at the coarse parity (scaleFineToCoarseCompressible.cu:76-124;
// eps = 2, coarse omega, pressure d000 - laplace / 4 of InterpolationFC.h:138).
template <typename S, bool EVEN_C, bool EVEN_F>
__device__ __forceinline__ void scale_fc_block(const Lat& C, const Lat& F, const ICellsDev& cells, double omega_c,
                                               double omega_f, uint32_t block, typename Compute<S>::T* sh)
{
    using T = typename Compute<S>::T;
    const int lane = threadIdx.x % kCellLanes;
    const int slot_in_block = threadIdx.x / kCellLanes;
    const uint32_t cell = block * kCellsPerBlock + slot_in_block;
    const bool active = cell < cells.n && cells.coarse[cell] < C.N && cells.fine[cell] < F.N;
    T* src = sh + slot_in_block * kCellLanes * kMoments;
    load_cell<S, EVEN_F>(F, cells.direct_f, active ? cells.fine[cell] : 0u, lane, active, (T)omega_f, src);
    __syncthreads();
    if (!active) return;
    Coeffs<T> k;
    const T xo = cells.ox ? ((const T*)cells.ox)[cell] : T(0);
    const T yo = cells.oy ? ((const T*)cells.oy)[cell] : T(0);
    const T zo = cells.oz ? ((const T*)cells.oz)[cell] : T(0);
    cell_coefficients(src, xo, yo, zo, k);
    const T drho = k.d[P000] - T(2) * k.laplace_rho * T(0.125);
    const T gx[3] = {k.a[P100], k.a[P010], k.a[P001]};
    const T gy[3] = {k.b[P100], k.b[P010], k.b[P001]};
    const T gz[3] = {k.c[P100], k.c[P010], k.c[P001]};
    T m[27];
    const uint32_t kc = cells.coarse[cell];
    destination(drho, k.a[P000], k.b[P000], k.c[P000], gx, gy, gz, T(2), node_omega<T>(C, kc, (T)omega_c), m);
    uint32_t kk[8];
    if (cells.direct_c) nbrs_direct(C, kc, kk);
    else nbrs_sparse(C, kc, kk);
#pragma unroll
    for (int s = 0; s < 27; ++s) {
        if (s % kCellLanes != lane) continue;
        const int d = dir_at(s);
        Storage<S, T>::st(plane<S>(C, slot(d, EVEN_C)), kk[nbr(d)], m[s], C.st[weight_class(d)], C.ss[d]);
    }
}

template <typename S, bool EVEN_C, bool EVEN_F>
__global__ void __launch_bounds__(256) k_scale_fc(Lat C, Lat F, ICellsDev cells, double omega_c, double omega_f)
{
    __shared__ typename Compute<S>::T sh[kCellsPerBlock * kCellLanes * kMoments];
    scale_fc_block<S, EVEN_C, EVEN_F>(C, F, cells, omega_c, omega_f, blockIdx.x, sh);
}

// Fine-to-coarse and coarse-to-fine of one nest in one launch: the first
// nfc blocks restrict, the rest prolongate.  Valid when no coarse node the
// restriction writes is a corner of a coarse-to-fine cell and no fine node the
// prolongation writes is a corner of a fine-to-coarse cell (the nest layout
// keeps them apart: restriction two coarse cells inside the nest, prolongation
// on its outermost layer); the values are those of the two launches in order.
template <typename S, bool EVEN_C, bool EVEN_F>
__global__ void __launch_bounds__(256)
    k_scale_both(Lat C, Lat F, ICellsDev fc, ICellsDev cf, double omega_c, double omega_f, uint32_t nfc)
{
    __shared__ typename Compute<S>::T sh[kCellsPerBlock * kCellLanes * kMoments];
    if (blockIdx.x < nfc) scale_fc_block<S, EVEN_C, EVEN_F>(C, F, fc, omega_c, omega_f, blockIdx.x, sh);
    else scale_cf_block<S, EVEN_C, EVEN_F>(C, F, cf, omega_c, omega_f, blockIdx.x - nfc, sh);
}

template <typename S>
hipError_t launch_scale_both(const Lat& C, const Lat& F, const ICellsDev& fc, const ICellsDev& cf, bool even_c,
                             bool even_f, double omega_c, double omega_f, hipStream_t st)
{
    const uint32_t nfc = (fc.n + kCellsPerBlock - 1) / kCellsPerBlock, ncf = (cf.n + kCellsPerBlock - 1) / kCellsPerBlock;
    if (nfc + ncf == 0) return hipSuccess;
    const dim3 g(nfc + ncf), b(kCellsPerBlock * kCellLanes);
    const int v = (even_c ? 2 : 0) + (even_f ? 1 : 0);
    switch (v) {
    case 3: hipLaunchKernelGGL((k_scale_both<S, true, true>), g, b, 0, st, C, F, fc, cf, omega_c, omega_f, nfc); break;
    case 2: hipLaunchKernelGGL((k_scale_both<S, true, false>), g, b, 0, st, C, F, fc, cf, omega_c, omega_f, nfc); break;
    case 1: hipLaunchKernelGGL((k_scale_both<S, false, true>), g, b, 0, st, C, F, fc, cf, omega_c, omega_f, nfc); break;
    default: hipLaunchKernelGGL((k_scale_both<S, false, false>), g, b, 0, st, C, F, fc, cf, omega_c, omega_f, nfc); break;
    }
    return hipGetLastError();
}

template <typename S>
hipError_t launch_scale(const Lat& C, const Lat& F, const ICellsDev& cells, bool coarse_to_fine, bool even_c,
                        bool even_f, double omega_c, double omega_f, hipStream_t st)
{
    if (cells.n == 0) return hipSuccess;
    const dim3 g((cells.n + kCellsPerBlock - 1) / kCellsPerBlock), b(kCellsPerBlock * kCellLanes);
    const int v = (even_c ? 2 : 0) + (even_f ? 1 : 0);
    if (coarse_to_fine) {
        switch (v) {
        case 3: hipLaunchKernelGGL((k_scale_cf<S, true, true>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        case 2: hipLaunchKernelGGL((k_scale_cf<S, true, false>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        case 1: hipLaunchKernelGGL((k_scale_cf<S, false, true>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        default: hipLaunchKernelGGL((k_scale_cf<S, false, false>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        }
    } else {
        switch (v) {
        case 3: hipLaunchKernelGGL((k_scale_fc<S, true, true>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        case 2: hipLaunchKernelGGL((k_scale_fc<S, true, false>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        case 1: hipLaunchKernelGGL((k_scale_fc<S, false, true>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        default: hipLaunchKernelGGL((k_scale_fc<S, false, false>), g, b, 0, st, C, F, cells, omega_c, omega_f); break;
        }
    }
    return hipGetLastError();
}

} // namespace vfx
