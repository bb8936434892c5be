// collision.hpp -- node-local collision operators for the gfx950 stepper.
//
// * k17<T>  : cumulant K17 (Geier et al. 2017), the operator VirtualFluids
//             implements in src/lbm/collision/K17CompressibleNavierStokes.h:87-591.
// * bgk<T>  : compressible BGK, VirtualFluids' "BGKCompSP27"
//             (src/gpu/core/Kernel/Compressible/NavierStokes/B92/
//              B92CompressibleNavierStokes_Device.cu:189-234).
//
// Both work on shifted ("well-conditioned") populations f_d - w_d held in a
// 27-register array indexed by velocity cube position (see cube()).  Every
// arithmetic expression keeps VirtualFluids' evaluation order, so with
// T = double and -ffp-contract=off the result is the reference CPU path's bit
// for bit.  For T = float the same sequence runs in FP32; the relaxation rates
// that depend only on omega are computed once on the host (Rates<T>).
#pragma once

#include "d3q27.hpp"

namespace vfx {

// Register slot of the population with velocity (i-1, j-1, k-1), i,j,k in {0,1,2}.
// Before the forward transform slot (i,j,k) holds f of that direction; after it,
// slot (a,b,c) holds the central moment of order (a,b,c).
__host__ __device__ constexpr int cube(int i, int j, int k) { return 9 * i + 3 * j + k; }

// Direction number (VirtualFluids order) stored at cube slot s.
__host__ __device__ constexpr int dir_at(int s)
{
    constexpr int t[27] = {dMMM, dMM0, dMMP, dM0M, dM00, dM0P, dMPM, dMP0, dMPP,
                           d0MM, d0M0, d0MP, d00M, d000, d00P, d0PM, d0P0, d0PP,
                           dPMM, dPM0, dPMP, dP0M, dP00, dP0P, dPPM, dPP0, dPPP};
    return t[s];
}
// Cube slot of direction d (inverse of dir_at).
__host__ __device__ constexpr int slot_of(int d) { return cube(cx(d) + 1, cy(d) + 1, cz(d) + 1); }

// Relaxation rates that depend on omega only
// (K17CompressibleNavierStokes.h:248-273), computed once per launch.
template <typename T> struct Rates {
    T omega, OxyyPxzz, OxyyMxzz, Oxyz, factorA, factorB;
    T sgs; // SGSconstant of the eddy-viscosity models
    T limP, limM, limD;
    T fx, fy, fz; // half force added to the velocity (:164-166)
    // folded constants of the reduced-operation form (k17_fast)
    T om1;                // 1 - omega
    T c3;                 // 3 (1 - omega/2), Galilean correction weight
    T limP1, limM1, limD1; // (1 - O3) * lambda of the three limiter groups
    T A4, B4;             // (1/omega - 1/2) * 2/3 * factorA, (1/omega - 1/2) * 1/3 * factorB
};

// the rates of one omega, in T with the reference's expression order (so
// T=double matches it exactly), on the host per launch or on the device per
// node (eddy-viscosity models)
template <typename T> __host__ __device__ inline void omega_rates(T w, Rates<T>& r)
{
    r.omega = w;
    r.OxyyPxzz = T(8) * (-T(2) + w) * (T(1) + T(2) * w) / (-T(8) - T(14) * w + T(7) * w * w);
    r.OxyyMxzz = T(8) * (-T(2) + w) * (-T(7) + T(4) * w) / (T(56) - T(50) * w + T(9) * w * w);
    r.Oxyz = T(24) * (-T(2) + w) * (-T(2) - T(7) * w + T(3) * w * w) /
             (T(48) + T(152) * w - T(130) * w * w + T(29) * w * w * w);
    r.factorA = (T(4) + T(2) * w - T(3) * w * w) / (T(2) - T(7) * w + T(5) * w * w);
    r.factorB = (T(4) + T(28) * w - T(14) * w * w) / (T(6) - T(21) * w + T(15) * w * w);
}

template <typename T>
inline Rates<T> make_rates(double omega, const double lim[3], const double force[3], double sgs = 0.0)
{
    Rates<T> r;
    omega_rates<T>((T)omega, r);
    r.sgs = (T)sgs;
    r.limP = (T)lim[0];
    r.limM = (T)lim[1];
    r.limD = (T)lim[2];
    r.fx = (T)force[0];
    r.fy = (T)force[1];
    r.fz = (T)force[2];
    // folded constants, evaluated in double and rounded once
    const double od = omega;
    const double oPP = 8. * (-2. + od) * (1. + 2. * od) / (-8. - 14. * od + 7. * od * od);
    const double oMM = 8. * (-2. + od) * (-7. + 4. * od) / (56. - 50. * od + 9. * od * od);
    const double oD = 24. * (-2. + od) * (-2. - 7. * od + 3. * od * od) /
                      (48. + 152. * od - 130. * od * od + 29. * od * od * od);
    const double fA = (4. + 2. * od - 3. * od * od) / (2. - 7. * od + 5. * od * od);
    const double fB = (4. + 28. * od - 14. * od * od) / (6. - 21. * od + 15. * od * od);
    r.om1 = (T)(1. - od);
    r.c3 = (T)(3. * (1. - 0.5 * od));
    r.limP1 = (T)((1. - oPP) * lim[0]);
    r.limM1 = (T)((1. - oMM) * lim[1]);
    r.limD1 = (T)((1. - oD) * lim[2]);
    r.A4 = (T)((1. / od - 0.5) * (2. / 3.) * fA);
    r.B4 = (T)((1. / od - 0.5) * (1. / 3.) * fB);
    return r;
}

// ------------------------------------------------------------ chimera ----
// Three-point moment transforms along one axis (ChimeraTransformation.h:58-121).
template <typename T> __device__ __forceinline__ void fwd_k(T& a, T& b, T& c, T vv, T v2, T kinv, T k)
{
    const T s = a + c;
    const T d = c - a;
    T m0 = s + b;
    a = m0;
    m0 *= kinv;
    m0 += T(1);
    b = (d * kinv - m0 * vv) * k;
    c = ((s - T(2) * d * vv) * kinv + v2 * m0) * k;
}
template <typename T> __device__ __forceinline__ void bwd_k(T& a, T& b, T& c, T vv, T v2, T kinv, T k)
{
    const T h = T(0.5);
    const T r0 = (((c - b) * h + b * vv) * kinv + (a * kinv + T(1)) * (v2 - vv) * h) * k;
    const T r1 = (((a - c) - T(2) * b * vv) * kinv + (a * kinv + T(1)) * (-v2)) * k;
    c = (((c + b) * h + b * vv) * kinv + (a * kinv + T(1)) * (v2 + vv) * h) * k;
    a = r0;
    b = r1;
}
template <typename T> __device__ __forceinline__ void fwd(T& a, T& b, T& c, T vv, T v2)
{
    const T s = (a + c) + b;
    const T d = c - a;
    c = (c + a) + (v2 * s - T(2) * vv * d);
    b = d - vv * s;
    a = s;
}
template <typename T> __device__ __forceinline__ void bwd(T& a, T& b, T& c, T vv, T v2)
{
    const T h = T(0.5);
    const T ra = (c + a * (v2 - vv)) * h + b * (vv - h);
    const T rb = ((a - c) - a * v2) - T(2) * b * vv;
    c = (c + a * (v2 + vv)) * h + b * (vv + h);
    b = rb;
    a = ra;
}

// chimera constants of one row; the two axes not being transformed contribute
// M/P: 1/6, 0: 2/3 before their own sweep and order 0: 1, order 2: 1/3 after it
template <typename T> struct KPair { T kinv, k; };

#define VFX_M(a, b, c) m[cube(a, b, c)]

// density deviation and momentum of cube-ordered populations in the
// reference's pyramid order (MacroscopicQuantities.h:55-85,117-127)
template <typename T>
__device__ __forceinline__ void pyramid_sums(const T (&m)[27], T& drho, T& jx, T& jy, T& jz)
{
#define F(d) m[slot_of(d)]
    const T corners = ((F(dPPP) + F(dMMM)) + (F(dMPM) + F(dPMP))) + ((F(dMPP) + F(dPMM)) + (F(dMMP) + F(dPPM)));
    const T edges = (((F(d0MP) + F(d0PM)) + (F(d0MM) + F(d0PP))) + ((F(dM0P) + F(dP0M)) + (F(dM0M) + F(dP0P)))) +
                    ((F(dMP0) + F(dPM0)) + (F(dMM0) + F(dPP0)));
    const T faces = ((F(dM00) + F(dP00)) + (F(d0M0) + F(d0P0))) + (F(d00M) + F(d00P));
    drho = ((corners + edges) + F(d000)) + faces;
    jx = (((F(dPPP) - F(dMMM)) + (F(dPMP) - F(dMPM))) + ((F(dPMM) - F(dMPP)) + (F(dPPM) - F(dMMP)))) +
         (((F(dP0M) - F(dM0P)) + (F(dP0P) - F(dM0M))) + ((F(dPM0) - F(dMP0)) + (F(dPP0) - F(dMM0)))) +
         (F(dP00) - F(dM00));
    jy = (((F(dPPP) - F(dMMM)) + (F(dMPM) - F(dPMP))) + ((F(dMPP) - F(dPMM)) + (F(dPPM) - F(dMMP)))) +
         (((F(d0PM) - F(d0MP)) + (F(d0PP) - F(d0MM))) + ((F(dMP0) - F(dPM0)) + (F(dPP0) - F(dMM0)))) +
         (F(d0P0) - F(d0M0));
    jz = (((F(dPPP) - F(dMMM)) + (F(dPMP) - F(dMPM))) + ((F(dMPP) - F(dPMM)) + (F(dMMP) - F(dPPM)))) +
         (((F(d0MP) - F(d0PM)) + (F(d0PP) - F(d0MM))) + ((F(dM0P) - F(dP0M)) + (F(dP0P) - F(dM0M)))) +
         (F(d00P) - F(d00M));
#undef F
}

// central moments (orders (i,j,k) at cube(i,j,k)) -> cube-ordered shifted
// populations: backward chimera sweeps x, y, z with the well-conditioning
// constants of K17CompressibleNavierStokes.h:525-557 (the same sweeps close
// InterpolationCF.h:255-293 and InterpolationFC.h:210-248)
template <typename T> __device__ __forceinline__ void backward_sweeps(T (&m)[27], T vx, T vy, T vz)
{
    const T vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (b == 1 || c == 1) {
                bwd(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), vx, vx2);
            } else {
                const T kinv = (b == 0 && c == 0) ? T(1) : (b == 2 && c == 2) ? T(9) : T(3);
                const T k = (b == 0 && c == 0) ? T(1) : (b == 2 && c == 2) ? T(1. / 9.) : T(1. / 3.);
                bwd_k(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), vx, vx2, kinv, k);
            }
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (c == 1) {
                bwd(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), vy, vy2);
            } else {
                const T kinv = (i == 1) ? (c == 0 ? T(3. / 2.) : T(9. / 2.)) : (c == 0 ? T(6) : T(18));
                const T k = (i == 1) ? (c == 0 ? T(2. / 3.) : T(2. / 9.)) : (c == 0 ? T(1. / 6.) : T(1. / 18.));
                bwd_k(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), vy, vy2, kinv, k);
            }
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T kinv = (i == 1 && j == 1) ? T(9. / 4.) : (i == 1 || j == 1) ? T(9) : T(36);
            const T k = (i == 1 && j == 1) ? T(4. / 9.) : (i == 1 || j == 1) ? T(1. / 9.) : T(1. / 36.);
            bwd_k(VFX_M(i, j, 0), VFX_M(i, j, 1), VFX_M(i, j, 2), vz, vz2, kinv, k);
        }
}

// ------------------------------------------------------------------ K17 ----
// m[27]: populations in cube order on entry, post-collision populations on exit.
// Returns drho and the (force-shifted) velocity in mv[4].
// Eddy-viscosity models of the K17 kernel (vf::lbm::TurbulenceModel,
// src/lbm/collision/TurbulentViscosity.h): the node relaxes with
// omega / (1 + 3 omega nu_t) (calculateOmegaWithturbulentViscosity), and the
// Smagorinsky and QR models replace nu_t by their value from the new velocity
// gradients (K17CompressibleNavierStokes.h:245, 355-368).  AMD's nu_t comes
// from a separate kernel between steps.
enum TurbModel : int { kTurbNone = 0, kTurbSmagorinsky = 1, kTurbAMD = 2, kTurbQR = 3 };

template <typename T> __device__ __forceinline__ T turbulent_viscosity_smagorinsky(T Cs, T dxux, T dyuy, T dzuz, T Dxy,
                                                                                   T Dxz, T Dyz)
{
    return Cs * Cs * sqrt(T(2) * (dxux * dxux + dyuy * dyuy + dzuz * dzuz) + Dxy * Dxy + Dxz * Dxz + Dyz * Dyz);
}
// Verstappen's QR model: C max(R, 0) / Q with the second and third invariants
// of the strain rate (0/0 at rest, as in the reference)
template <typename T>
__device__ __forceinline__ T turbulent_viscosity_qr(T C, T dxux, T dyuy, T dzuz, T Dxy, T Dxz, T Dyz)
{
    const T Q = T(0.5) * (dxux * dxux + dyuy * dyuy + dzuz * dzuz) + T(0.25) * (Dxy * Dxy + Dxz * Dxz + Dyz * Dyz);
    const T Rr = -dxux * dyuy * dzuz + T(0.25) * (-Dxy * Dxz * Dyz + dxux * Dyz * Dyz + dyuy * Dxz * Dxz + dzuz * Dxy * Dxy);
    return C * (Rr > T(0) ? Rr : T(0)) / Q;
}

template <typename T, int TM = kTurbNone>
__device__ __forceinline__ void k17(T (&m)[27], const Rates<T>& R0, T (&mv)[4], T& nut)
{
    Rates<T> Rt;
    if constexpr (TM != kTurbNone) {
        Rt = R0;
        omega_rates<T>(R0.omega / (T(1) + T(3) * R0.omega * nut), Rt);
    }
    const Rates<T>& R = TM != kTurbNone ? Rt : R0;
    constexpr T c1o3 = T(1. / 3.), c2o3 = T(2. / 3.), c1o9 = T(1. / 9.), c1o27 = T(1. / 27.), h = T(0.5);
    constexpr T c3o2 = T(3. / 2.);

    // macroscopic values, pyramid summation (MacroscopicQuantities.h:55-85,117-127)
    T drho, vx, vy, vz;
    pyramid_sums(m, drho, vx, vy, vz);
    const T oneOverRho = T(1) / (drho + T(1));
    vx *= oneOverRho;
    vy *= oneOverRho;
    vz *= oneOverRho;
    vx += R.fx;
    vy += R.fy;
    vz += R.fz;
    const T vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;

    // forward sweeps z, y, x (K17CompressibleNavierStokes.h:190-222)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T kinv = (i == 1 && j == 1) ? T(9. / 4.) : (i == 1 || j == 1) ? T(9) : T(36);
            const T k = (i == 1 && j == 1) ? T(4. / 9.) : (i == 1 || j == 1) ? T(1. / 9.) : T(1. / 36.);
            fwd_k(VFX_M(i, j, 0), VFX_M(i, j, 1), VFX_M(i, j, 2), vz, vz2, kinv, k);
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (c == 1) {
                fwd(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), vy, vy2);
            } else {
                const T kinv = (i == 1) ? (c == 0 ? T(3. / 2.) : T(9. / 2.)) : (c == 0 ? T(6) : T(18));
                const T k = (i == 1) ? (c == 0 ? T(2. / 3.) : T(2. / 9.)) : (c == 0 ? T(1. / 6.) : T(1. / 18.));
                fwd_k(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), vy, vy2, kinv, k);
            }
        }
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (b == 1 || c == 1) {
                fwd(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), vx, vx2);
            } else {
                // (b,c)=(2,2) uses K^-1 = 3 with K = 1/9 exactly as the reference (:222)
                const T kinv = (b == 0 && c == 0) ? T(1) : T(3);
                const T k = (b == 0 && c == 0) ? T(1) : (b == 2 && c == 2) ? T(1. / 9.) : T(1. / 3.);
                fwd_k(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), vx, vx2, kinv, k);
            }
        }

    T& m000 = VFX_M(0, 0, 0); T& m100 = VFX_M(1, 0, 0); T& m010 = VFX_M(0, 1, 0); T& m001 = VFX_M(0, 0, 1);
    T& m200 = VFX_M(2, 0, 0); T& m020 = VFX_M(0, 2, 0); T& m002 = VFX_M(0, 0, 2);
    T& m110 = VFX_M(1, 1, 0); T& m101 = VFX_M(1, 0, 1); T& m011 = VFX_M(0, 1, 1);
    T& m111 = VFX_M(1, 1, 1);
    T& m210 = VFX_M(2, 1, 0); T& m012 = VFX_M(0, 1, 2); T& m201 = VFX_M(2, 0, 1);
    T& m021 = VFX_M(0, 2, 1); T& m120 = VFX_M(1, 2, 0); T& m102 = VFX_M(1, 0, 2);
    T& m211 = VFX_M(2, 1, 1); T& m121 = VFX_M(1, 2, 1); T& m112 = VFX_M(1, 1, 2);
    T& m220 = VFX_M(2, 2, 0); T& m202 = VFX_M(2, 0, 2); T& m022 = VFX_M(0, 2, 2);
    T& m122 = VFX_M(1, 2, 2); T& m212 = VFX_M(2, 1, 2); T& m221 = VFX_M(2, 2, 1);
    T& m222 = VFX_M(2, 2, 2);

    const T omega = R.omega;
    const T OxxPyyPzz = T(1), O4 = T(1), O5 = T(1), O6 = T(1);

    // cumulants of order 4..6 (:282-316)
    T cm211 = m211 - ((m200 + c1o3) * m011 + T(2) * m110 * m101) * oneOverRho;
    T cm121 = m121 - ((m020 + c1o3) * m101 + T(2) * m110 * m011) * oneOverRho;
    T cm112 = m112 - ((m002 + c1o3) * m110 + T(2) * m101 * m011) * oneOverRho;
    T cm220 = m220 - (((m200 * m020 + T(2) * m110 * m110) + c1o3 * (m200 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm202 = m202 - (((m200 * m002 + T(2) * m101 * m101) + c1o3 * (m200 + m002)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm022 = m022 - (((m002 * m020 + T(2) * m011 * m011) + c1o3 * (m002 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    T cm122 = m122 - ((m002 * m120 + m020 * m102 + T(4) * m011 * m111 + T(2) * (m101 * m021 + m110 * m012)) +
                      c1o3 * (m120 + m102)) * oneOverRho;
    T cm212 = m212 - ((m002 * m210 + m200 * m012 + T(4) * m101 * m111 + T(2) * (m011 * m201 + m110 * m102)) +
                      c1o3 * (m210 + m012)) * oneOverRho;
    T cm221 = m221 - ((m200 * m021 + m020 * m201 + T(4) * m110 * m111 + T(2) * (m101 * m120 + m011 * m210)) +
                      c1o3 * (m021 + m201)) * oneOverRho;
    T cm222 = m222 + ((-T(4) * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                       T(4) * (m011 * m211 + m101 * m121 + m110 * m112) -
                       T(2) * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                      (T(4) * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                       T(2) * (m200 * m020 * m002) + T(16) * m110 * m101 * m011) * oneOverRho * oneOverRho -
                      c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                      (T(2) * (m101 * m101 + m011 * m011 + m110 * m110) +
                       (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                          oneOverRho * oneOverRho * c2o3 +
                      c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    // second and third order linear combinations (:323-335)
    T mxxPyyPzz = m200 + m020 + m002;
    T mxxMyy = m200 - m020;
    T mxxMzz = m200 - m002;
    T mxxyPyzz = m210 + m012, mxxyMyzz = m210 - m012;
    T mxxzPyyz = m201 + m021, mxxzMyyz = m201 - m021;
    T mxyyPxzz = m120 + m102, mxyyMxzz = m120 - m102;

    // velocity gradients times rho (:347-352)
    const T Dxy = -T(3) * omega * m110;
    const T Dxz = -T(3) * omega * m101;
    const T Dyz = -T(3) * omega * m011;
    const T dxux = h * (-omega) * (mxxMyy + mxxMzz) + h * OxxPyyPzz * (m000 - mxxPyyPzz);
    const T dyuy = dxux + omega * c3o2 * mxxMyy;
    const T dzuz = dxux + omega * c3o2 * mxxMzz;
    if constexpr (TM == kTurbSmagorinsky) nut = turbulent_viscosity_smagorinsky<T>(R.sgs, dxux, dyuy, dzuz, Dxy, Dxz, Dyz);
    if constexpr (TM == kTurbQR) nut = turbulent_viscosity_qr<T>(R.sgs, dxux, dyuy, dzuz, Dxy, Dxz, Dyz);

    // second order with Galilean correction (:375-387)
    mxxPyyPzz += OxxPyyPzz * (m000 - mxxPyyPzz) - T(3) * (T(1) - h * OxxPyyPzz) * (vx2 * dxux + vy2 * dyuy + vz2 * dzuz);
    mxxMyy += omega * (-mxxMyy) - T(3) * (T(1) + h * (-omega)) * (vx2 * dxux - vy2 * dyuy);
    mxxMzz += omega * (-mxxMzz) - T(3) * (T(1) + h * (-omega)) * (vx2 * dxux - vz2 * dzuz);
    m011 += omega * (-m011);
    m101 += omega * (-m101);
    m110 += omega * (-m110);

    // limited third order (:397-410)
    T wadj = R.Oxyz + (T(1) - R.Oxyz) * fabs(m111) / (fabs(m111) + R.limD);
    m111 += wadj * (-m111);
    wadj = R.OxyyPxzz + (T(1) - R.OxyyPxzz) * fabs(mxxyPyzz) / (fabs(mxxyPyzz) + R.limP);
    mxxyPyzz += wadj * (-mxxyPyzz);
    wadj = R.OxyyMxzz + (T(1) - R.OxyyMxzz) * fabs(mxxyMyzz) / (fabs(mxxyMyzz) + R.limM);
    mxxyMyzz += wadj * (-mxxyMyzz);
    wadj = R.OxyyPxzz + (T(1) - R.OxyyPxzz) * fabs(mxxzPyyz) / (fabs(mxxzPyyz) + R.limP);
    mxxzPyyz += wadj * (-mxxzPyyz);
    wadj = R.OxyyMxzz + (T(1) - R.OxyyMxzz) * fabs(mxxzMyyz) / (fabs(mxxzMyyz) + R.limM);
    mxxzMyyz += wadj * (-mxxzMyyz);
    wadj = R.OxyyPxzz + (T(1) - R.OxyyPxzz) * fabs(mxyyPxzz) / (fabs(mxyyPxzz) + R.limP);
    mxyyPxzz += wadj * (-mxyyPxzz);
    wadj = R.OxyyMxzz + (T(1) - R.OxyyMxzz) * fabs(mxyyMxzz) / (fabs(mxyyMxzz) + R.limM);
    mxyyMxzz += wadj * (-mxyyMxzz);

    // back from linear combinations (:424-433)
    m200 = c1o3 * (mxxMyy + mxxMzz + mxxPyyPzz);
    m020 = c1o3 * (-T(2) * mxxMyy + mxxMzz + mxxPyyPzz);
    m002 = c1o3 * (mxxMyy - T(2) * mxxMzz + mxxPyyPzz);
    m210 = (mxxyMyzz + mxxyPyzz) * h;
    m012 = (-mxxyMyzz + mxxyPyzz) * h;
    m201 = (mxxzMyyz + mxxzPyyz) * h;
    m021 = (-mxxzMyyz + mxxzPyyz) * h;
    m120 = (mxyyMxzz + mxyyPxzz) * h;
    m102 = (-mxyyMxzz + mxyyPxzz) * h;

    // fourth order to the modified equilibrium, 5th/6th to zero (:443-459)
    const T invOmH = T(1) / omega - h;
    cm022 = -O4 * invOmH * (dyuy + dzuz) * c2o3 * R.factorA + (T(1) - O4) * (cm022);
    cm202 = -O4 * invOmH * (dxux + dzuz) * c2o3 * R.factorA + (T(1) - O4) * (cm202);
    cm220 = -O4 * invOmH * (dyuy + dxux) * c2o3 * R.factorA + (T(1) - O4) * (cm220);
    cm112 = -O4 * invOmH * Dxy * c1o3 * R.factorB + (T(1) - O4) * (cm112);
    cm121 = -O4 * invOmH * Dxz * c1o3 * R.factorB + (T(1) - O4) * (cm121);
    cm211 = -O4 * invOmH * Dyz * c1o3 * R.factorB + (T(1) - O4) * (cm211);
    cm122 += O5 * (-cm122);
    cm212 += O5 * (-cm212);
    cm221 += O5 * (-cm221);
    cm222 += O6 * (-cm222);

    // central moments from post-collision cumulants (:469-505)
    m211 = cm211 + c1o3 * ((T(3) * m200 + T(1)) * m011 + T(6) * m110 * m101) * oneOverRho;
    m121 = cm121 + c1o3 * ((T(3) * m020 + T(1)) * m101 + T(6) * m110 * m011) * oneOverRho;
    m112 = cm112 + c1o3 * ((T(3) * m002 + T(1)) * m110 + T(6) * m101 * m011) * oneOverRho;
    m220 = cm220 + (((m200 * m020 + T(2) * m110 * m110) * T(9) + T(3) * (m200 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m202 = cm202 + (((m200 * m002 + T(2) * m101 * m101) * T(9) + T(3) * (m200 + m002)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m022 = cm022 + (((m002 * m020 + T(2) * m011 * m011) * T(9) + T(3) * (m002 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m122 = cm122 + c1o3 * (T(3) * (m002 * m120 + m020 * m102 + T(4) * m011 * m111 + T(2) * (m101 * m021 + m110 * m012)) +
                           (m120 + m102)) * oneOverRho;
    m212 = cm212 + c1o3 * (T(3) * (m002 * m210 + m200 * m012 + T(4) * m101 * m111 + T(2) * (m011 * m201 + m110 * m102)) +
                           (m210 + m012)) * oneOverRho;
    m221 = cm221 + c1o3 * (T(3) * (m200 * m021 + m020 * m201 + T(4) * m110 * m111 + T(2) * (m101 * m120 + m011 * m210)) +
                           (m021 + m201)) * oneOverRho;
    m222 = cm222 - ((-T(4) * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                     T(4) * (m011 * m211 + m101 * m121 + m110 * m112) -
                     T(2) * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                    (T(4) * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                     T(2) * (m200 * m020 * m002) + T(16) * m110 * m101 * m011) * oneOverRho * oneOverRho -
                    c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                    (T(2) * (m101 * m101 + m011 * m011 + m110 * m110) +
                     (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                        oneOverRho * oneOverRho * c2o3 +
                    c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    m100 = -m100;
    m010 = -m010;
    m001 = -m001;

    backward_sweeps(m, vx, vy, vz);

    mv[0] = drho;
    mv[1] = vx;
    mv[2] = vy;
    mv[3] = vz;
}

// ------------------------------------------------------------------ BGK ----
// Compressible BGK on shifted populations, B92CompressibleNavierStokes_Device.cu:189-234.
template <typename T> __device__ __forceinline__ void bgk(T (&m)[27], T omega, T (&mv)[4])
{
#define F(d) m[slot_of(d)]
    const T drho = (F(dPPP) + F(dMMM)) + (F(dMMP) + F(dPPM)) + (F(dPMP) + F(dMPM)) + (F(dMPP) + F(dPMM)) +
                   (F(dPP0) + F(dMM0)) + (F(dMP0) + F(dPM0)) + (F(dP0P) + F(dM0M)) + (F(dP0M) + F(dM0P)) +
                   (F(d0PP) + F(d0MM)) + (F(d0PM) + F(d0MP)) + (F(dP00) + F(dM00)) + (F(d0P0) + F(d0M0)) +
                   (F(d00P) + F(d00M)) + F(d000);
    const T rho = drho + T(1);
    const T oor = T(1) / rho;
    const T vx1 = oor * ((F(dPPP) - F(dMMM)) + (F(dPPM) - F(dMMP)) + (F(dPMP) - F(dMPM)) + (F(dPMM) - F(dMPP)) +
                         (F(dPP0) - F(dMM0)) + (F(dPM0) - F(dMP0)) + (F(dP0P) - F(dM0M)) + (F(dP0M) - F(dM0P)) +
                         (F(dP00) - F(dM00)));
    const T vx2 = oor * ((F(dPPP) - F(dMMM)) + (F(dPPM) - F(dMMP)) + (F(dMPM) - F(dPMP)) + (F(dMPP) - F(dPMM)) +
                         (F(dPP0) - F(dMM0)) + (F(dMP0) - F(dPM0)) + (F(d0PP) - F(d0MM)) + (F(d0PM) - F(d0MP)) +
                         (F(d0P0) - F(d0M0)));
    const T vx3 = oor * ((F(dPPP) - F(dMMM)) + (F(dMMP) - F(dPPM)) + (F(dPMP) - F(dMPM)) + (F(dMPP) - F(dPMM)) +
                         (F(dP0P) - F(dM0M)) + (F(dM0P) - F(dP0M)) + (F(d0PP) - F(d0MM)) + (F(d0MP) - F(d0PM)) +
                         (F(d00P) - F(d00M)));
    const T cusq = T(1.5) * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
    const T a = T(1) + (-omega);
    F(d000) = F(d000) * a - (-omega) * T(8. / 27.) * (drho - rho * cusq);
#pragma unroll
    for (int d = 1; d < 27; ++d) {
        // projected velocity with the reference's left-to-right component order
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
        m[slot_of(d)] = m[slot_of(d)] * a - (-omega) * weight<T>(d) * (drho + rho * (T(3) * cu + T(4.5) * cu * cu - cusq));
    }
#undef F
    mv[0] = drho;
    mv[1] = vx1;
    mv[2] = vx2;
    mv[3] = vx3;
}

#undef VFX_M

// ------------------------------------------------------ K17, reduced form ----
// The same collision as k17<T> with the algebra folded for FP32 throughput:
//  * the pre-collision cumulants of order 4..6 are never formed: K17 relaxes
//    them with rate 1 (O4 = O5 = O6 = 1, K17CompressibleNavierStokes.h:258-264),
//    so their post-collision values do not depend on them, and the forward
//    transform skips every moment of total order >= 4;
//  * chimera constants enter as K only (K^-1 * K = 1), as in the reference's
//    forwardChimeraWithK/backwardChimeraWithK (ChimeraTransformation.h:124-147);
//  * omega-only factors are folded on the host (Rates::om1, c3, A4, B4, lim*1)
//    and the limiter becomes x (1-O) lambda / (|x| + lambda);
//  * divisions use the hardware reciprocal (1 ulp).
// Macroscopic sums are the partial sums of the forward sweeps (z rows, then
// y, then x), which are as well conditioned as the pyramid sums.  Not bit-identical to the reference order;
// tested against the FP64 oracle by error norms.
__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// forward three-point transform (k = 0: no weight constant); `full` = 2 also
// forms the second-order output, 1 only the first, 0 only the sum
template <int OUT>
__device__ __forceinline__ void fwdf(float& a, float& b, float& c, float vv, float v2, float tv, float k)
{
    const float s = a + c;
    const float d = c - a;
    const float m0 = s + b;
    const float mk = k == 0.f ? m0 : m0 + k; // k is a compile-time constant after unrolling
    a = m0;
    if (OUT >= 1) b = d - mk * vv;
    if (OUT >= 2) c = (s - tv * d) + mk * v2;
}
// the same transform with the sums s = a + c, d = c - a, m0 = s + b already formed
template <int OUT>
__device__ __forceinline__ void fwdf_sums(float& a, float& b, float& c, float s, float d, float m0, float vv, float v2,
                                          float tv, float k)
{
    const float mk = k == 0.f ? m0 : m0 + k;
    a = m0;
    if (OUT >= 1) b = d - mk * vv;
    if (OUT >= 2) c = (s - tv * d) + mk * v2;
}
// backward three-point transform; hvm = (v2 - vv)/2, hvp = (v2 + vv)/2,
// vmh = vv - 1/2, vph = vv + 1/2.  The centre output uses r0 + r1 + r2 = a
// (the zeroth moment): r1 = a - (r0 + r2), two operations instead of four.
__device__ __forceinline__ void bwdf(float& a, float& b, float& c, float k, float hvm, float hvp, float vmh,
                                     float vph)
{
    const float ak = k == 0.f ? a : a + k; // k is a compile-time constant after unrolling
    const float ch = 0.5f * c;
    const float r0 = ch + b * vmh + ak * hvm;
    const float r2 = ch + b * vph + ak * hvp;
    const float r1 = a - (r0 + r2);
    a = r0;
    b = r1;
    c = r2;
}

// The same transform returning the outer outputs doubled, (2 r0, r1, 2 r2):
// with hv2m = v2 - vv, hv2p = v2 + vv, v2mh = 2 vv - 1, v2ph = 2 vv + 1 the
// factor 1/2 of the outer rows disappears and the centre output becomes
// a - (2 r0 + 2 r2)/2.  Every value is exactly twice the one bwdf forms
// (scaling by 2 commutes with rounding), so a sweep over doubled inputs, with
// its weight constant doubled too, is exact bookkeeping: after x, y and z a
// population of direction d comes out as 2^n f_d, n = number of non-zero
// components of c_d, and the store folds 2^-n into its scale factor.
__device__ __forceinline__ void bwdf2(float& a, float& b, float& c, float k, float hv2m, float hv2p, float v2mh,
                                      float v2ph)
{
    const float ak = k == 0.f ? a : a + k; // k is a compile-time constant after unrolling
    const float r0 = c + b * v2mh + ak * hv2m;
    const float r2 = c + b * v2ph + ak * hv2p;
    const float r1 = a - 0.5f * (r0 + r2);
    a = r0;
    b = r1;
    c = r2;
}

#define VFX_M(a, b, c) m[cube(a, b, c)]

// OUTER2: run the backward sweeps with bwdf2, so that population d leaves
// scaled by 2^weight_class(d) (the caller's store removes the factor; used
// where the store multiplies by a scale anyway, i.e. FP16 storage).
template <bool OUTER2>
__device__ __forceinline__ void k17_fast(float (&m)[27], const Rates<float>& R, float (&mv)[4])
{
    constexpr float c1o3 = 1.f / 3.f, c1o9 = 1.f / 9.f, c1o27 = 1.f / 27.f, c2o3 = 2.f / 3.f;
    // ---- macroscopic values from the sums the sweeps form anyway.  Per (x,y)
    // row of the z sweep: s = f(z-) + f(z+), d = f(z+) - f(z-), m0 = s + f(z0).
    // The zeroth-moment rows enter the y sweep unchanged (their sums do not
    // depend on the velocity), so its sums ys, yd, ym per x plane are formed
    // here too: drho = (ym0 + ym2) + ym1 and jx = ym2 - ym0 are the sums of the
    // x sweep's first line, jy = sum of yd; jz sums the nine z differences.
    float zs[3][3], zd[3][3], zm[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            zs[i][j] = VFX_M(i, j, 0) + VFX_M(i, j, 2);
            zd[i][j] = VFX_M(i, j, 2) - VFX_M(i, j, 0);
            zm[i][j] = zs[i][j] + VFX_M(i, j, 1);
        }
    float ys[3], yd[3], ym[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ys[i] = zm[i][0] + zm[i][2];
        yd[i] = zm[i][2] - zm[i][0];
        ym[i] = ys[i] + zm[i][1];
    }
    const float xs = ym[0] + ym[2];
    const float jx = ym[2] - ym[0];
    const float drho = xs + ym[1];
    const float jy = (yd[0] + yd[2]) + yd[1];
    const float jz = (((zd[0][0] + zd[2][2]) + (zd[0][2] + zd[2][0])) + ((zd[0][1] + zd[2][1]) + (zd[1][0] + zd[1][2]))) + zd[1][1];
    const float oneOverRho = rcpf(drho + 1.f);
    const float vx = jx * oneOverRho + R.fx;
    const float vy = jy * oneOverRho + R.fy;
    const float vz = jz * oneOverRho + R.fz;
    const float vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const float tx = 2.f * vx, ty = 2.f * vy, tz = 2.f * vz;

    // ---- forward: z, y, x; moments of total order >= 4 are not formed
    constexpr float wv[3] = {1.f / 6.f, 2.f / 3.f, 1.f / 6.f};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float mk = zm[i][j] + wv[i] * wv[j];
            VFX_M(i, j, 0) = zm[i][j];
            VFX_M(i, j, 1) = zd[i][j] - mk * vz;
            VFX_M(i, j, 2) = (zs[i][j] - tz * zd[i][j]) + mk * vz2;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        fwdf_sums<2>(VFX_M(i, 0, 0), VFX_M(i, 1, 0), VFX_M(i, 2, 0), ys[i], yd[i], ym[i], vy, vy2, ty, wv[i]);
        fwdf<2>(VFX_M(i, 0, 1), VFX_M(i, 1, 1), VFX_M(i, 2, 1), vy, vy2, ty, 0.f);
        fwdf<1>(VFX_M(i, 0, 2), VFX_M(i, 1, 2), VFX_M(i, 2, 2), vy, vy2, ty, wv[i] * c1o3);
    }
    fwdf_sums<2>(VFX_M(0, 0, 0), VFX_M(1, 0, 0), VFX_M(2, 0, 0), xs, jx, drho, vx, vx2, tx, 1.f);
    fwdf<2>(VFX_M(0, 1, 0), VFX_M(1, 1, 0), VFX_M(2, 1, 0), vx, vx2, tx, 0.f);
    fwdf<2>(VFX_M(0, 0, 1), VFX_M(1, 0, 1), VFX_M(2, 0, 1), vx, vx2, tx, 0.f);
    fwdf<1>(VFX_M(0, 2, 0), VFX_M(1, 2, 0), VFX_M(2, 2, 0), vx, vx2, tx, c1o3);
    fwdf<1>(VFX_M(0, 0, 2), VFX_M(1, 0, 2), VFX_M(2, 0, 2), vx, vx2, tx, c1o3);
    fwdf<1>(VFX_M(0, 1, 1), VFX_M(1, 1, 1), VFX_M(2, 1, 1), vx, vx2, tx, 0.f);
    fwdf<0>(VFX_M(0, 1, 2), VFX_M(1, 1, 2), VFX_M(2, 1, 2), vx, vx2, tx, 0.f);
    fwdf<0>(VFX_M(0, 2, 1), VFX_M(1, 2, 1), VFX_M(2, 2, 1), vx, vx2, tx, 0.f);

    float& m000 = VFX_M(0, 0, 0); float& m100 = VFX_M(1, 0, 0); float& m010 = VFX_M(0, 1, 0); float& m001 = VFX_M(0, 0, 1);
    float& m200 = VFX_M(2, 0, 0); float& m020 = VFX_M(0, 2, 0); float& m002 = VFX_M(0, 0, 2);
    float& m110 = VFX_M(1, 1, 0); float& m101 = VFX_M(1, 0, 1); float& m011 = VFX_M(0, 1, 1);
    float& m111 = VFX_M(1, 1, 1);
    float& m210 = VFX_M(2, 1, 0); float& m012 = VFX_M(0, 1, 2); float& m201 = VFX_M(2, 0, 1);
    float& m021 = VFX_M(0, 2, 1); float& m120 = VFX_M(1, 2, 0); float& m102 = VFX_M(1, 0, 2);
    float& m211 = VFX_M(2, 1, 1); float& m121 = VFX_M(1, 2, 1); float& m112 = VFX_M(1, 1, 2);
    float& m220 = VFX_M(2, 2, 0); float& m202 = VFX_M(2, 0, 2); float& m022 = VFX_M(0, 2, 2);
    float& m122 = VFX_M(1, 2, 2); float& m212 = VFX_M(2, 1, 2); float& m221 = VFX_M(2, 2, 1);
    float& m222 = VFX_M(2, 2, 2);
    const float omega = R.omega;

    // ---- second order (:323-387)
    float mxxPyyPzz = m200 + m020 + m002;
    float mxxMyy = m200 - m020;
    float mxxMzz = m200 - m002;
    const float Dxy = -3.f * omega * m110;
    const float Dxz = -3.f * omega * m101;
    const float Dyz = -3.f * omega * m011;
    const float dxux = -0.5f * omega * (mxxMyy + mxxMzz) + 0.5f * (m000 - mxxPyyPzz);
    const float dyuy = dxux + 1.5f * omega * mxxMyy;
    const float dzuz = dxux + 1.5f * omega * mxxMzz;
    mxxPyyPzz = m000 - 1.5f * (vx2 * dxux + vy2 * dyuy + vz2 * dzuz);
    mxxMyy = R.om1 * mxxMyy - R.c3 * (vx2 * dxux - vy2 * dyuy);
    mxxMzz = R.om1 * mxxMzz - R.c3 * (vx2 * dxux - vz2 * dzuz);
    m011 *= R.om1;
    m101 *= R.om1;
    m110 *= R.om1;

    // ---- third order with limiter: x (1-O) lambda / (|x| + lambda) (:397-410)
    float mxxyPyzz = m210 + m012, mxxyMyzz = m210 - m012;
    float mxxzPyyz = m201 + m021, mxxzMyyz = m201 - m021;
    float mxyyPxzz = m120 + m102, mxyyMxzz = m120 - m102;
    m111 *= R.limD1 * rcpf(fabsf(m111) + R.limD);
    mxxyPyzz *= R.limP1 * rcpf(fabsf(mxxyPyzz) + R.limP);
    mxxyMyzz *= R.limM1 * rcpf(fabsf(mxxyMyzz) + R.limM);
    mxxzPyyz *= R.limP1 * rcpf(fabsf(mxxzPyyz) + R.limP);
    mxxzMyyz *= R.limM1 * rcpf(fabsf(mxxzMyyz) + R.limM);
    mxyyPxzz *= R.limP1 * rcpf(fabsf(mxyyPxzz) + R.limP);
    mxyyMxzz *= R.limM1 * rcpf(fabsf(mxyyMxzz) + R.limM);

    m200 = c1o3 * (mxxMyy + mxxMzz + mxxPyyPzz);
    m020 = c1o3 * (-2.f * mxxMyy + mxxMzz + mxxPyyPzz);
    m002 = c1o3 * (mxxMyy - 2.f * mxxMzz + mxxPyyPzz);
    m210 = 0.5f * (mxxyMyzz + mxxyPyzz);
    m012 = 0.5f * (-mxxyMyzz + mxxyPyzz);
    m201 = 0.5f * (mxxzMyyz + mxxzPyyz);
    m021 = 0.5f * (-mxxzMyyz + mxxzPyyz);
    m120 = 0.5f * (mxyyMxzz + mxyyPxzz);
    m102 = 0.5f * (-mxyyMxzz + mxyyPxzz);

    // ---- fourth order at the modified equilibrium, fifth/sixth cumulants zero (:443-505)
    const float R1 = oneOverRho;
    const float dr9 = c1o9 * drho;
    m211 = -R.B4 * Dyz + ((m200 + c1o3) * m011 + 2.f * m110 * m101) * R1;
    m121 = -R.B4 * Dxz + ((m020 + c1o3) * m101 + 2.f * m110 * m011) * R1;
    m112 = -R.B4 * Dxy + ((m002 + c1o3) * m110 + 2.f * m101 * m011) * R1;
    m220 = -R.A4 * (dyuy + dxux) + ((m200 * m020 + 2.f * m110 * m110) + c1o3 * (m200 + m020) - dr9) * R1;
    m202 = -R.A4 * (dxux + dzuz) + ((m200 * m002 + 2.f * m101 * m101) + c1o3 * (m200 + m002) - dr9) * R1;
    m022 = -R.A4 * (dyuy + dzuz) + ((m002 * m020 + 2.f * m011 * m011) + c1o3 * (m002 + m020) - dr9) * R1;
    m122 = ((m002 * m120 + m020 * m102 + 4.f * m011 * m111 + 2.f * (m101 * m021 + m110 * m012)) + c1o3 * (m120 + m102)) * R1;
    m212 = ((m002 * m210 + m200 * m012 + 4.f * m101 * m111 + 2.f * (m011 * m201 + m110 * m102)) + c1o3 * (m210 + m012)) * R1;
    m221 = ((m200 * m021 + m020 * m201 + 4.f * m110 * m111 + 2.f * (m101 * m120 + m011 * m210)) + c1o3 * (m021 + m201)) * R1;
    const float R2 = R1 * R1;
    const float s2 = m002 + m020 + m200;
    m222 = -((-4.f * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
              4.f * (m011 * m211 + m101 * m121 + m110 * m112) - 2.f * (m120 * m102 + m210 * m012 + m201 * m021)) * R1 +
             (4.f * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) + 2.f * (m200 * m020 * m002) +
              16.f * m110 * m101 * m011) * R2 -
             c1o3 * (m022 + m202 + m220) * R1 - c1o9 * s2 * R1 +
             (2.f * (m101 * m101 + m011 * m011 + m110 * m110) + (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * s2) *
                 R2 * c2o3 +
             c1o27 * ((drho * drho - drho) * R2));

    m100 = -m100;
    m010 = -m010;
    m001 = -m001;

    // ---- backward: x, y, z
    if constexpr (OUTER2) {
        // a line's inputs carry the factor 2 of every outer index already
        // swept (lam), so its weight constant is scaled alike
        constexpr float lam[3] = {2.f, 1.f, 2.f};
        {
            const float hm = vx2 - vx, hp = vx2 + vx, bm = tx - 1.f, bp = tx + 1.f;
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float k = (b == 1 || c == 1) ? 0.f : (b == 0 && c == 0) ? 1.f : (b == 2 && c == 2) ? c1o9 : c1o3;
                    bwdf2(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), k, hm, hp, bm, bp);
                }
        }
        {
            const float hm = vy2 - vy, hp = vy2 + vy, bm = ty - 1.f, bp = ty + 1.f;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float k = (c == 1) ? 0.f : wv[i] * (c == 0 ? 1.f : c1o3) * lam[i];
                    bwdf2(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), k, hm, hp, bm, bp);
                }
        }
        {
            const float hm = vz2 - vz, hp = vz2 + vz, bm = tz - 1.f, bp = tz + 1.f;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    bwdf2(VFX_M(i, j, 0), VFX_M(i, j, 1), VFX_M(i, j, 2), wv[i] * wv[j] * (lam[i] * lam[j]), hm, hp, bm,
                          bp);
        }
    } else {
        {
            const float hvm = 0.5f * (vx2 - vx), hvp = 0.5f * (vx2 + vx), vmh = vx - 0.5f, vph = vx + 0.5f;
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float k = (b == 1 || c == 1) ? 0.f : (b == 0 && c == 0) ? 1.f : (b == 2 && c == 2) ? c1o9 : c1o3;
                    bwdf(VFX_M(0, b, c), VFX_M(1, b, c), VFX_M(2, b, c), k, hvm, hvp, vmh, vph);
                }
        }
        {
            const float hvm = 0.5f * (vy2 - vy), hvp = 0.5f * (vy2 + vy), vmh = vy - 0.5f, vph = vy + 0.5f;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float k = (c == 1) ? 0.f : wv[i] * (c == 0 ? 1.f : c1o3);
                    bwdf(VFX_M(i, 0, c), VFX_M(i, 1, c), VFX_M(i, 2, c), k, hvm, hvp, vmh, vph);
                }
        }
        {
            const float hvm = 0.5f * (vz2 - vz), hvp = 0.5f * (vz2 + vz), vmh = vz - 0.5f, vph = vz + 0.5f;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    bwdf(VFX_M(i, j, 0), VFX_M(i, j, 1), VFX_M(i, j, 2), wv[i] * wv[j], hvm, hvp, vmh, vph);
        }
    }
    mv[0] = drho;
    mv[1] = vx;
    mv[2] = vy;
    mv[3] = vz;
}

#undef VFX_M

} // namespace vfx
