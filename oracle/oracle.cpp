// oracle.cpp -- CPU restatement of the VirtualFluids D3Q27 hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (virtualfluids_amd/, include/)
// may include, link or call this file; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, and only as the checker.
//
// What it restates (reference paths relative to the VirtualFluids source root):
//   * K17 cumulant collision      src/lbm/collision/K17CompressibleNavierStokes.h:87-591
//   * chimera transforms          src/lbm/ChimeraTransformation.h:58-121
//   * macroscopic sums            src/lbm/MacroscopicQuantities.h:55-133
//   * compressible BGK            src/gpu/core/Kernel/Compressible/NavierStokes/B92/
//                                   B92CompressibleNavierStokes_Device.cu:189-234
//   * shifted equilibrium init    src/gpu/core/PreProcessor/PreProcessorStrategy/
//                                   InitNavierStokesCompressible/InitNavierStokesCompressible_Device.cu:170-210
//   * step order                  src/cpu/core/Simulation/Simulation.cpp:249-364
//                                 (pre-BC -> collide -> stream -> post-BC -> swap)
//   * velocity bounce-back        src/gpu/core/BoundaryConditions/Velocity/VelocityBounceBack.cu:159-184
//                                 src/cpu/core/BoundaryConditions/VelocityBounceBack.cpp:59-79
//                                 (+ float-rounded wall term, BoundaryConditions.h:195-262)
//
// Storage in the oracle is deliberately NOT the esoteric twist: the lattice is a
// plain node-major array F[node][27] of pre-collision shifted populations
// (f_i - w_i) and a step is collide -> push-stream into a second array -> swap.
// Streaming is a pure copy, so this gives bit-identical populations to any
// in-place scheme (EsoSplit on the CPU, the GPU esoteric twist) while sharing no
// index arithmetic with them.
//
// Arithmetic: every expression keeps the reference's evaluation order, and the
// file is compiled with -ffp-contract=off, so the FP64 results are the
// reference's bit for bit.
//
// Parity pins: see tests/test_oracle_*.py -- K17 and the full step are checked
// bit-exact against the compiled reference CPU closure (oracle/_ref) and against
// committed golden vectors; the chimera/macroscopic helpers against the
// reference unit tests' known answers; BGK against an independent textbook
// formulation (unshifted populations) plus conservation/fixed-point invariants.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <array>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    ORC_BC_VELOCITY_BOUNCE_BACK = 0,
    ORC_BC_NOSLIP_INTERPOLATED = 1,
    ORC_BC_VELOCITY_INTERPOLATED = 2,
    ORC_BC_SLIP = 3,
    ORC_BC_PRESSURE_NEQ = 4,
    ORC_BC_VELOCITY_PRESSURE_INTERPOLATED = 5,
    ORC_BC_OUTFLOW_NON_REFLECTING = 6,
    ORC_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION = 7,
    ORC_BC_SLIP_TURBULENT_VISCOSITY = 8,
    ORC_BC_STRESS_INTERPOLATED = 9,
    ORC_BC_STRESS_BOUNCE_BACK = 10,
    ORC_BC_STRESS_BOUNCE_BACK_PRESSURE = 11,
    ORC_BC_PRECURSOR_VELOCITY = 12,
    ORC_BC_PRECURSOR_DISTRIBUTIONS = 13
};

namespace {

// ---------------------------------------------------------------- D3Q27 ----
// Direction numbering of src/lbm/constants/D3Q27.h:42-72.
enum : int {
    d000 = 0, dP00, dM00, d0P0, d0M0, d00P, d00M, dPP0, dMM0, dPM0, dMP0,
    dP0P, dM0M, dP0M, dM0P, d0PP, d0MM, d0PM, d0MP,
    dPPP, dMPP, dPMP, dMMP, dPPM, dMPM, dPMM, dMMM
};
constexpr int CX[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1};
constexpr int CY[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1,-1};
constexpr int CZ[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1, 1, 1, 1, 1,-1,-1,-1,-1};

int inverse_of(int d)
{
    for (int e = 0; e < 27; ++e)
        if (CX[e] == -CX[d] && CY[e] == -CY[d] && CZ[e] == -CZ[d]) return e;
    return -1;
}

// direction with velocity components (i-1, j-1, k-1), i,j,k in {0,1,2}
int dir_of(int i, int j, int k)
{
    for (int e = 0; e < 27; ++e)
        if (CX[e] == i - 1 && CY[e] == j - 1 && CZ[e] == k - 1) return e;
    return -1;
}

struct Tables {
    int inv[27];
    int cube[3][3][3];
    Tables()
    {
        for (int d = 0; d < 27; ++d) inv[d] = inverse_of(d);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) cube[i][j][k] = dir_of(i, j, k);
    }
};
const Tables T;

// rational constants as the reference spells them (NumericConstants.h)
constexpr double c1o2 = 1. / 2., c3o2 = 3. / 2., c1o3 = 1. / 3., c2o3 = 2. / 3., c1o6 = 1. / 6.;
constexpr double c1o9 = 1. / 9., c2o9 = 2. / 9., c4o9 = 4. / 9., c1o18 = 1. / 18., c1o27 = 1. / 27.;
constexpr double c1o36 = 1. / 36., c8o27 = 8. / 27., c2o27 = 2. / 27., c1o54 = 1. / 54., c1o216 = 1. / 216.;
constexpr double c9o4 = 9. / 4., c9o2 = 9. / 2.;

double weight(int d)
{
    int n = std::abs(CX[d]) + std::abs(CY[d]) + std::abs(CZ[d]);
    return n == 0 ? c8o27 : n == 1 ? c2o27 : n == 2 ? c1o54 : c1o216;
}

// -------------------------------------------------------- chimera helpers ----
// ChimeraTransformation.h:58-70
inline void fwd_k(double& a, double& b, double& c, double vv, double v2, double kinv, double k)
{
    const double s = a + c;
    const double d = c - a;
    double m0 = s + b;
    a = m0;
    m0 *= kinv;
    m0 += 1.0;
    b = (d * kinv - m0 * vv) * k;
    c = ((s - 2.0 * d * vv) * kinv + v2 * m0) * k;
}
// ChimeraTransformation.h:78-87
inline void bwd_k(double& a, double& b, double& c, double vv, double v2, double kinv, double k)
{
    const double r0 = (((c - b) * c1o2 + b * vv) * kinv + (a * kinv + 1.0) * (v2 - vv) * c1o2) * k;
    const double r1 = (((a - c) - 2.0 * b * vv) * kinv + (a * kinv + 1.0) * (-v2)) * k;
    c = (((c + b) * c1o2 + b * vv) * kinv + (a * kinv + 1.0) * (v2 + vv) * c1o2) * k;
    a = r0;
    b = r1;
}
// ChimeraTransformation.h:96-104
inline void fwd(double& a, double& b, double& c, double vv, double v2)
{
    const double s = (a + c) + b;
    const double d = c - a;
    c = (c + a) + (v2 * s - 2.0 * vv * d);
    b = d - vv * s;
    a = s;
}
// ChimeraTransformation.h:113-121
inline void bwd(double& a, double& b, double& c, double vv, double v2)
{
    const double ra = (c + a * (v2 - vv)) * c1o2 + b * (vv - c1o2);
    const double rb = ((a - c) - a * v2) - 2.0 * b * vv;
    c = (c + a * (v2 + vv)) * c1o2 + b * (vv + c1o2);
    b = rb;
    a = ra;
}

// ------------------------------------------------------ macroscopic sums ----
// MacroscopicQuantities.h:55-64 (pyramid summation)
double density(const double* f)
{
    const double corners = ((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]));
    const double edges = (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P]))) +
                         ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]));
    const double faces = ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0])) + (f[d00M] + f[d00P]);
    return ((corners + edges) + f[d000]) + faces;
}
// MacroscopicQuantities.h:69-85
double mom_x(const double* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
           (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
           (f[dP00] - f[dM00]);
}
double mom_y(const double* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
           (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
           (f[d0P0] - f[d0M0]);
}
double mom_z(const double* f)
{
    return (((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
           (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
           (f[d00P] - f[d00M]);
}

// Central moments -> shifted populations: backward chimera sweeps x, y, z with
// the well-conditioning constants of K17CompressibleNavierStokes.h:525-557
// (the same sweeps close InterpolationCF.h:255-293 and InterpolationFC.h:210-248).
// m[i][j][k] holds the moment of orders (i, j, k); on return the population of
// direction (i-1, j-1, k-1).
void backward_sweeps(double (&m)[3][3][3], double vx, double vy, double vz)
{
    const double vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    // x sweep (:525-533): rows (cy,cz)
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
            if (b == 1 || c == 1) { bwd(m[0][b][c], m[1][b][c], m[2][b][c], vx, vx2); continue; }
            double kinv, k;
            if (b == 0 && c == 0) { kinv = 1.0; k = 1.0; }
            else if (b == 2 && c == 2) { kinv = 9.0; k = c1o9; }
            else { kinv = 3.0; k = c1o3; }
            bwd_k(m[0][b][c], m[1][b][c], m[2][b][c], vx, vx2, kinv, k);
        }
    // y sweep (:537-545): rows (i, cz)
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            if (c == 1) { bwd(m[i][0][c], m[i][1][c], m[i][2][c], vy, vy2); continue; }
            double kinv, k;
            if (i == 1) { kinv = (c == 0) ? c3o2 : c9o2; k = (c == 0) ? c2o3 : c2o9; }
            else        { kinv = (c == 0) ? 6.0 : 18.0; k = (c == 0) ? c1o6 : c1o18; }
            bwd_k(m[i][0][c], m[i][1][c], m[i][2][c], vy, vy2, kinv, k);
        }
    // z sweep (:549-557): rows (i,j)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double kinv, k;
            if (i == 1 && j == 1) { kinv = c9o4; k = c4o9; }
            else if (i == 1 || j == 1) { kinv = 9.0; k = c1o9; }
            else { kinv = 36.0; k = c1o36; }
            bwd_k(m[i][j][0], m[i][j][1], m[i][j][2], vz, vz2, kinv, k);
        }
}

// ------------------------------------------------------------- K17 -------
// Weight of a velocity-space slot along one axis before that axis is
// transformed: M/P -> 1/6, 0 -> 2/3; after it is transformed the slot holds a
// moment of order 0,1,2 whose weight factor is 1, (none), 1/3.  The chimera
// constant K of a row is the product of the factors of the two other axes.
// Reproduces the K arguments of K17CompressibleNavierStokes.h:190-222 and
// :525-557, including the (K^-1 = 3, K = 1/9) pair the reference passes for
// the (y,z) = (2,2) row of the forward x sweep (:222).
// Eddy-viscosity models (vf::lbm::TurbulenceModel, src/lbm/collision/TurbulentViscosity.h:50-89)
enum { TURB_NONE = 0, TURB_SMAGORINSKY = 1, TURB_AMD = 2, TURB_QR = 3 };

// :355-368 with TurbulentViscosity.h:63-84
double smagorinsky_viscosity(double Cs, double dxux, double dyuy, double dzuz, double Dxy, double Dxz, double Dyz)
{
    return Cs * Cs * std::sqrt(2.0 * (dxux * dxux + dyuy * dyuy + dzuz * dzuz) + Dxy * Dxy + Dxz * Dxz + Dyz * Dyz);
}
double qr_viscosity(double C, double dxux, double dyuy, double dzuz, double Dxy, double Dxz, double Dyz)
{
    const double Q = c1o2 * (dxux * dxux + dyuy * dyuy + dzuz * dzuz) + 0.25 * (Dxy * Dxy + Dxz * Dxz + Dyz * Dyz);
    const double R = -dxux * dyuy * dzuz + 0.25 * (-Dxy * Dxz * Dyz + dxux * Dyz * Dyz + dyuy * Dxz * Dxz + dzuz * Dxy * Dxy);
    return C * (R > 0.0 ? R : 0.0) / Q;
}

// turb != TURB_NONE: the node relaxes with omega / (1 + 3 omega nu_t) (:245);
// Smagorinsky / QR replace *nut by the model's value of the new gradients.
void k17_collide(double* f, double omega_in, const double* lim, const double* force, double* macro,
                 int turb = TURB_NONE, double sgs = 0.0, double* nut = nullptr)
{
    const double omega = turb == TURB_NONE ? omega_in : omega_in / (1.0 + 3.0 * omega_in * *nut);
    double m[3][3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) m[i][j][k] = f[T.cube[i][j][k]];

    // :157 macroscopic values
    const double drho = density(f);
    double vx = mom_x(f), vy = mom_y(f), vz = mom_z(f);
    const double oneOverRho = 1.0 / (drho + 1.0);
    vx *= oneOverRho;
    vy *= oneOverRho;
    vz *= oneOverRho;
    // :164-166 half force
    vx += force[0];
    vy += force[1];
    vz += force[2];
    const double vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const double limP = lim[0], limM = lim[1], limD = lim[2];

    // forward z sweep (:190-198): rows (i,j), K = w(i)*w(j) with w(M)=w(P)=1/6, w(0)=2/3
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double kinv, k;
            if (i == 1 && j == 1) { kinv = c9o4; k = c4o9; }
            else if (i == 1 || j == 1) { kinv = 9.0; k = c1o9; }
            else { kinv = 36.0; k = c1o36; }
            fwd_k(m[i][j][0], m[i][j][1], m[i][j][2], vz, vz2, kinv, k);
        }
    // forward y sweep (:202-210): rows (i, cz); cz==1 has no K
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) {
            if (c == 1) { fwd(m[i][0][c], m[i][1][c], m[i][2][c], vy, vy2); continue; }
            double kinv, k;
            if (i == 1) { kinv = (c == 0) ? c3o2 : c9o2; k = (c == 0) ? c2o3 : c2o9; }
            else        { kinv = (c == 0) ? 6.0 : 18.0; k = (c == 0) ? c1o6 : c1o18; }
            fwd_k(m[i][0][c], m[i][1][c], m[i][2][c], vy, vy2, kinv, k);
        }
    // forward x sweep (:214-222): rows (cy, cz)
    for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
            if (b == 1 || c == 1) { fwd(m[0][b][c], m[1][b][c], m[2][b][c], vx, vx2); continue; }
            double kinv, k;
            if (b == 0 && c == 0) { kinv = 1.0; k = 1.0; }
            else if (b == 2 && c == 2) { kinv = 3.0; k = c1o9; }   // as passed at :222
            else { kinv = 3.0; k = c1o3; }
            fwd_k(m[0][b][c], m[1][b][c], m[2][b][c], vx, vx2, kinv, k);
        }

    // aliases: m[a][b][c] is the central moment of order (a,b,c)
    double& m000 = m[0][0][0]; double& m100 = m[1][0][0]; double& m010 = m[0][1][0]; double& m001 = m[0][0][1];
    double& m200 = m[2][0][0]; double& m020 = m[0][2][0]; double& m002 = m[0][0][2];
    double& m110 = m[1][1][0]; double& m101 = m[1][0][1]; double& m011 = m[0][1][1];
    double& m111 = m[1][1][1];
    double& m210 = m[2][1][0]; double& m012 = m[0][1][2]; double& m201 = m[2][0][1];
    double& m021 = m[0][2][1]; double& m120 = m[1][2][0]; double& m102 = m[1][0][2];
    double& m211 = m[2][1][1]; double& m121 = m[1][2][1]; double& m112 = m[1][1][2];
    double& m220 = m[2][2][0]; double& m202 = m[2][0][2]; double& m022 = m[0][2][2];
    double& m122 = m[1][2][2]; double& m212 = m[2][1][2]; double& m221 = m[2][2][1];
    double& m222 = m[2][2][2];

    // :245-273 relaxation rates
    const double OxxPyyPzz = 1.0;
    const double OxyyPxzz = 8.0 * (-2.0 + omega) * (1.0 + 2.0 * omega) / (-8.0 - 14.0 * omega + 7.0 * omega * omega);
    const double OxyyMxzz = 8.0 * (-2.0 + omega) * (-7.0 + 4.0 * omega) / (56.0 - 50.0 * omega + 9.0 * omega * omega);
    const double Oxyz = 24.0 * (-2.0 + omega) * (-2.0 - 7.0 * omega + 3.0 * omega * omega) /
                        (48.0 + 152.0 * omega - 130.0 * omega * omega + 29.0 * omega * omega * omega);
    const double O4 = 1.0, O5 = 1.0, O6 = 1.0;
    const double factorA = (4.0 + 2.0 * omega - 3.0 * omega * omega) / (2.0 - 7.0 * omega + 5.0 * omega * omega);
    const double factorB = (4.0 + 28.0 * omega - 14.0 * omega * omega) / (6.0 - 21.0 * omega + 15.0 * omega * omega);

    // :282-316 cumulants of order 4..6
    double cm211 = m211 - ((m200 + c1o3) * m011 + 2.0 * m110 * m101) * oneOverRho;
    double cm121 = m121 - ((m020 + c1o3) * m101 + 2.0 * m110 * m011) * oneOverRho;
    double cm112 = m112 - ((m002 + c1o3) * m110 + 2.0 * m101 * m011) * oneOverRho;
    double cm220 = m220 - (((m200 * m020 + 2.0 * m110 * m110) + c1o3 * (m200 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    double cm202 = m202 - (((m200 * m002 + 2.0 * m101 * m101) + c1o3 * (m200 + m002)) * oneOverRho - c1o9 * (drho * oneOverRho));
    double cm022 = m022 - (((m002 * m020 + 2.0 * m011 * m011) + c1o3 * (m002 + m020)) * oneOverRho - c1o9 * (drho * oneOverRho));
    double cm122 = m122 - ((m002 * m120 + m020 * m102 + 4.0 * m011 * m111 + 2.0 * (m101 * m021 + m110 * m012)) +
                           c1o3 * (m120 + m102)) * oneOverRho;
    double cm212 = m212 - ((m002 * m210 + m200 * m012 + 4.0 * m101 * m111 + 2.0 * (m011 * m201 + m110 * m102)) +
                           c1o3 * (m210 + m012)) * oneOverRho;
    double cm221 = m221 - ((m200 * m021 + m020 * m201 + 4.0 * m110 * m111 + 2.0 * (m101 * m120 + m011 * m210)) +
                           c1o3 * (m021 + m201)) * oneOverRho;
    double cm222 = m222 + ((-4.0 * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                            4.0 * (m011 * m211 + m101 * m121 + m110 * m112) -
                            2.0 * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                           (4.0 * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                            2.0 * (m200 * m020 * m002) + 16.0 * m110 * m101 * m011) * oneOverRho * oneOverRho -
                           c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                           (2.0 * (m101 * m101 + m011 * m011 + m110 * m110) +
                            (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                               oneOverRho * oneOverRho * c2o3 +
                           c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    // :323-335 linear combinations
    double mxxPyyPzz = m200 + m020 + m002;
    double mxxMyy = m200 - m020;
    double mxxMzz = m200 - m002;
    double mxxyPyzz = m210 + m012, mxxyMyzz = m210 - m012;
    double mxxzPyyz = m201 + m021, mxxzMyyz = m201 - m021;
    double mxyyPxzz = m120 + m102, mxyyMxzz = m120 - m102;

    // :347-352 velocity gradients (times rho)
    const double Dxy = -3.0 * omega * m110;
    const double Dxz = -3.0 * omega * m101;
    const double Dyz = -3.0 * omega * m011;
    const double dxux = c1o2 * (-omega) * (mxxMyy + mxxMzz) + c1o2 * OxxPyyPzz * (m000 - mxxPyyPzz);
    const double dyuy = dxux + omega * c3o2 * mxxMyy;
    const double dzuz = dxux + omega * c3o2 * mxxMzz;
    if (turb == TURB_SMAGORINSKY) *nut = smagorinsky_viscosity(sgs, dxux, dyuy, dzuz, Dxy, Dxz, Dyz);
    if (turb == TURB_QR) *nut = qr_viscosity(sgs, dxux, dyuy, dzuz, Dxy, Dxz, Dyz);

    // :375-387 second order with correction
    mxxPyyPzz += OxxPyyPzz * (m000 - mxxPyyPzz) - 3.0 * (1.0 - c1o2 * OxxPyyPzz) * (vx2 * dxux + vy2 * dyuy + vz2 * dzuz);
    mxxMyy += omega * (-mxxMyy) - 3.0 * (1.0 + c1o2 * (-omega)) * (vx2 * dxux - vy2 * dyuy);
    mxxMzz += omega * (-mxxMzz) - 3.0 * (1.0 + c1o2 * (-omega)) * (vx2 * dxux - vz2 * dzuz);
    m011 += omega * (-m011);
    m101 += omega * (-m101);
    m110 += omega * (-m110);

    // :397-410 limited third order
    auto limited = [](double rate, double x, double l) { return rate + (1.0 - rate) * std::fabs(x) / (std::fabs(x) + l); };
    double w = limited(Oxyz, m111, limD);
    m111 += w * (-m111);
    w = limited(OxyyPxzz, mxxyPyzz, limP); mxxyPyzz += w * (-mxxyPyzz);
    w = limited(OxyyMxzz, mxxyMyzz, limM); mxxyMyzz += w * (-mxxyMyzz);
    w = limited(OxyyPxzz, mxxzPyyz, limP); mxxzPyyz += w * (-mxxzPyyz);
    w = limited(OxyyMxzz, mxxzMyyz, limM); mxxzMyyz += w * (-mxxzMyyz);
    w = limited(OxyyPxzz, mxyyPxzz, limP); mxyyPxzz += w * (-mxyyPxzz);
    w = limited(OxyyMxzz, mxyyMxzz, limM); mxyyMxzz += w * (-mxyyMxzz);

    // :424-433 back from linear combinations
    m200 = c1o3 * (mxxMyy + mxxMzz + mxxPyyPzz);
    m020 = c1o3 * (-2.0 * mxxMyy + mxxMzz + mxxPyyPzz);
    m002 = c1o3 * (mxxMyy - 2.0 * mxxMzz + mxxPyyPzz);
    m210 = (mxxyMyzz + mxxyPyzz) * c1o2;
    m012 = (-mxxyMyzz + mxxyPyzz) * c1o2;
    m201 = (mxxzMyyz + mxxzPyyz) * c1o2;
    m021 = (-mxxzMyyz + mxxzPyyz) * c1o2;
    m120 = (mxyyMxzz + mxyyPxzz) * c1o2;
    m102 = (-mxyyMxzz + mxyyPxzz) * c1o2;

    // :443-459 fourth order to modified equilibrium, 5th/6th to zero
    cm022 = -O4 * (1.0 / omega - c1o2) * (dyuy + dzuz) * c2o3 * factorA + (1.0 - O4) * (cm022);
    cm202 = -O4 * (1.0 / omega - c1o2) * (dxux + dzuz) * c2o3 * factorA + (1.0 - O4) * (cm202);
    cm220 = -O4 * (1.0 / omega - c1o2) * (dyuy + dxux) * c2o3 * factorA + (1.0 - O4) * (cm220);
    cm112 = -O4 * (1.0 / omega - c1o2) * Dxy * c1o3 * factorB + (1.0 - O4) * (cm112);
    cm121 = -O4 * (1.0 / omega - c1o2) * Dxz * c1o3 * factorB + (1.0 - O4) * (cm121);
    cm211 = -O4 * (1.0 / omega - c1o2) * Dyz * c1o3 * factorB + (1.0 - O4) * (cm211);
    cm122 += O5 * (-cm122);
    cm212 += O5 * (-cm212);
    cm221 += O5 * (-cm221);
    cm222 += O6 * (-cm222);

    // :469-505 central moments from post-collision cumulants
    m211 = cm211 + c1o3 * ((3.0 * m200 + 1.0) * m011 + 6.0 * m110 * m101) * oneOverRho;
    m121 = cm121 + c1o3 * ((3.0 * m020 + 1.0) * m101 + 6.0 * m110 * m011) * oneOverRho;
    m112 = cm112 + c1o3 * ((3.0 * m002 + 1.0) * m110 + 6.0 * m101 * m011) * oneOverRho;
    m220 = cm220 + (((m200 * m020 + 2.0 * m110 * m110) * 9.0 + 3.0 * (m200 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m202 = cm202 + (((m200 * m002 + 2.0 * m101 * m101) * 9.0 + 3.0 * (m200 + m002)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m022 = cm022 + (((m002 * m020 + 2.0 * m011 * m011) * 9.0 + 3.0 * (m002 + m020)) * oneOverRho - (drho * oneOverRho)) * c1o9;
    m122 = cm122 + c1o3 * (3.0 * (m002 * m120 + m020 * m102 + 4.0 * m011 * m111 + 2.0 * (m101 * m021 + m110 * m012)) +
                           (m120 + m102)) * oneOverRho;
    m212 = cm212 + c1o3 * (3.0 * (m002 * m210 + m200 * m012 + 4.0 * m101 * m111 + 2.0 * (m011 * m201 + m110 * m102)) +
                           (m210 + m012)) * oneOverRho;
    m221 = cm221 + c1o3 * (3.0 * (m200 * m021 + m020 * m201 + 4.0 * m110 * m111 + 2.0 * (m101 * m120 + m011 * m210)) +
                           (m021 + m201)) * oneOverRho;
    m222 = cm222 - ((-4.0 * m111 * m111 - (m200 * m022 + m020 * m202 + m002 * m220) -
                     4.0 * (m011 * m211 + m101 * m121 + m110 * m112) -
                     2.0 * (m120 * m102 + m210 * m012 + m201 * m021)) * oneOverRho +
                    (4.0 * (m101 * m101 * m020 + m011 * m011 * m200 + m110 * m110 * m002) +
                     2.0 * (m200 * m020 * m002) + 16.0 * m110 * m101 * m011) * oneOverRho * oneOverRho -
                    c1o3 * (m022 + m202 + m220) * oneOverRho - c1o9 * (m200 + m020 + m002) * oneOverRho +
                    (2.0 * (m101 * m101 + m011 * m011 + m110 * m110) +
                     (m002 * m020 + m002 * m200 + m020 * m200) + c1o3 * (m002 + m020 + m200)) *
                        oneOverRho * oneOverRho * c2o3 +
                    c1o27 * ((drho * drho - drho) * oneOverRho * oneOverRho));

    // :512-514
    m100 = -m100;
    m010 = -m010;
    m001 = -m001;

    backward_sweeps(m, vx, vy, vz);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) f[T.cube[i][j][k]] = m[i][j][k];

    if (macro) { macro[0] = drho; macro[1] = vx; macro[2] = vy; macro[3] = vz; }
}

// ----------------------------------------------------------- BGK (B92) ----
// B92CompressibleNavierStokes_Device.cu:189-234; the naming there is
// E=dP00 W=dM00 N=d0P0 S=d0M0 T=d00P B=d00M and compounds thereof.
void bgk_collide(double* f, double omega, double* macro)
{
    const double drho = (f[dPPP] + f[dMMM]) + (f[dMMP] + f[dPPM]) + (f[dPMP] + f[dMPM]) + (f[dMPP] + f[dPMM]) +
                        (f[dPP0] + f[dMM0]) + (f[dMP0] + f[dPM0]) + (f[dP0P] + f[dM0M]) + (f[dP0M] + f[dM0P]) +
                        (f[d0PP] + f[d0MM]) + (f[d0PM] + f[d0MP]) + (f[dP00] + f[dM00]) + (f[d0P0] + f[d0M0]) +
                        (f[d00P] + f[d00M]) + f[d000];
    const double rho = drho + 1.0;
    const double oor = 1.0 / rho;
    const double vx1 = oor * ((f[dPPP] - f[dMMM]) + (f[dPPM] - f[dMMP]) + (f[dPMP] - f[dMPM]) + (f[dPMM] - f[dMPP]) +
                              (f[dPP0] - f[dMM0]) + (f[dPM0] - f[dMP0]) + (f[dP0P] - f[dM0M]) + (f[dP0M] - f[dM0P]) +
                              (f[dP00] - f[dM00]));
    const double vx2 = oor * ((f[dPPP] - f[dMMM]) + (f[dPPM] - f[dMMP]) + (f[dMPM] - f[dPMP]) + (f[dMPP] - f[dPMM]) +
                              (f[dPP0] - f[dMM0]) + (f[dMP0] - f[dPM0]) + (f[d0PP] - f[d0MM]) + (f[d0PM] - f[d0MP]) +
                              (f[d0P0] - f[d0M0]));
    const double vx3 = oor * ((f[dPPP] - f[dMMM]) + (f[dMMP] - f[dPPM]) + (f[dPMP] - f[dMPM]) + (f[dMPP] - f[dPMM]) +
                              (f[dP0P] - f[dM0M]) + (f[dM0P] - f[dP0M]) + (f[d0PP] - f[d0MM]) + (f[d0MP] - f[d0PM]) +
                              (f[d00P] - f[d00M]));
    const double cusq = c3o2 * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
    const double a = 1.0 + (-omega);
    for (int d = 0; d < 27; ++d) {
        const double wt = weight(d);
        if (d == d000) { f[d] = f[d] * a - (-omega) * wt * (drho - rho * cusq); continue; }
        // projected velocity spelled as the reference writes it: components in
        // x,y,z order with the direction's signs
        double cu;
        const int c[3] = {CX[d], CY[d], CZ[d]};
        const double v[3] = {vx1, vx2, vx3};
        bool first = true;
        cu = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] == 0) continue;
            if (first) { cu = c[ax] > 0 ? v[ax] : -v[ax]; first = false; }
            else cu = c[ax] > 0 ? cu + v[ax] : cu - v[ax];
        }
        f[d] = f[d] * a - (-omega) * wt * (drho + rho * (3.0 * cu + c9o2 * cu * cu - cusq));
    }
    if (macro) { macro[0] = drho; macro[1] = vx1; macro[2] = vx2; macro[3] = vx3; }
}

// ------------------------------------------------------ shifted equilibrium ----
// InitNavierStokesCompressible_Device.cu:170-210
void equilibrium(double* f, double drho, double vx1, double vx2, double vx3)
{
    const double cu_sq = c3o2 * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
    const double v[3] = {vx1, vx2, vx3};
    for (int d = 0; d < 27; ++d) {
        if (d == d000) { f[d] = c8o27 * (drho - cu_sq * (1.0 + drho)); continue; }
        const int c[3] = {CX[d], CY[d], CZ[d]};
        double cu = 0.0;
        bool first = true;
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] == 0) continue;
            if (first) { cu = c[ax] > 0 ? v[ax] : -v[ax]; first = false; }
            else cu = c[ax] > 0 ? cu + v[ax] : cu - v[ax];
        }
        f[d] = weight(d) * (drho + (1.0 + drho) * (3.0 * cu + c9o2 * cu * cu - cu_sq));
    }
}

// Wall term of the moving-wall bounce-back for link direction d (towards the
// wall).  GPU (VelocityBounceBack.cu:159-184): 6 w_d (c_d . u) in working
// precision.  CPU (BoundaryConditions.h:195-262): the wall velocity is stored
// as float, the signed component sum is formed in float, scaled by the double
// constant and rounded to float again.
double wall_term(int d, double ux, double uy, double uz, bool cpu_float)
{
    const int n = std::abs(CX[d]) + std::abs(CY[d]) + std::abs(CZ[d]);
    const double coef = n == 1 ? c4o9 : n == 2 ? c1o9 : c1o36;
    const int c[3] = {CX[d], CY[d], CZ[d]};
    if (cpu_float) {
        const float u[3] = {(float)ux, (float)uy, (float)uz};
        float s = 0.0f;
        bool first = true;
        for (int ax = 0; ax < 3; ++ax) {
            if (c[ax] == 0) continue;
            if (first) { s = c[ax] > 0 ? +u[ax] : -u[ax]; first = false; }
            else s = c[ax] > 0 ? s + u[ax] : s - u[ax];
        }
        return (double)(float)(coef * (double)s);
    }
    const double u[3] = {ux, uy, uz};
    double s = 0.0;
    bool first = true;
    for (int ax = 0; ax < 3; ++ax) {
        if (c[ax] == 0) continue;
        if (first) { s = c[ax] > 0 ? u[ax] : -u[ax]; first = false; }
        else s = c[ax] > 0 ? s + u[ax] : s - u[ax];
    }
    return coef * s;
}

// c_d . u spelled as the reference writes it (first non-zero component with
// its sign, the others added/subtracted in x, y, z order)
double proj(int d, double ux, double uy, double uz)
{
    const int c[3] = {CX[d], CY[d], CZ[d]};
    const double u[3] = {ux, uy, uz};
    double v = 0.0;
    bool first = true;
    for (int a = 0; a < 3; ++a) {
        if (c[a] == 0) continue;
        if (first) { v = c[a] > 0 ? u[a] : -u[a]; first = false; }
        else v = c[a] > 0 ? v + u[a] : v - u[a];
    }
    return v;
}

// Interpolated bounce-back family on the post-collision populations p of one
// boundary node (NoSlipInterpolatedCompressible.cu:44-378,
// VelocityInterpolatedCompressible.cu, SlipCompressible.cu; formulas
// KernelUtilities.h:171-197).  out[d'] receives the population that re-enters
// the node along d' = inv(d) for every wall link d; returns a mask of written d'.
void interp_bc(int kind, const double* p, const double* q, int n, int i, double omega, double VX, double VY, double VZ,
               double* out, bool* written)
{
    const double drho = p[dPMP] + p[dMPP] + p[dPPP] + p[dMMP] + p[dPMM] + p[dMPM] + p[dPPM] + p[dMMM] + p[d0PM] +
                        p[d0MP] + p[d0PP] + p[d0MM] + p[dP0M] + p[dM0P] + p[dP0P] + p[dM0M] + p[dPM0] + p[dMP0] +
                        p[dPP0] + p[dMM0] + p[d00P] + p[d00M] + p[d0P0] + p[d0M0] + p[dP00] + p[dM00] + p[d000];
    const double vx1 = (((p[dPMP] - p[dMPM]) - (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                        ((p[dP0M] - p[dM0P]) + (p[dP0P] - p[dM0M])) + ((p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                        (p[dP00] - p[dM00])) / (1.0 + drho);
    const double vx2 = ((-(p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                        ((p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + (-(p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                        (p[d0P0] - p[d0M0])) / (1.0 + drho);
    const double vx3 = (((p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) + (p[dMMP] - p[dPPM])) +
                        (-(p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + ((p[dP0P] - p[dM0M]) - (p[dP0M] - p[dM0P])) +
                        (p[d00P] - p[d00M])) / (1.0 + drho);
    const double cu_sq = c3o2 * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3) * (1.0 + drho);
    bool fx = false, fy = false, fz = false;
    for (int d = 1; d < 27; ++d) {
        const double qd = (double)q[(int64_t)d * n + i];
        if (!(qd >= 0.0 && qd <= 1.0)) continue;
        const double w = weight(d);
        const double vlb = proj(d, vx1, vx2, vx3);
        const double feq = w * (drho + c9o2 * vlb * vlb * (1.0 + drho) - cu_sq);
        double f = p[d], fi = p[T.inv[d]];
        if (kind == ORC_BC_PRECURSOR_VELOCITY) {
            // PrecursorNonReflectiveCompressible.cu:383, :392: f_0PP stands in
            // for f_0MP on the d0PM and d0MP links
            if (d == d0PM) fi = p[d0PP];
            if (d == d0MP) f = p[d0PP];
        }
        const double a = (1.0 - qd) / (1.0 + qd) * (f - fi + (f + fi - 2.0 * feq * omega) / (1.0 - omega)) * c1o2;
        double v;
        if (kind == ORC_BC_NOSLIP_INTERPOLATED) {
            v = a + (qd * (f + fi)) / (1.0 + qd);
        } else {
            double vbc;
            if (kind == ORC_BC_SLIP || kind == ORC_BC_SLIP_TURBULENT_VISCOSITY) {
                if (std::abs(CX[d]) + std::abs(CY[d]) + std::abs(CZ[d]) == 1) {
                    if (CX[d]) fx = true;
                    if (CY[d]) fy = true;
                    if (CZ[d]) fz = true;
                }
                vbc = proj(d, fx ? 0.0 : vx1, fy ? 0.0 : vx2, fz ? 0.0 : vx3);
            } else {
                vbc = proj(d, VX, VY, VZ);
            }
            v = a + (qd * (f + fi) - 6.0 * w * vbc) / (1.0 + qd);
            // getInterpolatedDistributionForVeloWithPressureBC (KernelUtilities.h:207-213)
            if (kind == ORC_BC_VELOCITY_PRESSURE_INTERPOLATED || kind == ORC_BC_PRECURSOR_VELOCITY) v = v - w * drho;
        }
        out[T.inv[d]] = v;
        written[T.inv[d]] = true;
    }
}

// Non-reflecting outflow (OutflowNonReflecting.cu:44-280,
// OutflowNonReflectingPressureCorrection.cu:46-245): from the pre-collision
// populations f of the boundary node and fn of its inward neighbour, the
// population entering through the boundary in every direction d with
// c_d . normal < 0 (normal one of the six face directions) becomes
//   fn[d] cs + (1 - cs) f[d]  [- w_d * factor * drho(f)],  cs = 1 / sqrtf(3).
// The pressure-corrected kernel reads its "neighbour" populations of the
// directions with a negative component from the boundary node's own slots
// (its neighbour indices are built from the boundary node), so for those d
// fn[d] is f[d]; restated as is.  Returns the values in out[d], written[d].
void outflow_bc(int kind, const double* f, const double* fn_true, int normal, double factor, double* out,
                bool* written)
{
    if (std::abs(CX[normal]) + std::abs(CY[normal]) + std::abs(CZ[normal]) != 1) return;
    const double cs = 1.0 / (double)std::sqrt(3.0f);
    const bool corr = kind == ORC_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION;
    const double rc = corr ? factor * density(f) : 0.0;
    for (int d = 1; d < 27; ++d) {
        if (CX[d] * CX[normal] + CY[d] * CY[normal] + CZ[d] * CZ[normal] >= 0) continue;
        const bool own = corr && (CX[d] < 0 || CY[d] < 0 || CZ[d] < 0);
        const double fn = own ? f[d] : fn_true[d];
        double v = fn * cs + (1.0 - cs) * f[d];
        if (corr) v = v - weight(d) * rc;
        out[d] = v;
        written[d] = true;
    }
}

double feq_shift(int d, double drho, double vx1, double vx2, double vx3, double cusq)
{
    if (d == d000) return c8o27 * (drho - (drho + 1.0) * cusq);
    const double cu = proj(d, vx1, vx2, vx3);
    return weight(d) * (drho + (drho + 1.0) * (3.0 * cu + c9o2 * cu * cu - cusq));
}

// PressureNonEquilibriumCompressible.cu: neighbour populations with their
// equilibrium replaced by the one of the prescribed density; returned for
// every direction with c_d . normal <= 0.
void pressure_bc(const double* fnb, double rho_bc, int normal, double* out, bool* written)
{
    double f[27];
    std::memcpy(f, fnb, sizeof(f));
    double drho1 = density(f);
    const double oor = 1.0 / (drho1 + 1.0);
    const double vx1 = mom_x(f) * oor, vx2 = mom_y(f) * oor, vx3 = mom_z(f) * oor;
    const double cusq = c3o2 * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3);
    for (int d = 0; d < 27; ++d) f[d] -= feq_shift(d, drho1, vx1, vx2, vx3, cusq);
    drho1 = rho_bc;
    for (int d = 0; d < 27; ++d) f[d] += feq_shift(d, drho1, vx1, vx2, vx3, cusq);
    for (int d = 0; d < 27; ++d) {
        if (CX[d] * CX[normal] + CY[d] * CY[normal] + CZ[d] * CZ[normal] > 0) continue;
        out[d] = f[d];
        written[d] = true;
    }
}


// ------------------------------------------------------------ refinement --
// Interface interpolation between a coarse level and its 2:1 fine level
// (src/lbm/interpolation/InterpolationCoefficients.h:74-432,
// src/lbm/refinement/InterpolationCF.h:57-306, InterpolationFC.h:56-263).
//
// The 8 source nodes of an interpolation cell sit at the corners
// (sx/2, sy/2, sz/2), s = +-1, of a unit cube centred on the cell; corner c
// has bit 0 = +x, bit 1 = +y, bit 2 = +z (c = 0 is the reference's MMM, 7 is
// PPP).  From each corner the reference takes drho, u and the five
// non-equilibrium second moments k.. = -3 omega (sum c_a c_b f / rho - u_a u_b)
// (kxx-yy, kxx-zz with -3/2 omega), :89-128.
//
// Each velocity component is then a polynomial
//   v(x,y,z) = v000 + v100 x + v010 y + v001 z + v200 x^2 + v020 y^2 + v002 z^2
//            + v110 xy + v101 xz + v011 yz + v111 xyz
// whose trilinear part interpolates the 8 corner values and whose curvature
// terms follow from the corner strain rates; drho is trilinear.  Written with
// the corner sums  S q = sum_c q_c  and the signed sums  D_x q = sum_c s_x q_c,
// D_xy q = sum_c s_x s_y q_c, ... the coefficients of :242-319 are:
//   v100 = D_x v / 4,  v110 = D_xy v / 2,  v111 = D_xyz v          (all v, drho)
//   a200 = (2 D_xy vy + 2 D_xz vz + D_x kxxMyy + D_x kxxMzz) / 16
//   a020 = (D_y kxy - 2 D_xy vy) / 8,   a002 = (D_z kxz - 2 D_xz vz) / 8
//   b200 = (D_x kxy - 2 D_xy vx) / 8,   b002 = (D_z kyz - 2 D_yz vz) / 8
//   b020 = (2 D_xy vx + 2 D_yz vz - 2 D_y kxxMyy + D_y kxxMzz) / 16
//   c200 = (D_x kxz - 2 D_xz vx) / 8,   c020 = (D_y kyz - 2 D_yz vy) / 8
//   c002 = (2 D_xz vx + 2 D_yz vy - 2 D_z kxxMzz + D_z kxxMyy) / 16
//   v000 = S v / 8 - (v200 + v020 + v002) / 4,   d000 = S drho / 8
// (a, b, c: the x, y, z velocity).  These are algebraically the reference's
// expressions; the association of the sums differs, so results agree to
// rounding (checked against the compiled reference header).
struct SourceMoments {
    double drho, vx, vy, vz, kxy, kyz, kxz, kxxMyy, kxxMzz;
};

void source_moments(const double* f, double omega, SourceMoments& s)
{
    s.drho = density(f);
    const double oneOverRho = 1.0 / (s.drho + 1.0);
    s.vx = mom_x(f) * oneOverRho;
    s.vy = mom_y(f) * oneOverRho;
    s.vz = mom_z(f) * oneOverRho;
    double pxy = 0, pyz = 0, pxz = 0, pxxMyy = 0, pxxMzz = 0;
    for (int d = 1; d < 27; ++d) {
        pxy += CX[d] * CY[d] * f[d];
        pyz += CY[d] * CZ[d] * f[d];
        pxz += CX[d] * CZ[d] * f[d];
        pxxMyy += (CX[d] * CX[d] - CY[d] * CY[d]) * f[d];
        pxxMzz += (CX[d] * CX[d] - CZ[d] * CZ[d]) * f[d];
    }
    s.kxy = -3.0 * omega * (pxy * oneOverRho - s.vx * s.vy);
    s.kyz = -3.0 * omega * (pyz * oneOverRho - s.vy * s.vz);
    s.kxz = -3.0 * omega * (pxz * oneOverRho - s.vx * s.vz);
    s.kxxMyy = -1.5 * omega * (pxxMyy * oneOverRho - (s.vx * s.vx - s.vy * s.vy));
    s.kxxMzz = -1.5 * omega * (pxxMzz * oneOverRho - (s.vx * s.vx - s.vz * s.vz));
}

// coefficient slots of one polynomial
enum { P000, P100, P010, P001, P200, P020, P002, P110, P101, P011, P111, NP };

struct Coefficients {
    double a[NP], b[NP], c[NP], d[NP]; // vx, vy, vz, drho (d has no square terms)
    double laplace_rho;
};

struct CornerSums {
    double S, Dx, Dy, Dz, Dxy, Dxz, Dyz, Dxyz;
};

CornerSums corner_sums(const double q[8])
{
    CornerSums r{0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 8; ++c) {
        const double sx = (c & 1) ? 1.0 : -1.0, sy = (c & 2) ? 1.0 : -1.0, sz = (c & 4) ? 1.0 : -1.0;
        r.S += q[c];
        r.Dx += sx * q[c];
        r.Dy += sy * q[c];
        r.Dz += sz * q[c];
        r.Dxy += sx * sy * q[c];
        r.Dxz += sx * sz * q[c];
        r.Dyz += sy * sz * q[c];
        r.Dxyz += sx * sy * sz * q[c];
    }
    return r;
}

void linear_part(const CornerSums& s, double* p)
{
    p[P100] = s.Dx / 4;
    p[P010] = s.Dy / 4;
    p[P001] = s.Dz / 4;
    p[P110] = s.Dxy / 2;
    p[P101] = s.Dxz / 2;
    p[P011] = s.Dyz / 2;
    p[P111] = s.Dxyz;
}

// MomentsOnSourceNodeSet::calculateCoefficients (:174-428), including the
// Taylor shift to the offset point (xoff, yoff, zoff) of :389-424, which, as
// in the reference, leaves the second-order coefficients and the xyz terms
// unchanged and drops the xyz contributions to the shifted first-order terms.
void interpolation_coefficients(const SourceMoments src[8], double xoff, double yoff, double zoff, Coefficients& k)
{
    double q[8];
    auto sums = [&](double SourceMoments::*field) {
        for (int c = 0; c < 8; ++c) q[c] = src[c].*field;
        return corner_sums(q);
    };
    const CornerSums vx = sums(&SourceMoments::vx), vy = sums(&SourceMoments::vy), vz = sums(&SourceMoments::vz);
    const CornerSums rho = sums(&SourceMoments::drho);
    const CornerSums kxy = sums(&SourceMoments::kxy), kyz = sums(&SourceMoments::kyz), kxz = sums(&SourceMoments::kxz);
    const CornerSums kMyy = sums(&SourceMoments::kxxMyy), kMzz = sums(&SourceMoments::kxxMzz);

    double* a = k.a;
    double* b = k.b;
    double* c = k.c;
    double* d = k.d;
    linear_part(vx, a);
    linear_part(vy, b);
    linear_part(vz, c);
    linear_part(rho, d);
    a[P200] = (2 * vy.Dxy + 2 * vz.Dxz + kMyy.Dx + kMzz.Dx) / 16;
    a[P020] = (kxy.Dy - 2 * vy.Dxy) / 8;
    a[P002] = (kxz.Dz - 2 * vz.Dxz) / 8;
    b[P200] = (kxy.Dx - 2 * vx.Dxy) / 8;
    b[P020] = (2 * vx.Dxy + 2 * vz.Dyz - 2 * kMyy.Dy + kMzz.Dy) / 16;
    b[P002] = (kyz.Dz - 2 * vz.Dyz) / 8;
    c[P200] = (kxz.Dx - 2 * vx.Dxz) / 8;
    c[P020] = (kyz.Dy - 2 * vy.Dyz) / 8;
    c[P002] = (2 * vx.Dxz + 2 * vy.Dyz - 2 * kMzz.Dz + kMyy.Dz) / 16;
    a[P000] = vx.S / 8 - (a[P200] + a[P020] + a[P002]) / 4;
    b[P000] = vy.S / 8 - (b[P200] + b[P020] + b[P002]) / 4;
    c[P000] = vz.S / 8 - (c[P200] + c[P020] + c[P002]) / 4;
    d[P000] = rho.S / 8;
    d[P200] = d[P020] = d[P002] = 0.0;

    // :378-380: the pressure curvature is used only for an unshifted cell
    k.laplace_rho = (xoff != 0.0 || yoff != 0.0 || zoff != 0.0)
                        ? 0.0
                        : -3.0 * (a[P100] * a[P100] + b[P010] * b[P010] + c[P001] * c[P001]) -
                              6.0 * (b[P100] * a[P010] + c[P100] * a[P001] + c[P010] * b[P001]);

    for (double* p : {a, b, c, d}) {
        p[P000] += xoff * p[P100] + yoff * p[P010] + zoff * p[P001] + xoff * xoff * p[P200] + yoff * yoff * p[P020] +
                   zoff * zoff * p[P002] + xoff * yoff * p[P110] + xoff * zoff * p[P101] + yoff * zoff * p[P011];
        p[P100] += 2 * xoff * p[P200] + yoff * p[P110] + zoff * p[P101];
        p[P010] += 2 * yoff * p[P020] + xoff * p[P110] + zoff * p[P011];
        p[P001] += 2 * zoff * p[P002] + xoff * p[P101] + yoff * p[P011];
    }
}

double poly_value(const double* p, double x, double y, double z)
{
    return p[P000] + x * p[P100] + y * p[P010] + z * p[P001] + x * x * p[P200] + y * y * p[P020] + z * z * p[P002] +
           x * y * p[P110] + x * z * p[P101] + y * z * p[P011] + x * y * z * p[P111];
}
void poly_gradient(const double* p, double x, double y, double z, double g[3])
{
    g[0] = p[P100] + 2 * x * p[P200] + y * p[P110] + z * p[P101] + y * z * p[P111];
    g[1] = p[P010] + 2 * y * p[P020] + x * p[P110] + z * p[P011] + x * z * p[P111];
    g[2] = p[P001] + 2 * z * p[P002] + x * p[P101] + y * p[P011] + x * y * p[P111];
}

// Destination node from (drho, u, grad u): the second-order central moments
// carry the non-equilibrium stress of the destination level, -(eps/omega)
// (1+drho) S_ab with the source-level strain rate S and eps the ratio of grid
// spacings (1/2 coarse->fine, 2 fine->coarse); third order is zero, fourth
// and sixth order take their equilibrium values drho/9 and drho/27
// (InterpolationCF.h:180-236, InterpolationFC.h:148-202); then back to
// shifted populations (backward sweeps).
void destination_populations(double drho, const double u[3], const double gx[3], const double gy[3], const double gz[3],
                             double eps, double omega, double* f)
{
    // gx = grad vx, gy = grad vy, gz = grad vz
    const double s = eps / omega * (1.0 + drho);
    const double mxxMyy = -c2o3 * (gx[0] - gy[1]) * s;
    const double mxxMzz = -c2o3 * (gx[0] - gz[2]) * s;
    double m[3][3][3] = {};
    m[0][0][0] = drho;
    m[0][1][1] = -c1o3 * (gy[2] + gz[1]) * s;
    m[1][0][1] = -c1o3 * (gx[2] + gz[0]) * s;
    m[1][1][0] = -c1o3 * (gx[1] + gy[0]) * s;
    m[2][0][0] = c1o3 * (mxxMyy + mxxMzz + drho);
    m[0][2][0] = c1o3 * (-2.0 * mxxMyy + mxxMzz + drho);
    m[0][0][2] = c1o3 * (mxxMyy - 2.0 * mxxMzz + drho);
    m[0][2][2] = m[2][0][2] = m[2][2][0] = drho * c1o9;
    m[2][2][2] = drho * c1o27;
    backward_sweeps(m, u[0], u[1], u[2]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk) f[T.cube[i][j][kk]] = m[i][j][kk];
}

// interpolateCF at (x, y, z) in cell units (fine nodes sit at +-1/4).  The
// pressure carries the curvature term 3 x^2 laplace_rho of :168-169.
void interpolate_cf(const Coefficients& k, double omega_fine, double eps, double x, double y, double z, double* f)
{
    const double drho = poly_value(k.d, x, y, z) + 3.0 * x * x * k.laplace_rho;
    const double u[3] = {poly_value(k.a, x, y, z), poly_value(k.b, x, y, z), poly_value(k.c, x, y, z)};
    double gx[3], gy[3], gz[3];
    poly_gradient(k.a, x, y, z, gx);
    poly_gradient(k.b, x, y, z, gy);
    poly_gradient(k.c, x, y, z, gz);
    destination_populations(drho, u, gx, gy, gz, eps, omega_fine, f);
}

// interpolateFC at the cell centre, pressure d000 - laplace_rho / 4 (:138)
void interpolate_fc(const Coefficients& k, double omega_coarse, double eps, double* f)
{
    const double drho = k.d[P000] - 2.0 * k.laplace_rho * 0.125;
    const double u[3] = {k.a[P000], k.b[P000], k.c[P000]};
    const double gx[3] = {k.a[P100], k.a[P010], k.a[P001]};
    const double gy[3] = {k.b[P100], k.b[P010], k.b[P001]};
    const double gz[3] = {k.c[P100], k.c[P010], k.c[P001]};
    destination_populations(drho, u, gx, gy, gz, eps, omega_coarse, f);
}

int64_t box_node(int x, int y, int z, int nx, int ny, int nz)
{
    x = ((x % nx) + nx) % nx;
    y = ((y % ny) + ny) % ny;
    z = ((z % nz) + nz) % nz;
    return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
}

// calculateOmegaWithturbulentViscosity (TurbulentViscosity.h:86-89) where a
// turbulent viscosity array is given (hasTurbulentViscosity of the scaling)
double node_omega(double omega, const double* nut, int64_t k)
{
    return nut ? omega / (1.0 + 3.0 * omega * nut[k]) : omega;
}

void cell_moments(const double* F, int64_t base, int nx, int ny, int nz, double omega, SourceMoments src[8],
                  const double* nut = nullptr)
{
    const int bx = (int)(base % nx), by = (int)((base / nx) % ny), bz = (int)(base / ((int64_t)nx * ny));
    for (int c = 0; c < 8; ++c) {
        const int64_t k = box_node(bx + (c & 1), by + ((c >> 1) & 1), bz + ((c >> 2) & 1), nx, ny, nz);
        source_moments(F + 27 * k, node_omega(omega, nut, k), src[c]);
    }
}

} // namespace

// ============================================================== C ABI =====
extern "C" {

enum { ORC_K17 = 0, ORC_BGK = 1 };

typedef struct {
    int32_t kind;             // ORC_BC_*
    int32_t n;                // number of boundary nodes
    const int64_t* node;      // dense node index of each boundary node
    const double* q;           // [27*n] structure of arrays; link d of node i is a wall link iff 0 <= q[d*n+i] <= 1
    const double* ux;         // [n] wall velocity (NULL = 0)
    const double* uy;
    const double* uz;
    int32_t cpu_float_velocity;
    const double* rho;        // pressure: prescribed density deviation
    const int64_t* nbr;       // pressure / outflow: inward neighbour node
    int32_t normal;           // pressure / outflow: outward direction
    double correction;        // outflow with pressure correction: density correction factor
    // stress kinds (per boundary node): wall normal, filtered velocities (updated),
    // sampling offset, roughness length, monitors (NULL: none)
    const double* wm_normal[3];
    double* wm_el[3];
    double* wm_wall[3];
    const int32_t* wm_offset;
    const double* wm_z0;
    double* wm_ustar;
    double* wm_force[3];
    // precursor kinds: stencil 0PP/0PM/0MP/0MM per boundary node, plane arrays
    // (components x n_pre), time ratio, velocity ratio
    int32_t n_pre;
    const int64_t* pn[4];
    const double* pw[4];
    const double* plast;
    const double* pcur;
    double tratio;
    double vratio;
    // optional read cycle in closed form: n_series planes (components x n_pre
    // each), plane m at time m * period; at the step of time t = t0 + s the
    // boundary reads last = plane(t / period), current = the next one, with
    // ratio (t - m period) / period (the cycle of runPrecursorBCKernelPost,
    // BoundaryConditionKernelManager.cpp:346-382).  NULL: plast/pcur/tratio fixed.
    const double* pseries;
    int32_t n_series;
    int32_t period;
    int32_t t0;
} orc_bc;

static bool stress_kind(int kind)
{
    return kind >= ORC_BC_STRESS_INTERPOLATED && kind <= ORC_BC_STRESS_BOUNCE_BACK_PRESSURE;
}

// Plane value of component c at boundary node i from the plane array a
// (components x n_pre): the inverse-distance weighted four precursor nodes,
// or the 0PP node alone when its weight is >= 1e6
// (PrecursorNonReflectiveCompressible.cu:82-110, PrecursorDistributions.cu:95-152).
static double plane_value(int n_pre, const int64_t* const pn[4], const double* const pw[4], int i, const double* a,
                          int c)
{
    const double* ac = a + (int64_t)c * n_pre;
    const double w0 = pw[0][i];
    if (!(w0 < 1e6)) return ac[pn[0][i]];
    const double w1 = pw[1][i], w2 = pw[2][i], w3 = pw[3][i];
    const double inv = 1.0 / (w0 + w1 + w2 + w3);
    return (ac[pn[0][i]] * w0 + ac[pn[1][i]] * w1 + ac[pn[2][i]] * w2 + ac[pn[3][i]] * w3) * inv;
}

// Blended plane values out[c*n + i] = last*(1 - tr) + current*tr for every
// boundary node and component (PrecursorDistributions.cu:163-171).
void orc_precursor_blend(int n, int n_pre, int comps, const int64_t* pn0, const int64_t* pn1, const int64_t* pn2,
                         const int64_t* pn3, const double* pw0, const double* pw1, const double* pw2,
                         const double* pw3, const double* last, const double* cur, double tr, double* out)
{
    const int64_t* pn[4] = {pn0, pn1, pn2, pn3};
    const double* pw[4] = {pw0, pw1, pw2, pw3};
    for (int c = 0; c < comps; ++c)
        for (int i = 0; i < n; ++i)
            out[(int64_t)c * n + i] =
                plane_value(n_pre, pn, pw, i, last, c) * (1.0 - tr) + plane_value(n_pre, pn, pw, i, cur, c) * tr;
}

// The inlet velocity of the precursor velocity kind at boundary node i:
// (velocityX + (1 - tr) last + tr current) / velocityRatio
// (PrecursorNonReflectiveCompressible.cu:112-114).
static void precursor_velocity(const orc_bc& bc, int i, double* v3)
{
    const double u[3] = {bc.ux ? bc.ux[i] : 0.0, bc.uy ? bc.uy[i] : 0.0, bc.uz ? bc.uz[i] : 0.0};
    for (int c = 0; c < 3; ++c)
        v3[c] = (u[c] + (1.0 - bc.tratio) * plane_value(bc.n_pre, bc.pn, bc.pw, i, bc.plast, c) +
                 bc.tratio * plane_value(bc.n_pre, bc.pn, bc.pw, i, bc.pcur, c)) /
                bc.vratio;
}

// The nine populations the precursor distributions kind sets, in plane order
static const int PRECURSOR_DIRS[9] = {dP00, dPP0, dPM0, dP0P, dP0M, dPPP, dPMP, dPPM, dPMM};

// Wall-model boundaries (StressCompressible.cu, StressBounceBackCompressible.cu,
// StressBounceBackPressureCompressible.cu with iMEM.cuh:43-141), restated:
// links visited E W N S T B NE SW SE NW TE BW BE TW TN BS BN TS TNE BSW BNE TSW
// TSE BNW BSE TNW; the entering population of a wall link is the interpolated
// no-slip with pressure (mode 0, the NE link's pressure weight 2/27), the
// bounce-back (1) or the bounce-back minus w drho (2); the wall model (the B
// link's q, eps = 0.001f, kappa 0.4, log law, clip at twice the sampled
// tangential velocity) gives a wall velocity whose term 6 w c.U (/ (1 + q) in
// mode 0) is subtracted.  vel: 3*N velocities of this step's collision; the
// sampling node's z velocity is taken from the y array (Stress.cu:57-59).
static const int STRESS_ORDER[26] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 26, 23, 22, 21, 24, 25, 20};

void stress_bc(const orc_bc& bc, int i, const double* p, double omega, const double* vel, int64_t N, double* out,
               bool* written)
{
    const int mode = bc.kind - ORC_BC_STRESS_INTERPOLATED;  // 0, 1, 2
    const double drho = p[dPMP] + p[dMPP] + p[dPPP] + p[dMMP] + p[dPMM] + p[dMPM] + p[dPPM] + p[dMMM] + p[d0PM] +
                        p[d0MP] + p[d0PP] + p[d0MM] + p[dP0M] + p[dM0P] + p[dP0P] + p[dM0M] + p[dPM0] + p[dMP0] +
                        p[dPP0] + p[dMM0] + p[d00P] + p[d00M] + p[d0P0] + p[d0M0] + p[dP00] + p[dM00] + p[d000];
    const double vx1 = (((p[dPMP] - p[dMPM]) - (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                        ((p[dP0M] - p[dM0P]) + (p[dP0P] - p[dM0M])) + ((p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                        (p[dP00] - p[dM00])) / (1.0 + drho);
    const double vx2 = ((-(p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) - (p[dMMP] - p[dPPM])) +
                        ((p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + (-(p[dPM0] - p[dMP0]) + (p[dPP0] - p[dMM0])) +
                        (p[d0P0] - p[d0M0])) / (1.0 + drho);
    const double vx3 = (((p[dPMP] - p[dMPM]) + (p[dMPP] - p[dPMM])) + ((p[dPPP] - p[dMMM]) + (p[dMMP] - p[dPPM])) +
                        (-(p[d0PM] - p[d0MP]) + (p[d0PP] - p[d0MM])) + ((p[dP0P] - p[dM0M]) - (p[dP0M] - p[dM0P])) +
                        (p[d00P] - p[d00M])) / (1.0 + drho);
    const double cu_sq = c3o2 * (vx1 * vx1 + vx2 * vx2 + vx3 * vx3) * (1.0 + drho);
    const double* q = bc.q;
    auto qd = [&](int d) { return (double)q[(int64_t)d * bc.n + i]; };
    double fin[27] = {0};
    double wm[3] = {0, 0, 0};
    auto add = [&](int d, double s) {
        const int c[3] = {CX[d], CY[d], CZ[d]};
        for (int a = 0; a < 3; ++a) {
            if (c[a] > 0) wm[a] += s;
            if (c[a] < 0) wm[a] -= s;
        }
    };
    for (int j = 0; j < 26; ++j) {
        const int d = STRESS_ORDER[j];
        const double qq = qd(d);
        if (!(qq >= 0.0 && qq <= 1.0)) continue;
        const double f = p[d], fi = p[T.inv[d]];
        double v;
        if (mode == 0) {
            const double w = weight(d);
            const double vlb = proj(d, vx1, vx2, vx3);
            const double feq = w * (drho + c9o2 * vlb * vlb * (1.0 + drho) - cu_sq);
            const double wp = d == dPP0 ? 2.0 / 27.0 : w;
            v = (1.0 - qq) / (1.0 + qq) * (f - fi + (f + fi - 2.0 * feq * omega) / (1.0 - omega)) * c1o2 +
                (qq * (f + fi)) / (1.0 + qq) - wp * drho;
        } else if (mode == 1) {
            v = f;
        } else {
            v = f - weight(d) * drho;
        }
        fin[T.inv[d]] = v;
        add(d, f + v);
    }
    // iMEM
    const int64_t kN = bc.nbr[i];
    const double qB = qd(d00M);
    const double forceFactor = mode == 0 ? 1.0 + qB : 1.0;
    const double eps = (double)0.001f;
    const double nrm[3] = {bc.wm_normal[0][i], bc.wm_normal[1][i], bc.wm_normal[2][i]};
    const double vs[3] = {vel[kN], vel[N + kN], vel[N + kN]}; // z from the y array
    const double vw[3] = {vx1, vx2, vx3};
    double el[3], mw[3];
    for (int a = 0; a < 3; ++a) {
        el[a] = eps * vs[a] + (1.0 - eps) * bc.wm_el[a][i];
        bc.wm_el[a][i] = el[a];
    }
    for (int a = 0; a < 3; ++a) {
        mw[a] = eps * vw[a] + (1.0 - eps) * bc.wm_wall[a][i];
        bc.wm_wall[a][i] = mw[a];
    }
    const double dn_el = el[0] * nrm[0] + el[1] * nrm[1] + el[2] * nrm[2];
    for (int a = 0; a < 3; ++a) el[a] -= dn_el * nrm[a];
    const double mag_el = std::sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
    const double dn_mw = mw[0] * nrm[0] + mw[1] * nrm[1] + mw[2] * nrm[2];
    for (int a = 0; a < 3; ++a) mw[a] -= dn_mw * nrm[a];
    const double mag_mw = std::sqrt(mw[0] * mw[0] + mw[1] * mw[1] + mw[2] * mw[2]);
    const double dn_w = vw[0] * nrm[0] + vw[1] * nrm[1] + vw[2] * nrm[2];
    const double wt[3] = {vw[0] - dn_w * nrm[0], vw[1] - dn_w * nrm[1], vw[2] - dn_w * nrm[2]};
    const double z = (double)bc.wm_offset[i] + qB;
    const double u_star = mag_el * 0.4 / (std::log(z / bc.wm_z0[i]));
    if (bc.wm_ustar) bc.wm_ustar[i] = u_star;
    const double tau_w = u_star * u_star;
    const double mdn = wm[0] * nrm[0] + wm[1] * nrm[1] + wm[2] * nrm[2];
    const double rho = 1.0 + drho;
    double U[3];
    for (int a = 0; a < 3; ++a) {
        const double Fw = (tau_w * 1.0) * (wt[a] / mag_mw);
        const double F = Fw - (wm[a] - mdn * nrm[a]) / rho;
        const double c = 2.0 * el[a];
        const double v = -3.0 * F * forceFactor;
        U[a] = c > -c ? std::fmin(c, std::fmax(-c, v)) : std::fmax(c, std::fmin(-c, v));
    }
    for (int j = 0; j < 26; ++j) {
        const int d = STRESS_ORDER[j];
        const double qq = qd(d);
        if (!(qq >= 0.0 && qq <= 1.0)) continue;
        double t = 6.0 * weight(d) * proj(d, U[0], U[1], U[2]);
        if (mode == 0) t = t / (1.0 + qq);
        const int e = T.inv[d];
        out[e] = fin[e] - t;
        written[e] = true;
        add(d, -t);
    }
    if (bc.wm_force[0])
        for (int a = 0; a < 3; ++a) bc.wm_force[a][i] = wm[a];
}

static bool pre_collision_kind(int kind)
{
    return kind == ORC_BC_PRESSURE_NEQ || kind == ORC_BC_OUTFLOW_NON_REFLECTING ||
           kind == ORC_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION;
}

void orc_k17_collide_turb(double* f27, double omega, const double* lim3, const double* force3, double* macro4,
                          int turb, double sgs, double* nut)
{
    static const double zero[3] = {0, 0, 0};
    k17_collide(f27, omega, lim3, force3 ? force3 : zero, macro4, turb, sgs, nut);
}

// one boundary node: post-collision populations p27 (interpolated kinds) or
// pre-collision f27 / neighbour fn27 (outflow kinds); out27 / written27 (0/1)
void orc_interp_bc(int kind, const double* p27, const double* q27, double omega, double vx, double vy, double vz,
                   double* out27, int32_t* written27)
{
    bool wr[27] = {false};
    interp_bc(kind, p27, q27, 1, 0, omega, vx, vy, vz, out27, wr);
    for (int d = 0; d < 27; ++d) written27[d] = wr[d];
}
// non-equilibrium pressure boundary of one node from its inward neighbour's
// populations fn27 (PressureNonEquilibriumCompressible.cu)
void orc_pressure_bc(const double* fn27, double rho_bc, int normal, double* out27, int32_t* written27)
{
    bool wr[27] = {false};
    pressure_bc(fn27, rho_bc, normal, out27, wr);
    for (int d = 0; d < 27; ++d) written27[d] = wr[d];
}
void orc_outflow_bc(int kind, const double* f27, const double* fn27, int normal, double factor, double* out27,
                    int32_t* written27)
{
    bool wr[27] = {false};
    outflow_bc(kind, f27, fn27, normal, factor, out27, wr);
    for (int d = 0; d < 27; ++d) written27[d] = wr[d];
}

// calculateMacroscopicQuantitiesCompressible_device per node
// (MacroscopicQuantities.cu:222-236 with MacroscopicQuantities.h:55-149):
// out5 = {rho, vx, vy, vz, pressure}
void orc_macroscopic_quantities(const double* f, double* out5)
{
    const double rho = density(f);
    const double vx = mom_x(f) / (rho + 1.0), vy = mom_y(f) / (rho + 1.0), vz = mom_z(f) / (rho + 1.0);
    const double press = (f[dP00] + f[dM00] + f[d0P0] + f[d0M0] + f[d00P] + f[d00M] +
                          2.0 * (f[dPP0] + f[dMM0] + f[dPM0] + f[dMP0] + f[dP0P] + f[dM0M] + f[dP0M] + f[dM0P] +
                                 f[d0PP] + f[d0MM] + f[d0PM] + f[d0MP]) +
                          3.0 * (f[dPPP] + f[dMMP] + f[dPMP] + f[dMPP] + f[dPPM] + f[dMMM] + f[dPMM] + f[dMPM]) -
                          rho - (vx * vx + vy * vy + vz * vz) * (1.0 + rho)) *
                             c1o2 +
                         rho;
    out5[0] = rho;
    out5[1] = vx;
    out5[2] = vy;
    out5[3] = vz;
    out5[4] = press;
}

// calculateMeanCompressible_device per fluid node (MacroscopicQuantities.cu:390-536):
// acc5 = {rho, vx, vy, vz, pressure} running sums in and out.  The density and
// velocity of f are added to the sums in the kernel's summation order; the
// pressure sum gets the pressure expression with the new density and velocity
// sums in place of the node's own (the kernel reads rhoD/vxD back after
// writing them).
void orc_mean_add(const double* f, double* acc5)
{
    const double drho =
        ((((f[dPPP] + f[dMMM]) + (f[dMPM] + f[dPMP])) + ((f[dMPP] + f[dPMM]) + (f[dMMP] + f[dPPM]))) +
         (((f[d0MP] + f[d0PM]) + (f[d0MM] + f[d0PP])) + ((f[dM0P] + f[dP0M]) + (f[dM0M] + f[dP0P])) +
          ((f[dMP0] + f[dPM0]) + (f[dMM0] + f[dPP0]))) +
         ((f[dM00] + f[dP00]) + (f[d0M0] + f[d0P0]) + (f[d00M] + f[d00P]))) +
        f[d000];
    const double rho = 1.0 + drho;
    const double r = drho + acc5[0];
    const double x = (((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dPMM] - f[dMPP]) + (f[dPPM] - f[dMMP]))) +
                       (((f[dP0M] - f[dM0P]) + (f[dP0P] - f[dM0M])) + ((f[dPM0] - f[dMP0]) + (f[dPP0] - f[dMM0]))) +
                       (f[dP00] - f[dM00])) /
                      rho) +
                     acc5[1];
    const double y = (((((f[dPPP] - f[dMMM]) + (f[dMPM] - f[dPMP])) + ((f[dMPP] - f[dPMM]) + (f[dPPM] - f[dMMP]))) +
                       (((f[d0PM] - f[d0MP]) + (f[d0PP] - f[d0MM])) + ((f[dMP0] - f[dPM0]) + (f[dPP0] - f[dMM0]))) +
                       (f[d0P0] - f[d0M0])) /
                      rho) +
                     acc5[2];
    const double z = (((((f[dPPP] - f[dMMM]) + (f[dPMP] - f[dMPM])) + ((f[dMPP] - f[dPMM]) + (f[dMMP] - f[dPPM]))) +
                       (((f[d0MP] - f[d0PM]) + (f[d0PP] - f[d0MM])) + ((f[dM0P] - f[dP0M]) + (f[dP0P] - f[dM0M]))) +
                       (f[d00P] - f[d00M])) /
                      rho) +
                     acc5[3];
    acc5[4] = (f[dP00] + f[dM00] + f[d0P0] + f[d0M0] + f[d00P] + f[d00M] +
               2.0 * (f[dPP0] + f[dMM0] + f[dPM0] + f[dMP0] + f[dP0P] + f[dM0M] + f[dP0M] + f[dM0P] + f[d0PP] +
                      f[d0MM] + f[d0PM] + f[d0MP]) +
               3.0 * (f[dPPP] + f[dMMP] + f[dPMP] + f[dMPP] + f[dPPM] + f[dMMM] + f[dPMM] + f[dMPM]) - r -
               (x * x + y * y + z * z) * (1.0 + r)) *
                  c1o2 +
              r + acc5[4];
    acc5[0] = r;
    acc5[1] = x;
    acc5[2] = y;
    acc5[3] = z;
}

// one node of a wall-model boundary: post-collision p27, q27, the rate, the
// sampling node's velocity vs3 (its z component as the kernels read it),
// normal3, offset, z0; el3 / wall3 filtered velocities in and out; out27,
// written27; ustar1 and force3 monitors (may be NULL)
void orc_stress_bc(int kind, const double* p27, const double* q27, double omega, const double* vs3, const double* normal3,
                   int32_t offset, double z0, double* el3, double* wall3, double* out27, int32_t* written27,
                   double* ustar1, double* force3)
{
    orc_bc bc{};
    bc.kind = kind;
    bc.n = 1;
    bc.q = q27;
    int64_t kN = 0;
    bc.nbr = &kN;
    for (int a = 0; a < 3; ++a) {
        bc.wm_normal[a] = normal3 + a;
        bc.wm_el[a] = el3 + a;
        bc.wm_wall[a] = wall3 + a;
        bc.wm_force[a] = force3 ? force3 + a : nullptr;
    }
    bc.wm_offset = &offset;
    bc.wm_z0 = &z0;
    bc.wm_ustar = ustar1;
    const double vel[3] = {vs3[0], vs3[1], vs3[2]};
    bool wr[27] = {false};
    stress_bc(bc, 0, p27, omega, vel, 1, out27, wr);
    for (int d = 0; d < 27; ++d) written27[d] = wr[d];
}

void orc_k17_collide(double* f27, double omega, const double* lim3, const double* force3, double* macro4)
{
    static const double zero[3] = {0, 0, 0};
    k17_collide(f27, omega, lim3, force3 ? force3 : zero, macro4);
}

void orc_bgk_collide(double* f27, double omega, double* macro4) { bgk_collide(f27, omega, macro4); }

void orc_chimera(int kind, double* abc, double vv, double v2, double kinv, double k)
{
    switch (kind) {
    case 0: fwd_k(abc[0], abc[1], abc[2], vv, v2, kinv, k); break;
    case 1: bwd_k(abc[0], abc[1], abc[2], vv, v2, kinv, k); break;
    case 2: fwd(abc[0], abc[1], abc[2], vv, v2); break;
    default: bwd(abc[0], abc[1], abc[2], vv, v2); break;
    }
}

// out4 = {drho, vx, vy, vz} (compressible: velocities divided by 1+drho)
void orc_macroscopic(const double* f27, double* out4)
{
    const double drho = density(f27);
    const double oor = 1.0 / (drho + 1.0);
    out4[0] = drho;
    out4[1] = mom_x(f27) * oor;
    out4[2] = mom_y(f27) * oor;
    out4[3] = mom_z(f27) * oor;
}

void orc_equilibrium(double* f27, double drho, double ux, double uy, double uz) { equilibrium(f27, drho, ux, uy, uz); }

void orc_init_field(double* F, int64_t n, const double* drho, const double* ux, const double* uy, const double* uz)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) equilibrium(F + 27 * i, drho[i], ux[i], uy[i], uz[i]);
}

double orc_wall_term(int d, double ux, double uy, double uz, int cpu_float) { return wall_term(d, ux, uy, uz, cpu_float != 0); }

// Run nsteps on a dense nx*ny*nz box (x fastest, periodic index wrap on every
// axis; walls are solid nodes).  F holds node-major pre-collision populations
// and is overwritten with the state after nsteps.  macro (optional, 4*N)
// receives {drho,vx,vy,vz} of the last collision.  Returns 0, or a negative
// code when a fluid node streams into a solid node through a link no boundary
// condition covers.
static int run_impl(double* F, int nx, int ny, int nz, const uint8_t* solid, int collision, double omega,
                    const double* lim3, const double* force3, int n_bc, const orc_bc* bcs, int nsteps, double* macro,
                    int turb, double sgs, double* nut, const double* gforce = nullptr, double ff = 1.0,
                    double* bf = nullptr);

// orc_run with ApplyBodyForce (RunCollision.cuh:88-97): node i's force is
// ((gforce + bf[a*N + i]) / 2) * ff, and the collision resets bf to 0
int orc_run_body_force(double* F, int nx, int ny, int nz, const uint8_t* solid, double omega, const double* lim3,
                       const double* gforce3, double ff, double* bf, int n_bc, const orc_bc* bcs, int nsteps)
{
    return run_impl(F, nx, ny, nz, solid, ORC_K17, omega, lim3, nullptr, n_bc, bcs, nsteps, nullptr, TURB_NONE, 0.0,
                    nullptr, gforce3, ff, bf);
}

int orc_run(double* F, int nx, int ny, int nz, const uint8_t* solid, int collision, double omega,
            const double* lim3, const double* force3, int n_bc, const orc_bc* bcs, int nsteps, double* macro)
{
    return run_impl(F, nx, ny, nz, solid, collision, omega, lim3, force3, n_bc, bcs, nsteps, macro, TURB_NONE, 0.0,
                    nullptr);
}

// orc_run with an eddy-viscosity model of the K17 collision: nut (N values)
// holds the turbulent viscosity of every node in and out.  Smagorinsky / QR
// update it inside the collision; AMD recomputes it from the velocities the
// collision wrote after the streaming and the post-collision boundary
// conditions, before the pre-collision ones (UpdateGrid27.cpp:69-85).
int orc_run_turb(double* F, int nx, int ny, int nz, const uint8_t* solid, double omega, const double* lim3,
                 const double* force3, int n_bc, const orc_bc* bcs, int nsteps, double* macro, int turb, double sgs,
                 double* nut)
{
    return run_impl(F, nx, ny, nz, solid, ORC_K17, omega, lim3, force3, n_bc, bcs, nsteps, macro, turb, sgs, nut);
}

// calcAMD (src/gpu/core/PostProcessor/TurbulentViscosityKernels.cu:52-113) on
// a dense periodic box: fluid = not solid; derivatives central where both
// neighbours along an axis are fluid, one-sided where one is, 0 where none is.
void orc_turbulent_viscosity_amd(int nx, int ny, int nz, const uint8_t* solid, const double* vx, const double* vy,
                                 const double* vz, double C, double* nut)
{
    const int64_t N = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; ++k) {
        if (solid && solid[k]) continue;
        const int x = (int)(k % nx), y = (int)((k / nx) % ny), z = (int)(k / ((int64_t)nx * ny));
        auto at = [&](int a, int b, int c) {
            return (int64_t)((a + nx) % nx) + (int64_t)nx * ((int64_t)((b + ny) % ny) + (int64_t)ny * ((c + nz) % nz));
        };
        const int64_t kp[3] = {at(x + 1, y, z), at(x, y + 1, z), at(x, y, z + 1)};
        const int64_t km[3] = {at(x - 1, y, z), at(x, y - 1, z), at(x, y, z - 1)};
        double g[3][3]; // g[i][a] = d u_i / d x_a
        for (int a = 0; a < 3; ++a) {
            const bool fp = !(solid && solid[kp[a]]), fm = !(solid && solid[km[a]]);
            const double div = (fm && fp) ? c1o2 : 1.0;
            const double* v[3] = {vx, vy, vz};
            for (int i = 0; i < 3; ++i) g[i][a] = ((fp ? v[i][kp[a]] : v[i][k]) - (fm ? v[i][km[a]] : v[i][k])) * div;
        }
        const double dvxdx = g[0][0], dvxdy = g[0][1], dvxdz = g[0][2];
        const double dvydx = g[1][0], dvydy = g[1][1], dvydz = g[1][2];
        const double dvzdx = g[2][0], dvzdy = g[2][1], dvzdz = g[2][2];
        const double den = dvxdx * dvxdx + dvydx * dvydx + dvzdx * dvzdx + dvxdy * dvxdy + dvydy * dvydy +
                           dvzdy * dvzdy + dvxdz * dvxdz + dvydz * dvydz + dvzdz * dvzdz;
        const double num = (dvxdx * dvxdx + dvxdy * dvxdy + dvxdz * dvxdz) * dvxdx +
                           (dvydx * dvydx + dvydy * dvydy + dvydz * dvydz) * dvydy +
                           (dvzdx * dvzdx + dvzdy * dvzdy + dvzdz * dvzdz) * dvzdz +
                           (dvxdx * dvydx + dvxdy * dvydy + dvxdz * dvydz) * (dvxdy + dvydx) +
                           (dvxdx * dvzdx + dvxdy * dvzdy + dvxdz * dvzdz) * (dvxdz + dvzdx) +
                           (dvydx * dvzdx + dvydy * dvzdy + dvydz * dvzdz) * (dvydz + dvzdy);
        nut[k] = den != 0.0 ? std::max(0.0, -C * num) / den : 0.0;
    }
}

static int run_impl(double* F, int nx, int ny, int nz, const uint8_t* solid, int collision, double omega,
                    const double* lim3, const double* force3, int n_bc, const orc_bc* bcs, int nsteps, double* macro,
                    int turb, double sgs, double* nut, const double* gforce, double ff, double* bf)
{
    const int64_t N = (int64_t)nx * ny * nz;
    // a private copy of the tables: precursor plane series set plast/pcur/tratio per step
    std::vector<orc_bc> bcv(bcs, bcs + n_bc);
    bcs = bcv.data();
    bool stress = false;
    for (int b = 0; b < n_bc; ++b) stress = stress || stress_kind(bcs[b].kind);
    std::vector<double> vel(turb == TURB_AMD || stress ? (size_t)N * 3 : 0);
    std::vector<double> post((size_t)N * 27), G((size_t)N * 27);
    static const double zero[3] = {0, 0, 0};
    if (!force3) force3 = zero;

    // link coverage: covered[i*27+d] = 1 if node i has a boundary condition on link d
    std::vector<uint8_t> covered;
    if (solid) {
        covered.assign((size_t)N * 27, 0);
        for (int b = 0; b < n_bc; ++b)
            for (int i = 0; i < bcs[b].n; ++i)
                for (int d = 1; d < 27; ++d) {
                    if (pre_collision_kind(bcs[b].kind)) {
                        // populations leaving through the pressure / outflow boundary are dropped
                        const int nm = bcs[b].normal;
                        if (CX[d] * CX[nm] + CY[d] * CY[nm] + CZ[d] * CZ[nm] > 0)
                            covered[(size_t)bcs[b].node[i] * 27 + d] = 1;
                        continue;
                    }
                    if (bcs[b].kind == ORC_BC_PRECURSOR_DISTRIBUTIONS) {
                        // the populations entering with a positive x component are set
                        if (CX[d] < 0) covered[(size_t)bcs[b].node[i] * 27 + d] = 1;
                        continue;
                    }
                    const double q = bcs[b].q[(int64_t)d * bcs[b].n + i];
                    if (q >= 0.0 && q <= 1.0) covered[(size_t)bcs[b].node[i] * 27 + d] = 1;
                }
    }
    auto neighbour = [&](int64_t i, int d) {
        int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        x = (x + CX[d] + nx) % nx;
        y = (y + CY[d] + ny) % ny;
        z = (z + CZ[d] + nz) % nz;
        return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
    };

    // precursor velocity kind: the value computed at step s is written over
    // the outgoing population it read (the parity's own slot) and enters the
    // node two steps later; pv[b][parity] holds it in between.  The step after
    // the first reads the element as it was initially, the population of the
    // node across the link (kept, since that node does not collide).
    std::vector<std::array<std::vector<double>, 2>> pv(n_bc);
    std::vector<std::array<std::vector<uint8_t>, 2>> pv_wr(n_bc);
    for (int s = 0; s < nsteps; ++s) {
        for (orc_bc& bc : bcv) {
            if (!bc.pseries) continue;
            const int t = bc.t0 + s, m = t / bc.period;
            if (m + 1 >= bc.n_series) return -2; // the series ends before this step
            const int64_t psz = (int64_t)(bc.kind == ORC_BC_PRECURSOR_VELOCITY ? 3 : 9) * bc.n_pre;
            bc.plast = bc.pseries + m * psz;
            bc.pcur = bc.pseries + (m + 1) * psz;
            bc.tratio = (double)(t - m * bc.period) / bc.period;
        }
        // non-reflecting outflow: the populations entering the boundary nodes
        // in the next state come from this pre-collision state (the kernel
        // runs after a parity flip and writes the other parity's slots, which
        // the following collision neither reads nor writes)
        std::vector<std::vector<double>> pend(n_bc);
        std::vector<std::vector<uint8_t>> pend_wr(n_bc);
        for (int b = 0; b < n_bc; ++b) {
            const orc_bc& bc = bcs[b];
            if (bc.kind != ORC_BC_OUTFLOW_NON_REFLECTING && bc.kind != ORC_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION)
                continue;
            pend[b].assign((size_t)bc.n * 27, 0.0);
            pend_wr[b].assign((size_t)bc.n * 27, 0);
            for (int i = 0; i < bc.n; ++i) {
                bool wr[27] = {false};
                outflow_bc(bc.kind, F + 27 * bc.node[i], F + 27 * bc.nbr[i], bc.normal, bc.correction,
                           &pend[b][(size_t)i * 27], wr);
                for (int d = 0; d < 27; ++d) pend_wr[b][(size_t)i * 27 + d] = wr[d];
            }
        }
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int64_t i = 0; i < N; ++i) {
            if (solid && solid[i]) continue;
            double f[27];
            std::memcpy(f, F + 27 * i, sizeof(f));
            double mv[4];
            double fnode[3];
            const double* fo = force3;
            if (bf) {
                for (int a = 0; a < 3; ++a) {
                    fnode[a] = (gforce[a] + bf[(size_t)a * N + i]) * c1o2 * ff;
                    bf[(size_t)a * N + i] = 0.0;
                }
                fo = fnode;
            }
            if (collision == ORC_K17) k17_collide(f, omega, lim3, fo, mv, turb, sgs, nut ? nut + i : nullptr);
            else bgk_collide(f, omega, mv);
            if (!vel.empty())
                for (int a = 0; a < 3; ++a) vel[(size_t)a * N + i] = mv[1 + a];
            std::memcpy(&post[(size_t)i * 27], f, sizeof(f));
            if (macro && s == nsteps - 1) std::memcpy(macro + 4 * i, mv, sizeof(mv));
        }
        // push streaming
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int64_t i = 0; i < N; ++i) {
            if (solid && solid[i]) continue;
            for (int d = 0; d < 27; ++d) {
                const int64_t j = neighbour(i, d);
                if (solid && solid[j]) {
                    // solid = 2 marks an inactive node (e.g. a coarse node under a
                    // finer level): links into it are dropped without a boundary
                    if (solid[j] == 1 && !covered[(size_t)i * 27 + d]) bad += 1;
                    continue;
                }
                G[(size_t)j * 27 + d] = post[(size_t)i * 27 + d];
            }
        }
        if (bad) return -1;
        for (int b = 0; b < n_bc; ++b)
            for (int i = 0; i < (int)(pend_wr[b].size() / 27); ++i)
                for (int d = 0; d < 27; ++d)
                    if (pend_wr[b][(size_t)i * 27 + d]) G[(size_t)bcs[b].node[i] * 27 + d] = pend[b][(size_t)i * 27 + d];
        // post-collision boundary conditions (writes the population that
        // re-enters node i against the wall link d)
        for (int b = 0; b < n_bc; ++b) {
            const orc_bc& bc = bcs[b];
            if (pre_collision_kind(bc.kind)) continue;
            for (int i = 0; i < bc.n; ++i) {
                const int64_t k = bc.node[i];
                const double ux = bc.ux ? bc.ux[i] : 0.0, uy = bc.uy ? bc.uy[i] : 0.0, uz = bc.uz ? bc.uz[i] : 0.0;
                if (bc.kind == ORC_BC_VELOCITY_BOUNCE_BACK) {
                    for (int d = 1; d < 27; ++d) {
                        const double q = bc.q[(int64_t)d * bc.n + i];
                        if (!(q >= 0.0f && q <= 1.0f)) continue;
                        const double fd = post[(size_t)k * 27 + d];
                        G[(size_t)k * 27 + T.inv[d]] = fd - wall_term(d, ux, uy, uz, bc.cpu_float_velocity != 0);
                    }
                    continue;
                }
                if (bc.kind == ORC_BC_PRECURSOR_DISTRIBUTIONS) {
                    for (int c = 0; c < 9; ++c)
                        G[(size_t)k * 27 + PRECURSOR_DIRS[c]] =
                            plane_value(bc.n_pre, bc.pn, bc.pw, i, bc.plast, c) * (1.0 - bc.tratio) +
                            plane_value(bc.n_pre, bc.pn, bc.pw, i, bc.pcur, c) * bc.tratio;
                    continue;
                }
                double out[27];
                bool wr[27] = {false};
                if (bc.kind == ORC_BC_PRECURSOR_VELOCITY) {
                    auto& now = pv[b][s & 1];
                    auto& now_wr = pv_wr[b][s & 1];
                    const auto& prev = pv[b][(s + 1) & 1];
                    const auto& prev_wr = pv_wr[b][(s + 1) & 1];
                    if (now.empty()) {
                        now.assign((size_t)bc.n * 27, 0.0);
                        now_wr.assign((size_t)bc.n * 27, 0);
                    }
                    for (int d = 1; d < 27; ++d) {
                        const double q = bc.q[(int64_t)d * bc.n + i];
                        if (!(q >= 0.0f && q <= 1.0f)) continue;
                        const int e = T.inv[d];
                        if (s == 0) G[(size_t)k * 27 + e] = F[27 * neighbour(k, d) + d];
                        else if (prev_wr[(size_t)i * 27 + e]) G[(size_t)k * 27 + e] = prev[(size_t)i * 27 + e];
                    }
                    double v3[3];
                    precursor_velocity(bc, i, v3);
                    interp_bc(bc.kind, &post[(size_t)k * 27], bc.q, bc.n, i, omega, v3[0], v3[1], v3[2], out, wr);
                    for (int d = 0; d < 27; ++d) {
                        now[(size_t)i * 27 + d] = out[d];
                        now_wr[(size_t)i * 27 + d] = wr[d];
                    }
                    continue;
                }
                if (stress_kind(bc.kind)) {
                    const double om = nut ? omega / (1.0 + 3.0 * omega * std::max(0.0, nut[k])) : omega;
                    stress_bc(bc, i, &post[(size_t)k * 27], om, vel.data(), N, out, wr);
                    for (int d = 0; d < 27; ++d)
                        if (wr[d]) G[(size_t)k * 27 + d] = out[d];
                    continue;
                }
                // SlipTurbulentViscosityCompressible.cu:130-134: the node's own rate
                const double om = bc.kind == ORC_BC_SLIP_TURBULENT_VISCOSITY && nut
                                      ? omega / (1.0 + 3.0 * omega * std::max(0.0, nut[k]))
                                      : omega;
                interp_bc(bc.kind, &post[(size_t)k * 27], bc.q, bc.n, i, om, ux, uy, uz, out, wr);
                for (int d = 0; d < 27; ++d)
                    if (wr[d]) G[(size_t)k * 27 + d] = out[d];
            }
        }
        // solid nodes keep whatever they had (never read)
        if (solid)
            for (int64_t i = 0; i < N; ++i)
                if (solid[i]) std::memcpy(&G[(size_t)i * 27], F + 27 * i, 27 * sizeof(double));
        std::memcpy(F, G.data(), (size_t)N * 27 * sizeof(double));
        if (turb == TURB_AMD)
            orc_turbulent_viscosity_amd(nx, ny, nz, solid, vel.data(), vel.data() + N, vel.data() + 2 * N, sgs, nut);
        // pre-collision boundary conditions of the next step (after the parity
        // flip, UpdateGrid27.cpp:288-307): read the inward neighbours first
        for (int b = 0; b < n_bc; ++b) {
            const orc_bc& bc = bcs[b];
            if (bc.kind != ORC_BC_PRESSURE_NEQ) continue;
            std::vector<double> outs((size_t)bc.n * 27);
            std::vector<uint8_t> wrs((size_t)bc.n * 27);
            for (int i = 0; i < bc.n; ++i) {
                bool wr[27] = {false};
                pressure_bc(F + 27 * bc.nbr[i], bc.rho[i], bc.normal, &outs[(size_t)i * 27], wr);
                for (int d = 0; d < 27; ++d) wrs[(size_t)i * 27 + d] = wr[d];
            }
            for (int i = 0; i < bc.n; ++i)
                for (int d = 0; d < 27; ++d)
                    if (wrs[(size_t)i * 27 + d]) F[27 * bc.node[i] + d] = outs[(size_t)i * 27 + d];
        }
    }
    return 0;
}

// Global mass deviation and momentum of node-major populations (long double
// accumulation): out4 = {sum drho, sum jx, sum jy, sum jz} over non-solid nodes.
void orc_mass_momentum(const double* F, int64_t n, const uint8_t* solid, double* out4)
{
    long double a[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        if (solid && solid[i]) continue;
        const double* f = F + 27 * i;
        a[0] += density(f);
        a[1] += mom_x(f);
        a[2] += mom_y(f);
        a[3] += mom_z(f);
    }
    for (int c = 0; c < 4; ++c) out4[c] = (double)a[c];
}

// KineticEnergyAnalyzer.cu:137-154 / EnstrophyAnalyzer.cu:134-336 on a
// structured periodic box: out2 = {sum, count of nodes with a fluid 7-stencil}.
void orc_energy(const double* F, int nx, int ny, int nz, const uint8_t* solid, int enstrophy, double* out2)
{
    const int64_t N = (int64_t)nx * ny * nz;
    std::vector<double> r(N), u[3];
    for (int a = 0; a < 3; ++a) u[a].resize(N);
    for (int64_t i = 0; i < N; ++i) {
        const double* f = F + 27 * i;
        r[i] = density(f);
        const double oor = 1.0 / (r[i] + 1.0);
        u[0][i] = mom_x(f) * oor;
        u[1][i] = mom_y(f) * oor;
        u[2][i] = mom_z(f) * oor;
    }
    auto idx = [&](int x, int y, int z) {
        x = ((x % nx) + nx) % nx; y = ((y % ny) + ny) % ny; z = ((z % nz) + nz) % nz;
        return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
    };
    auto fluid = [&](int64_t k) { return !(solid && solid[k]); };
    long double sum = 0, cnt = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t k = idx(x, y, z);
                if (!fluid(k)) continue;
                const int p[3] = {x, y, z};
                auto at = [&](int ax, int s) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[ax] += s;
                    return idx(q[0], q[1], q[2]);
                };
                bool ok = true;
                for (int ax = 0; ax < 3; ++ax)
                    if (!fluid(at(ax, 1)) || !fluid(at(ax, -1))) ok = false;
                if (!ok) continue;
                if (!enstrophy) {
                    sum += 0.5 * (u[0][k] * u[0][k] + u[1][k] * u[1][k] + u[2][k] * u[2][k]) * (r[k] + 1.0);
                    cnt += 1;
                    continue;
                }
                auto deriv = [&](int comp, int ax) {
                    const std::vector<double>& v = u[comp];
                    int order = 8;
                    if (!fluid(at(ax, 2)) || !fluid(at(ax, -2))) order = 2;
                    if ((!fluid(at(ax, 3)) || !fluid(at(ax, -3))) && order > 4) order = 4;
                    if ((!fluid(at(ax, 4)) || !fluid(at(ax, -4))) && order > 6) order = 6;
                    auto d = [&](int s) { return v[at(ax, s)] - v[at(ax, -s)]; };
                    if (order == 2) return d(1) / 2.0;
                    if (order == 4) return ((8.0 * v[at(ax, 1)] - 8.0 * v[at(ax, -1)]) - d(2)) / 12.0;
                    if (order == 6) return (45.0 * d(1) - 9.0 * d(2) + d(3)) / 60.0;
                    return (224.0 * d(1) - 56.0 * d(2) + (32.0 / 3.0) * d(3) - d(4)) / 280.0;
                };
                const double wx = deriv(2, 1) - deriv(1, 2);
                const double wy = deriv(0, 2) - deriv(2, 0);
                const double wz = deriv(1, 0) - deriv(0, 1);
                sum += 0.5 * (r[k] + 1.0) * (wx * wx + wy * wy + wz * wz);
                cnt += 1;
            }
    out2[0] = (double)sum;
    out2[1] = (double)cnt;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}


// ---- refinement (test infrastructure)
void orc_source_moments(const double* f27, double omega, double* out9)
{
    SourceMoments s;
    source_moments(f27, omega, s);
    const double v[9] = {s.drho, s.vx, s.vy, s.vz, s.kxy, s.kyz, s.kxz, s.kxxMyy, s.kxxMzz};
    std::memcpy(out9, v, sizeof(v));
}

// out42: a[11], b[11], c[11] in the slot order 000,100,010,001,200,020,002,
// 110,101,011,111; d[8] in 000,100,010,001,110,101,011,111; laplace_rho
static void pack_coefficients(const Coefficients& k, double* out42)
{
    int n = 0;
    for (const double* p : {k.a, k.b, k.c})
        for (int i = 0; i < NP; ++i) out42[n++] = p[i];
    for (int i : {P000, P100, P010, P001, P110, P101, P011, P111}) out42[n++] = k.d[i];
    out42[n] = k.laplace_rho;
}
static void unpack_coefficients(const double* in42, Coefficients& k)
{
    int n = 0;
    for (double* p : {k.a, k.b, k.c})
        for (int i = 0; i < NP; ++i) p[i] = in42[n++];
    for (int i : {P000, P100, P010, P001, P110, P101, P011, P111}) k.d[i] = in42[n++];
    k.d[P200] = k.d[P020] = k.d[P002] = 0.0;
    k.laplace_rho = in42[n];
}

// f8x27: the 8 corner nodes (corner bit order) of one cell
void orc_interp_coefficients(const double* f8x27, double omega, double xoff, double yoff, double zoff, double* out42)
{
    SourceMoments src[8];
    for (int c = 0; c < 8; ++c) source_moments(f8x27 + 27 * c, omega, src[c]);
    Coefficients k;
    interpolation_coefficients(src, xoff, yoff, zoff, k);
    pack_coefficients(k, out42);
}

// as orc_interp_coefficients with one omega per corner
void orc_interp_coefficients_omegas(const double* f8x27, const double* omega8, double xoff, double yoff, double zoff,
                                    double* out42)
{
    SourceMoments src[8];
    for (int c = 0; c < 8; ++c) source_moments(f8x27 + 27 * c, omega8[c], src[c]);
    Coefficients k;
    interpolation_coefficients(src, xoff, yoff, zoff, k);
    pack_coefficients(k, out42);
}

void orc_interpolate_cf(const double* coeff42, double omega_fine, double eps, double x, double y, double z, double* f27)
{
    Coefficients k;
    unpack_coefficients(coeff42, k);
    interpolate_cf(k, omega_fine, eps, x, y, z, f27);
}

void orc_interpolate_fc(const double* coeff42, double omega_coarse, double eps, double* f27)
{
    Coefficients k;
    unpack_coefficients(coeff42, k);
    interpolate_fc(k, omega_coarse, eps, f27);
}

// Coarse-to-fine over a list of cells (scaleCoarseToFineCompressible.cu:305-529):
// node-major pre-collision populations of periodic boxes; coarse cell with
// corner MMM coarse_cells[i]; its 8 fine nodes form the fine cell with corner
// MMM fine_cells[i] and receive the polynomial at (+-1/4, +-1/4, +-1/4)
// (eps = 1/2, fine omega).  Offsets may be NULL (zero).
// nut_c / nut_f (may be NULL): turbulent viscosity of the coarse / fine level;
// given, every source corner and destination node uses its own node's omega
// (scaleCoarseToFineCompressible.cu:94, 369).
void orc_scale_cf_turb(const double* Fc, double* Ff, const int* dims_c, const int* dims_f, int64_t n_cells,
                       const int64_t* coarse_cells, const int64_t* fine_cells, const double* ox, const double* oy,
                       const double* oz, double omega_c, double omega_f, const double* nut_c, const double* nut_f)
{
    for (int64_t i = 0; i < n_cells; ++i) {
        SourceMoments src[8];
        cell_moments(Fc, coarse_cells[i], dims_c[0], dims_c[1], dims_c[2], omega_c, src, nut_c);
        Coefficients k;
        interpolation_coefficients(src, ox ? ox[i] : 0.0, oy ? oy[i] : 0.0, oz ? oz[i] : 0.0, k);
        const int64_t b = fine_cells[i];
        const int nx = dims_f[0], ny = dims_f[1], nz = dims_f[2];
        const int bx = (int)(b % nx), by = (int)((b / nx) % ny), bz = (int)(b / ((int64_t)nx * ny));
        for (int c = 0; c < 8; ++c) {
            const double x = (c & 1) ? 0.25 : -0.25, y = (c & 2) ? 0.25 : -0.25, z = (c & 4) ? 0.25 : -0.25;
            const int64_t kf = box_node(bx + (c & 1), by + ((c >> 1) & 1), bz + ((c >> 2) & 1), nx, ny, nz);
            interpolate_cf(k, node_omega(omega_f, nut_f, kf), 0.5, x, y, z, Ff + 27 * kf);
        }
    }
}

void orc_scale_cf(const double* Fc, double* Ff, const int* dims_c, const int* dims_f, int64_t n_cells,
                  const int64_t* coarse_cells, const int64_t* fine_cells, const double* ox, const double* oy,
                  const double* oz, double omega_c, double omega_f)
{
    orc_scale_cf_turb(Fc, Ff, dims_c, dims_f, n_cells, coarse_cells, fine_cells, ox, oy, oz, omega_c, omega_f,
                      nullptr, nullptr);
}

// Fine-to-coarse (scaleFineToCoarseCompressible.cu:76-124): fine cell with
// corner MMM fine_cells[i] -> coarse node coarse_nodes[i] at the cell centre
// (eps = 2, coarse omega).
void orc_scale_fc_turb(double* Fc, const double* Ff, const int* dims_c, const int* dims_f, int64_t n_cells,
                       const int64_t* coarse_nodes, const int64_t* fine_cells, const double* ox, const double* oy,
                       const double* oz, double omega_c, double omega_f, const double* nut_c, const double* nut_f)
{
    (void)dims_c;
    for (int64_t i = 0; i < n_cells; ++i) {
        SourceMoments src[8];
        cell_moments(Ff, fine_cells[i], dims_f[0], dims_f[1], dims_f[2], omega_f, src, nut_f);
        Coefficients k;
        interpolation_coefficients(src, ox ? ox[i] : 0.0, oy ? oy[i] : 0.0, oz ? oz[i] : 0.0, k);
        interpolate_fc(k, node_omega(omega_c, nut_c, coarse_nodes[i]), 2.0, Fc + 27 * coarse_nodes[i]);
    }
}

void orc_scale_fc(double* Fc, const double* Ff, const int* dims_c, const int* dims_f, int64_t n_cells,
                  const int64_t* coarse_nodes, const int64_t* fine_cells, const double* ox, const double* oy,
                  const double* oz, double omega_c, double omega_f)
{
    orc_scale_fc_turb(Fc, Ff, dims_c, dims_f, n_cells, coarse_nodes, fine_cells, ox, oy, oz, omega_c, omega_f,
                      nullptr, nullptr);
}

} // extern "C"
