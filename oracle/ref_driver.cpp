// ref_driver.cpp -- C-ABI driver over the VirtualFluids reference CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// sources where they lie (REF=/root/reference) into oracle/_ref/libvfref.so.
// It contains no reference code: it instantiates the reference's own
// K17CompressibleNavierStokes kernel, EsoSplit storage, VelocityBounceBack
// strategy and src/lbm collision/chimera/macroscopic functions and drives them
// the way src/cpu/core/Simulation/Simulation.cpp:249-364 does for one block:
//   calculate() -> exchange (periodic ghost fill) -> post-collision BCs -> swap.
//
// The periodic ghost fill stands in for the block connectors
// (src/cpu/core/Connectors/FullDirectConnector.cpp) of a single block that is
// its own neighbour: every EsoSplit slot in the ghost layer receives the value
// the periodic image node wrote.

#include "K17CompressibleNavierStokes.h"
#include "BCArray3D.h"
#include "BCSet.h"
#include "BoundaryConditions.h"
#include "DataSet3D.h"
#include "EsoSplit.h"
#include "VelocityBounceBack.h"

#include <lbm/ChimeraTransformation.h>
#include <lbm/MacroscopicQuantities.h>
#include <lbm/collision/K17CompressibleNavierStokes.h>
#include <lbm/interpolation/InterpolationCoefficients.h>
#include <lbm/refinement/InterpolationCF.h>
#include <lbm/refinement/InterpolationFC.h>

#include <chrono>
#include <cstdint>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// write offsets of the 13 local / 13 non-local EsoSplit slots
// (EsoSplit.cpp:96-129, setPostCollisionDistribution)
const int WO_LOCAL[13][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {1, 0, 0},
                             {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
const int WO_NONLOCAL[13][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 0}, {1, 0, 1}, {0, 0, 1},
                                {0, 1, 1}, {0, 0, 1}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {0, 0, 1}};

int wrap1(int w, int n) { return ((w - 1) % n + n) % n + 1; }

void periodic_fill(EsoSplit& eso, int nx, int ny, int nz)
{
    const int n[3] = {nx, ny, nz};
    for (int which = 0; which < 2; ++which) {
        auto arr = which == 0 ? eso.getLocalDistributions() : eso.getNonLocalDistributions();
        const auto& wo = which == 0 ? WO_LOCAL : WO_NONLOCAL;
        for (int e = 0; e < 13; ++e)
            for (int z = 0; z <= nz + 2; ++z)
                for (int y = 0; y <= ny + 2; ++y)
                    for (int x = 0; x <= nx + 2; ++x) {
                        const int p[3] = {x, y, z};
                        int src[3];
                        bool stale = false;
                        for (int a = 0; a < 3; ++a) {
                            const int w = p[a] - wo[e][a];
                            if (w < 1 || w > n[a]) stale = true;
                            src[a] = wrap1(w, n[a]) + wo[e][a];
                        }
                        if (stale) (*arr)(e, x, y, z) = (*arr)(e, src[0], src[1], src[2]);
                    }
    }
}

struct Block {
    std::shared_ptr<K17CompressibleNavierStokes> kernel;
    std::shared_ptr<EsoSplit> eso;
    std::vector<std::shared_ptr<BCStrategy>> bcs;
};

Block make_block(int nx, int ny, int nz, double omega)
{
    auto proto = std::make_shared<K17CompressibleNavierStokes>();
    proto->setNX({nx, ny, nz});
    proto->setCollisionFactor(omega);
    proto->setBCSet(std::make_shared<BCSet>());
    Block b;
    b.kernel = std::dynamic_pointer_cast<K17CompressibleNavierStokes>(proto->clone());
    b.eso = std::dynamic_pointer_cast<EsoSplit>(b.kernel->getDataSet()->getFdistributions());
    return b;
}

void load(Block& b, const double* F, int nx, int ny, int nz)
{
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
                b.eso->setPreCollisionDistribution(F + 27 * i, x + 1, y + 1, z + 1);
            }
}

void store(Block& b, double* F, int nx, int ny, int nz)
{
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
                b.eso->getPreCollisionDistribution(F + 27 * i, x + 1, y + 1, z + 1);
            }
}

} // namespace

extern "C" {

// vf::lbm::runK17CompressibleNavierStokes<None> on one node.
void ref_k17_collide(double* f27, double omega, const double* lim3, const double* force3, double* macro4)
{
    vf::lbm::CollisionParameter p;
    for (int d = 0; d < 27; ++d) p.distribution[d] = f27[d];
    double lim[3] = {lim3[0], lim3[1], lim3[2]};
    p.omega = omega;
    p.quadricLimiter = lim;
    p.forceX = force3 ? force3[0] : 0.0;
    p.forceY = force3 ? force3[1] : 0.0;
    p.forceZ = force3 ? force3[2] : 0.0;
    vf::lbm::MacroscopicValues mv;
    vf::lbm::TurbulentViscosity tv;
    vf::lbm::runK17CompressibleNavierStokes<vf::lbm::TurbulenceModel::None>(p, mv, tv);
    for (int d = 0; d < 27; ++d) f27[d] = p.distribution[d];
    if (macro4) { macro4[0] = mv.rho; macro4[1] = mv.vx; macro4[2] = mv.vy; macro4[3] = mv.vz; }
}

// the reference K17 with an eddy-viscosity model (turb: 1 Smagorinsky, 2 AMD,
// 3 QR): nut is the node's turbulent viscosity in and the model's value out
void ref_k17_collide_turb(double* f27, double omega, const double* lim3, const double* force3, double* macro4,
                          int turb, double sgs, double* nut)
{
    vf::lbm::CollisionParameter p;
    for (int d = 0; d < 27; ++d) p.distribution[d] = f27[d];
    double lim[3] = {lim3[0], lim3[1], lim3[2]};
    p.omega = omega;
    p.quadricLimiter = lim;
    p.forceX = force3 ? force3[0] : 0.0;
    p.forceY = force3 ? force3[1] : 0.0;
    p.forceZ = force3 ? force3[2] : 0.0;
    vf::lbm::MacroscopicValues mv;
    vf::lbm::TurbulentViscosity tv;
    tv.value = *nut;
    tv.SGSconstant = sgs;
    using TM = vf::lbm::TurbulenceModel;
    switch (turb) {
    case 1: vf::lbm::runK17CompressibleNavierStokes<TM::Smagorinsky>(p, mv, tv); break;
    case 2: vf::lbm::runK17CompressibleNavierStokes<TM::AMD>(p, mv, tv); break;
    default: vf::lbm::runK17CompressibleNavierStokes<TM::QR>(p, mv, tv); break;
    }
    *nut = tv.value;
    for (int d = 0; d < 27; ++d) f27[d] = p.distribution[d];
    if (macro4) { macro4[0] = mv.rho; macro4[1] = mv.vx; macro4[2] = mv.vy; macro4[3] = mv.vz; }
}

// the reference's macroscopic functions of calculateMacroscopicQuantitiesCompressible_device
void ref_macroscopic_quantities(const double* f27, double* out5)
{
    const double rho = vf::lbm::getDensity(f27);
    const double vx = vf::lbm::getCompressibleVelocityX1(f27, rho);
    const double vy = vf::lbm::getCompressibleVelocityX2(f27, rho);
    const double vz = vf::lbm::getCompressibleVelocityX3(f27, rho);
    out5[0] = rho;
    out5[1] = vx;
    out5[2] = vy;
    out5[3] = vz;
    out5[4] = vf::lbm::getPressure(f27, rho, vx, vy, vz);
}

void ref_chimera(int kind, double* abc, double vv, double v2, double kinv, double k)
{
    switch (kind) {
    case 0: vf::lbm::forwardInverseChimeraWithK(abc[0], abc[1], abc[2], vv, v2, kinv, k); break;
    case 1: vf::lbm::backwardInverseChimeraWithK(abc[0], abc[1], abc[2], vv, v2, kinv, k); break;
    case 2: vf::lbm::forwardChimera(abc[0], abc[1], abc[2], vv, v2); break;
    default: vf::lbm::backwardChimera(abc[0], abc[1], abc[2], vv, v2); break;
    }
}

void ref_macroscopic(const double* f27, double* out4)
{
    double drho, vx, vy, vz;
    vf::lbm::getCompressibleMacroscopicValues(f27, drho, vx, vy, vz);
    out4[0] = drho; out4[1] = vx; out4[2] = vy; out4[3] = vz;
}

// Fully periodic nx*ny*nz block: nsteps of K17CompressibleNavierStokes::calculate
// (limiter 0.01 hard-coded there, zero force) on node-major pre-collision F.
int ref_run_periodic(double* F, int nx, int ny, int nz, double omega, int nsteps)
{
    Block b = make_block(nx, ny, nz, omega);
    load(b, F, nx, ny, nz);
    for (int s = 0; s < nsteps; ++s) {
        b.kernel->calculate(s);
        periodic_fill(*b.eso, nx, ny, nz);
        b.kernel->swapDistributions();
    }
    store(b, F, nx, ny, nz);
    return 0;
}

// Closed box of nx*ny*nz fluid nodes surrounded by the block's ghost layer as
// walls.  Every wall link uses the reference VelocityBounceBack strategy
// (post-collision); links into the z = nz+1 plane move with (lux, luy, luz),
// all others have zero velocity.  One strategy object per boundary node and
// velocity group, applied in creation order.
int ref_run_cavity(double* F, int nx, int ny, int nz, double omega, double lux, double luy, double luz, int nsteps)
{
    using namespace vf::lbm::dir;
    Block b = make_block(nx, ny, nz, omega);
    load(b, F, nx, ny, nz);
    auto is_wall = [&](int x, int y, int z) { return x < 1 || x > nx || y < 1 || y > ny || z < 1 || z > nz; };
    for (int z = 1; z <= nz; ++z)
        for (int y = 1; y <= ny; ++y)
            for (int x = 1; x <= nx; ++x) {
                auto lid = std::make_shared<BoundaryConditions>();
                auto wall = std::make_shared<BoundaryConditions>();
                bool has_lid = false, has_wall = false;
                for (int fdir = D3Q27System::FSTARTDIR; fdir <= D3Q27System::FENDDIR; ++fdir) {
                    // flag fdir: the population arriving along fdir comes from the wall at x - c_fdir
                    const int wx = x - D3Q27System::DX1[fdir], wy = y - D3Q27System::DX2[fdir], wz = z - D3Q27System::DX3[fdir];
                    if (!is_wall(wx, wy, wz)) continue;
                    if (wz == nz + 1) { lid->setVelocityBoundaryFlag(fdir); has_lid = true; }
                    else { wall->setVelocityBoundaryFlag(fdir); has_wall = true; }
                }
                lid->setBoundaryVelocityX1((float)lux);
                lid->setBoundaryVelocityX2((float)luy);
                lid->setBoundaryVelocityX3((float)luz);
                for (int g = 0; g < 2; ++g) {
                    if ((g == 0 && !has_lid) || (g == 1 && !has_wall)) continue;
                    auto s = std::make_shared<VelocityBounceBack>();
                    s->addDistributions(b.eso);
                    s->setBcPointer(g == 0 ? lid : wall);
                    s->setNodeIndex(x, y, z);
                    s->setCompressible(true);
                    s->setCollFactor(omega);
                    b.bcs.push_back(s);
                }
            }
    for (int s = 0; s < nsteps; ++s) {
        b.kernel->calculate(s);
        for (auto& bc : b.bcs) bc->applyBC();
        b.kernel->swapDistributions();
    }
    store(b, F, nx, ny, nz);
    return 0;
}

// CPU baseline: nblocks independent periodic-less blocks of bs^3 nodes advanced
// nsteps times with K17CompressibleNavierStokes::calculate + swap, OpenMP
// guided over blocks as Simulation::calculateBlocks (Simulation.cpp:366-401).
// Returns wall seconds of the timed loop (after one warm-up step).
double ref_time_blocks(int nblocks, int bs, double omega, int nsteps, int* threads_used)
{
    std::vector<Block> blocks;
    std::vector<double> F((size_t)bs * bs * bs * 27);
    for (size_t i = 0; i < F.size(); ++i) F[i] = 1e-4 * (double)((i * 2654435761u) % 1000) / 1000.0;
    for (int i = 0; i < nblocks; ++i) {
        blocks.push_back(make_block(bs, bs, bs, omega));
        load(blocks.back(), F.data(), bs, bs, bs);
    }
    auto step = [&](int s) {
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < nblocks; ++i) {
            blocks[i].kernel->calculate(s);
            blocks[i].kernel->swapDistributions();
        }
    };
    step(0);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s <= nsteps; ++s) step(s);
    auto t1 = std::chrono::steady_clock::now();
#ifdef _OPENMP
    if (threads_used) *threads_used = omp_get_max_threads();
#else
    if (threads_used) *threads_used = 1;
#endif
    return std::chrono::duration<double>(t1 - t0).count();
}

} // extern "C"

// Refinement interpolation of one cell with the reference's header-only
// src/lbm/interpolation/InterpolationCoefficients.h and
// src/lbm/refinement/InterpolationCF.h / InterpolationFC.h.
// f8x27: the 8 source nodes, corner c with bit 0 = +x, bit 1 = +y, bit 2 = +z
// (c = 0 is MMM).  kind 0: coarse-to-fine, out = 8 x 27 (destination nodes at
// (+-1/4, +-1/4, +-1/4), same corner order, eps = 1/2); kind 1: fine-to-coarse,
// out = 27 (eps = 2).  coeff42 receives a, b, c (000,100,010,001,200,020,002,
// 110,101,011,111), d (000,100,010,001,110,101,011,111), LaplaceRho.
extern "C" void ref_interp_cell(int kind, const double* f8x27, double omega_src, double omega_dst, double xoff,
                                double yoff, double zoff, double* coeff42, double* out)
{
    vf::lbm::MomentsOnSourceNodeSet set;
    const double* f = f8x27;
    set.calculateMMM(f + 27 * 0, omega_src);
    set.calculatePMM(f + 27 * 1, omega_src);
    set.calculateMPM(f + 27 * 2, omega_src);
    set.calculatePPM(f + 27 * 3, omega_src);
    set.calculateMMP(f + 27 * 4, omega_src);
    set.calculatePMP(f + 27 * 5, omega_src);
    set.calculateMPP(f + 27 * 6, omega_src);
    set.calculatePPP(f + 27 * 7, omega_src);
    vf::lbm::InterpolationCoefficients k;
    set.calculateCoefficients(k, xoff, yoff, zoff);
    const double packed[42] = {k.a000, k.a100, k.a010, k.a001, k.a200, k.a020, k.a002, k.a110, k.a101, k.a011, k.a111,
                               k.b000, k.b100, k.b010, k.b001, k.b200, k.b020, k.b002, k.b110, k.b101, k.b011, k.b111,
                               k.c000, k.c100, k.c010, k.c001, k.c200, k.c020, k.c002, k.c110, k.c101, k.c011, k.c111,
                               k.d000, k.d100, k.d010, k.d001, k.d110, k.d101, k.d011, k.d111, k.LaplaceRho};
    for (int i = 0; i < 42; ++i) coeff42[i] = packed[i];
    if (kind == 0) {
        for (int c = 0; c < 8; ++c) {
            const double x = (c & 1) ? 0.25 : -0.25, y = (c & 2) ? 0.25 : -0.25, z = (c & 4) ? 0.25 : -0.25;
            vf::lbm::interpolateCF(out + 27 * c, omega_dst, 0.5, k, x, y, z);
        }
    } else {
        vf::lbm::interpolateFC(out, 2.0, omega_dst, k);
    }
}

