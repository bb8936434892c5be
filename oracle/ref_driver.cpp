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

#include "B92IncompressibleNavierStokes.h"
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
#include <cmath>
#include <cstdlib>
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
    std::shared_ptr<LBMKernel> kernel;
    std::shared_ptr<EsoSplit> eso;
    std::vector<std::shared_ptr<BCStrategy>> bcs;
};

// a block of the reference's K17CompressibleNavierStokes (default) or, with
// incompressible_bgk, of its B92IncompressibleNavierStokes CPU kernel
Block make_block(int nx, int ny, int nz, double omega, bool incompressible_bgk = false)
{
    std::shared_ptr<LBMKernel> proto;
    if (incompressible_bgk) proto = std::make_shared<B92IncompressibleNavierStokes>();
    else proto = std::make_shared<K17CompressibleNavierStokes>();
    proto->setNX({nx, ny, nz});
    proto->setCollisionFactor(omega);
    proto->setBCSet(std::make_shared<BCSet>());
    Block b;
    b.kernel = proto->clone();
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

// The same with the reference's incompressible BGK CPU kernel
// (B92IncompressibleNavierStokes::calculate, B92IncompressibleNavierStokes.cpp:74-338).
int ref_run_periodic_b92(double* F, int nx, int ny, int nz, double omega, int nsteps)
{
    Block b = make_block(nx, ny, nz, omega, true);
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

} // extern "C"

namespace {

// A periodic bx*by*bz grid of bs^3 blocks advanced the way Simulation::run
// does for one level (Simulation.cpp:249-364): K17CompressibleNavierStokes::calculate
// of every block (OpenMP guided, Simulation::calculateBlocks :366-401), the
// block exchange, then swap.  The exchange gives every ghost slot of a block
// the value its neighbouring block wrote there -- the copy the local
// FullDirectConnectors make (OneDistributionFullDirectConnector.h:86-146),
// OpenMP over blocks as connectorsSendLocal (:452-467); reads touch only
// non-ghost slots, so blocks exchange in parallel without races.
struct BlockGrid {
    int B[3];
    int n;
    std::vector<Block> blocks;

    BlockGrid(int bx, int by, int bz, int bs, double omega) : B{bx, by, bz}, n(bs)
    {
        for (int i = 0; i < bx * by * bz; ++i) blocks.push_back(make_block(bs, bs, bs, omega));
    }
    int block_of(const int (&c)[3]) const
    {
        int w[3];
        for (int a = 0; a < 3; ++a) w[a] = ((c[a] % B[a]) + B[a]) % B[a];
        return w[0] + B[0] * (w[1] + B[1] * w[2]);
    }
    void exchange_block(int ib)
    {
        const int bc[3] = {ib % B[0], (ib / B[0]) % B[1], ib / (B[0] * B[1])};
        const int xs[4] = {0, 1, n + 1, n + 2};
        // raw array pointers of the 27 neighbouring blocks (shift -1..1 per axis),
        // fetched once: the shared_ptr getters would contend on reference counts
        using A4 = decltype(blocks[0].eso->getLocalDistributions().get());
        using A3 = decltype(blocks[0].eso->getZeroDistributions().get());
        A4 nbr4[2][27];
        A3 nbr3[27];
        for (int c = 0; c < 27; ++c) {
            const int nbc[3] = {bc[0] + c % 3 - 1, bc[1] + (c / 3) % 3 - 1, bc[2] + c / 9 - 1};
            const Block& nb = blocks[block_of(nbc)];
            nbr4[0][c] = nb.eso->getLocalDistributions().get();
            nbr4[1][c] = nb.eso->getNonLocalDistributions().get();
            nbr3[c] = nb.eso->getZeroDistributions().get();
        }
        for (int which = 0; which < 2; ++which) {
            A4 to = nbr4[which][13];
            const auto& wo = which == 0 ? WO_LOCAL : WO_NONLOCAL;
            for (int e = 0; e < 13; ++e) {
                // per axis: the slots p with p - wo in 1..n are the block's own;
                // every other slot (within two layers of a face) is a ghost that
                // takes the neighbouring block's value at wrap(p - wo) + wo
                std::vector<int> shv(3 * (n + 3)), srcv(3 * (n + 3));
                int* sh[3] = {shv.data(), shv.data() + (n + 3), shv.data() + 2 * (n + 3)};
                int* src[3] = {srcv.data(), srcv.data() + (n + 3), srcv.data() + 2 * (n + 3)};
                for (int a = 0; a < 3; ++a)
                    for (int p = 0; p <= n + 2; ++p) {
                        const int w = p - wo[e][a];
                        sh[a][p] = w < 1 ? 0 : w > n ? 2 : 1;
                        src[a][p] = wrap1(w, n) + wo[e][a];
                    }
                const int ox = wo[e][0];
                for (int z = 0; z <= n + 2; ++z)
                    for (int y = 0; y <= n + 2; ++y) {
                        const int cyz = 9 * sh[2][z] + 3 * sh[1][y];
                        const bool inner = sh[2][z] == 1 && sh[1][y] == 1;
                        // inner rows: only x in 0..ox and n+ox+1..n+2 are ghosts
                        for (int x = 0; x <= n + 2; x = (inner && x == ox) ? n + ox + 1 : x + 1) {
                            const int c = cyz + sh[0][x];
                            (*to)(e, x, y, z) = (*nbr4[which][c])(e, src[0][x], src[1][y], src[2][z]);
                        }
                    }
            }
        }
        A3 zto = nbr3[13];
        for (int z = 0; z <= n + 1; ++z)
            for (int y = 0; y <= n + 1; ++y)
                for (int x = 0; x <= n + 1; ++x) {
                    if (x >= 1 && x <= n && y >= 1 && y <= n && z >= 1 && z <= n) continue;
                    const int p[3] = {x, y, z};
                    int c = 0, src[3];
                    for (int a = 2; a >= 0; --a) {
                        c = 3 * c + (p[a] < 1 ? 0 : p[a] > n ? 2 : 1);
                        src[a] = wrap1(p[a], n);
                    }
                    (*zto)(x, y, z) = (*nbr3[c])(src[0], src[1], src[2]);
                }
    }
    void step(int s, bool exchange = true)
    {
        const int nb = (int)blocks.size();
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < nb; ++i) blocks[i].kernel->calculate(s);
        if (exchange) {
#pragma omp parallel for schedule(guided)
            for (int i = 0; i < nb; ++i) exchange_block(i);
        }
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < nb; ++i) blocks[i].kernel->swapDistributions();
    }
    // node-major field of the whole (bx*bs, by*bs, bz*bs) domain, x fastest
    void transfer(double* F, bool to_blocks)
    {
        const int NX = B[0] * n, NY = B[1] * n;
        std::vector<double> tmp((size_t)n * n * n * 27);
        for (int ib = 0; ib < (int)blocks.size(); ++ib) {
            const int bc[3] = {ib % B[0], (ib / B[0]) % B[1], ib / (B[0] * B[1])};
            if (!to_blocks) store(blocks[ib], tmp.data(), n, n, n);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) {
                        const int64_t g = (bc[0] * n + x) + (int64_t)NX * ((bc[1] * n + y) + (int64_t)NY * (bc[2] * n + z));
                        const int64_t l = x + (int64_t)n * (y + (int64_t)n * z);
                        for (int d = 0; d < 27; ++d) {
                            if (to_blocks) tmp[27 * l + d] = F[27 * g + d];
                            else F[27 * g + d] = tmp[27 * l + d];
                        }
                    }
            if (to_blocks) load(blocks[ib], tmp.data(), n, n, n);
        }
    }
};

} // namespace

extern "C" {

// CPU baseline: nblocks blocks of bs^3 nodes (split into an as-cubic-as-possible
// periodic block grid) advanced nsteps times by BlockGrid::step (exchange = 0
// leaves the block exchange out).  Returns wall seconds of the timed loop
// (after one warm-up step) and the grid in dims3.
double ref_time_blocks(int nblocks, int bs, double omega, int nsteps, int exchange, int* threads_used, int* dims3)
{
    int B[3] = {1, 1, 1};
    int rest = nblocks;
    for (int a = 0; a < 2; ++a) {
        const double want = std::cbrt((double)rest);
        int best = 1;
        for (int f = 1; f <= rest; ++f)
            if (rest % f == 0 && std::fabs(f - want) < std::fabs(best - want)) best = f;
        B[a] = best;
        rest /= best;
    }
    B[2] = rest;
    BlockGrid g(B[0], B[1], B[2], bs, omega);
    std::vector<double> F((size_t)nblocks * bs * bs * bs * 27);
    for (size_t i = 0; i < F.size(); ++i) F[i] = 1e-4 * (double)((i * 2654435761u) % 1000) / 1000.0;
    g.transfer(F.data(), true);
    g.step(0, exchange != 0);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 1; s <= nsteps; ++s) g.step(s, exchange != 0);
    auto t1 = std::chrono::steady_clock::now();
#ifdef _OPENMP
    if (threads_used) *threads_used = omp_get_max_threads();
#else
    if (threads_used) *threads_used = 1;
#endif
    if (dims3)
        for (int a = 0; a < 3; ++a) dims3[a] = B[a];
    return std::chrono::duration<double>(t1 - t0).count();
}

// The baseline's block grid on one node-major periodic field F of
// (bx*bs, by*bs, bz*bs) nodes: nsteps of BlockGrid::step, F receives the
// pre-collision state.  Equal to ref_run_periodic on the same field when the
// exchange is the streaming a periodic domain sees.
int ref_run_blocks(double* F, int bx, int by, int bz, int bs, double omega, int nsteps)
{
    BlockGrid g(bx, by, bz, bs, omega);
    g.transfer(F, true);
    for (int s = 0; s < nsteps; ++s) g.step(s);
    g.transfer(F, false);
    return 0;
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

