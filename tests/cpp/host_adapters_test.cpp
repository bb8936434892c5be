// Drives the C++ host layer (KernelFactory / Kernel / BoundaryConditionFactory /
// UpdateGrid27) the way VirtualFluids' Simulation does.
//   host_adapters_test cpu : factory and error behaviour only (no device calls)
//   host_adapters_test gpu : a periodic K17 lattice stepped through the host
//                            layer equals the same lattice stepped through the
//                            C ABI directly, bit for bit
#include "../../virtualfluids_amd/csrc/host/vfx_host.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace vfx::host;

static int failures = 0;
#define EXPECT(c)                                                                                                      \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            std::printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #c);                                                  \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

static void cpu_tests()
{
    auto para = std::make_shared<Parameter>(1);
    KernelFactoryImp factory;
    para->setMainKernel("K17CompressibleNavierStokes");
    auto ks = factory.makeKernels(para);
    EXPECT(ks.size() == 2);
    EXPECT(dynamic_cast<K17CompressibleNavierStokes*>(ks[0].get()) != nullptr);
    EXPECT(dynamic_cast<KernelImp*>(ks[0].get())->getKernelUsesFluidNodeIndices());
    para->setMainKernel("BGKCompSP27");
    ks = factory.makeKernels(para);
    EXPECT(dynamic_cast<B92CompressibleNavierStokes*>(ks[1].get()) != nullptr);
    bool threw = false;
    try {
        factory.makeKernel(para, "K15CompressibleNavierStokes", 0);
    } catch (const std::runtime_error& e) {
        threw = std::strcmp(e.what(), "KernelFactory does not know the KernelType.") == 0;
    }
    EXPECT(threw);
    BoundaryConditionFactory bcf;
    EXPECT(!bcf.getVelocityBoundaryConditionPost());
    bcf.setVelocityBoundaryCondition(BoundaryConditionFactory::VelocityBC::VelocityInterpolatedCompressible);
    EXPECT((bool)bcf.getVelocityBoundaryConditionPost());
    // boundary nodes without a selected boundary condition are rejected
    auto p1 = std::make_shared<Parameter>(0);
    p1->getParD(0)->noSlipBC.numberOfBCnodes = 5;
    threw = false;
    try {
        UpdateGrid27 u(p1, factory.makeKernels(p1), BoundaryConditionFactory());
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    // the C ABI rejects bad arguments without touching a device
    vfx_lattice lat{};
    EXPECT(vfx_collide_stream(&lat, VFX_K17, 1, 0, nullptr) == VFX_EINVAL);
    EXPECT(vfx_collide_stream(nullptr, VFX_K17, 1, 0, nullptr) == VFX_EINVAL);
}

static void gpu_tests()
{
    const uint32_t nx = 10, ny = 8, nz = 6, N = nx * ny * nz + 1; // node 0: dummy
    std::vector<uint32_t> nbx(N, 0), nby(N, 0), nbz(N, 0), fluid;
    auto id = [&](uint32_t x, uint32_t y, uint32_t z) { return 1 + (x % nx) + nx * ((y % ny) + ny * (z % nz)); };
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const uint32_t k = id(x, y, z);
                nbx[k] = id(x + 1, y, z);
                nby[k] = id(x, y + 1, z);
                nbz[k] = id(x, y, z + 1);
                fluid.push_back(k);
            }
    std::vector<double> F((size_t)N * 27, 0.0), drho(N), ux(N), uy(N), uz(N);
    for (uint32_t k = 0; k < N; ++k) {
        drho[k] = 1e-3 * std::sin(0.3 * k);
        ux[k] = 0.03 * std::sin(0.11 * k);
        uy[k] = 0.02 * std::cos(0.07 * k);
        uz[k] = 0.01 * std::sin(0.05 * k + 1.0);
    }
    auto up = [&](const void* src, size_t bytes) {
        void* d = nullptr;
        check(vfx_device_alloc(bytes, &d), "alloc");
        check(vfx_copy_to_device(d, src, bytes), "h2d");
        return d;
    };
    void* dnbx = up(nbx.data(), N * 4);
    void* dnby = up(nby.data(), N * 4);
    void* dnbz = up(nbz.data(), N * 4);
    void* dfl = up(fluid.data(), fluid.size() * 4);
    void* fields[4] = {up(drho.data(), N * 8), up(ux.data(), N * 8), up(uy.data(), N * 8), up(uz.data(), N * 8)};
    void* fa = nullptr;
    void* fb = nullptr;
    check(vfx_device_alloc((uint64_t)N * 27 * 8, &fa), "alloc");
    check(vfx_device_alloc((uint64_t)N * 27 * 8, &fb), "alloc");

    vfx_lattice lat{};
    lat.storage = VFX_FP64;
    lat.n_nodes = N;
    lat.direct = 0;
    lat.neighbor_x = (const uint32_t*)dnbx;
    lat.neighbor_y = (const uint32_t*)dnby;
    lat.neighbor_z = (const uint32_t*)dnbz;
    lat.fluid_nodes = (const uint32_t*)dfl;
    lat.n_fluid = (uint32_t)fluid.size();
    lat.omega = 1.85;
    lat.quadric_limiter[0] = lat.quadric_limiter[1] = lat.quadric_limiter[2] = 0.01;

    // (a) through the host layer
    auto para = std::make_shared<Parameter>(0);
    LBMSimulationParameter* p = para->getParD(0);
    p->lattice = lat;
    p->lattice.f = fa;
    p->taggedFluidNodeIndices[CollisionTemplate::Default] = {(uint32_t*)dfl, (uint32_t)fluid.size()};
    check(vfx_init_equilibrium(&p->lattice, 1, (double*)fields[0], (double*)fields[1], (double*)fields[2],
                               (double*)fields[3], nullptr),
          "init");
    KernelFactoryImp factory;
    UpdateGrid27 ug(para, factory.makeKernels(para), BoundaryConditionFactory());
    for (unsigned t = 0; t < 20; ++t) ug.updateGrid(0, t);

    // (b) through the C ABI directly
    vfx_lattice lb = lat;
    lb.f = fb;
    check(vfx_init_equilibrium(&lb, 1, (double*)fields[0], (double*)fields[1], (double*)fields[2], (double*)fields[3],
                               nullptr),
          "init");
    bool even = true;
    for (int t = 0; t < 20; ++t) {
        check(vfx_k17_step(&lb, even ? 1 : 0, VFX_TEMPLATE_DEFAULT, nullptr), "step");
        even = !even;
    }
    check(vfx_synchronize(nullptr), "sync");
    std::vector<double> a((size_t)N * 27), b((size_t)N * 27);
    check(vfx_copy_to_host(a.data(), fa, a.size() * 8), "d2h");
    check(vfx_copy_to_host(b.data(), fb, b.size() * 8), "d2h");
    EXPECT(p->isEvenTimestep == even);
    EXPECT(std::memcmp(a.data(), b.data(), a.size() * 8) == 0);
    double amax = 0;
    for (double v : a) amax = std::fmax(amax, std::fabs(v));
    EXPECT(amax > 1e-4 && amax < 1.0);
    for (void* d : {dnbx, dnby, dnbz, dfl, fa, fb, fields[0], fields[1], fields[2], fields[3]}) vfx_device_free(d);
}

int main(int argc, char** argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    try {
        if (gpu) gpu_tests();
        else cpu_tests();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s host adapter tests: %d failure(s)\n", gpu ? "gpu" : "cpu", failures);
    return failures ? 1 : 0;
}
