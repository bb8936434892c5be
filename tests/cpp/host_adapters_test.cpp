// Drives the C++ host layer (KernelFactory / Kernel / BoundaryConditionFactory /
// UpdateGrid27) the way VirtualFluids' Simulation does.  Built with
// VF_DOUBLE_ACCURACY (real = double): the lattices here are FP64.
//   host_adapters_test cpu : factory and error behaviour only (no device calls)
//   host_adapters_test gpu : a periodic K17 lattice stepped through the host
//                            layer equals the same lattice stepped through the
//                            C ABI directly, bit for bit
#include "../../virtualfluids_amd/csrc/host/vfx_host.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <map>
#include <string>
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
    // more than one level needs a scaling factory (GridScalingKernelManager.cpp:45-50)
    auto p2 = std::make_shared<Parameter>(1);
    threw = false;
    try {
        UpdateGrid27 u(p2, factory.makeKernels(p2), BoundaryConditionFactory());
    } catch (const std::runtime_error& e) {
        threw = std::strcmp(e.what(), "There is more than one level, but no scalingFactory was provided.") == 0;
    }
    EXPECT(threw);
    GridScalingFactory gsf;
    EXPECT(!gsf.getGridScalingFC(false) && !gsf.getGridScalingCF(false));
    gsf.setScalingFactory(GridScalingFactory::GridScaling::ScaleCompressible);
    EXPECT((bool)gsf.getGridScalingFC(false) && (bool)gsf.getGridScalingCF(false));
    // with a turbulent viscosity the same scaling runs with per-node rates
    EXPECT((bool)gsf.getGridScalingFC(true) && (bool)gsf.getGridScalingCF(true));
    // TurbulenceModelFactory::readConfigFile keys
    auto pt = std::make_shared<Parameter>(0);
    TurbulenceModelFactory tmf(pt);
    EXPECT(pt->getTurbulenceModel() == TurbulenceModel::None && !pt->getUseTurbulentViscosity());
    tmf.readConfig({{"TurbulenceModel", "QR"}, {"SGSconstant", "0.3"}});
    EXPECT(pt->getTurbulenceModel() == TurbulenceModel::QR && pt->getUseTurbulentViscosity());
    EXPECT(pt->getSGSConstant() == 0.3);
    EXPECT(vfx_turbulence_model_of(TurbulenceModel::AMD) == VFX_TURB_AMD);
    threw = false;
    try {
        tmf.readConfig({{"TurbulenceModel", "WALE"}});
    } catch (const std::runtime_error& e) {
        threw = std::strcmp(e.what(), "TurbulenceModelFactory: Invalid turbulence model!") == 0;
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

    // (c) restart: serialize the level (double, as with VF_DOUBLE_ACCURACY),
    // clear the array, deserialize -- the distributions come back bit for bit
    const std::string rs = "host_adapters_restart";
    BinaryRestartObjectT<double> out;
    out.serialize(rs, para);
    EXPECT(out.fs.size() == 1 && out.fs[0].size() == a.size());
    EXPECT(out.fs.size() == 1 && std::memcmp(out.fs[0].data(), a.data(), a.size() * 8) == 0);
    std::vector<double> zeros(a.size(), 0.0);
    check(vfx_copy_to_device(fa, zeros.data(), zeros.size() * 8), "h2d");
    BinaryRestartObjectT<double> in;
    in.deserialize(rs, para);
    std::vector<double> c(a.size());
    check(vfx_copy_to_host(c.data(), fa, c.size() * 8), "d2h");
    EXPECT(std::memcmp(a.data(), c.data(), a.size() * 8) == 0);
    EXPECT(in.delete_restart_file(rs));
    for (void* d : {dnbx, dnby, dnbz, dfl, fa, fb, fields[0], fields[1], fields[2], fields[3]}) vfx_device_free(d);
}

// Two levels through UpdateGrid27 (fine twice, coarse, fine-to-coarse,
// coarse-to-fine) equal the same launches made through the C ABI directly.
static void gpu_refinement_tests()
{
    const uint32_t n = 8, N = n * n * n;
    std::vector<double> drho(N), ux(N), uy(N), uz(N);
    for (uint32_t k = 0; k < N; ++k) {
        drho[k] = 1e-3 * std::cos(0.2 * k);
        ux[k] = 0.02 * std::sin(0.13 * k);
        uy[k] = 0.015 * std::cos(0.05 * k);
        uz[k] = 0.01 * std::sin(0.09 * k + 0.5);
    }
    auto up = [&](const void* src, size_t bytes) {
        void* d = nullptr;
        check(vfx_device_alloc(bytes, &d), "alloc");
        check(vfx_copy_to_device(d, src, bytes), "h2d");
        return d;
    };
    void* fields[4] = {up(drho.data(), N * 8), up(ux.data(), N * 8), up(uy.data(), N * 8), up(uz.data(), N * 8)};
    // disjoint cells: coarse-to-fine writes the 2x2x2 fine blocks at even
    // coordinates, fine-to-coarse writes distinct coarse nodes
    std::vector<uint32_t> cf_c, cf_f, fc_c, fc_f;
    for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t bx = 2 * (i % 4), by = 2 * ((i / 4) % 4), bz = 2 * (i % 3);
        cf_f.push_back(bx + n * (by + n * bz));
        cf_c.push_back((7 * i + 3) % N);
        fc_c.push_back((11 * i + 5) % N);
        fc_f.push_back((13 * i + 1) % N);
    }
    void* d_cf_c = up(cf_c.data(), 64);
    void* d_cf_f = up(cf_f.data(), 64);
    void* d_fc_c = up(fc_c.data(), 64);
    void* d_fc_f = up(fc_f.data(), 64);
    vfx_lattice lat{};
    lat.storage = VFX_FP64;
    lat.n_nodes = N;
    lat.direct = 1;
    lat.dims[0] = lat.dims[1] = lat.dims[2] = n;
    lat.quadric_limiter[0] = lat.quadric_limiter[1] = lat.quadric_limiter[2] = 0.01;
    void* f[4];
    for (auto& a : f) check(vfx_device_alloc((uint64_t)N * 27 * 8, &a), "alloc");
    auto level = [&](void* fp, double omega) {
        vfx_lattice l = lat;
        l.f = fp;
        l.omega = omega;
        check(vfx_init_equilibrium(&l, 1, (double*)fields[0], (double*)fields[1], (double*)fields[2],
                                   (double*)fields[3], nullptr),
              "init");
        return l;
    };
    const double oc = 1.9, of = 1.0 / (6.0 * (1.0 / oc - 0.5) / 3.0 + 0.5);

    // (a) host layer
    auto para = std::make_shared<Parameter>(1);
    LBMSimulationParameter* c = para->getParD(0);
    LBMSimulationParameter* fi = para->getParD(1);
    c->lattice = level(f[0], oc);
    fi->lattice = level(f[1], of);
    c->coarseToFine = {(uint32_t*)d_cf_f, (uint32_t*)d_cf_c, 16};
    c->fineToCoarse = {(uint32_t*)d_fc_f, (uint32_t*)d_fc_c, 16};
    GridScalingFactory gsf;
    gsf.setScalingFactory(GridScalingFactory::GridScaling::ScaleCompressible);
    KernelFactoryImp factory;
    UpdateGrid27 ug(para, factory.makeKernels(para), BoundaryConditionFactory(), &gsf);
    for (unsigned t = 0; t < 6; ++t) ug.updateGrid(0, t);

    // (b) C ABI
    vfx_lattice lc = level(f[2], oc), lf = level(f[3], of);
    vfx_icells cf{(const uint32_t*)d_cf_f, (const uint32_t*)d_cf_c, 16, nullptr, nullptr, nullptr};
    vfx_icells fc{(const uint32_t*)d_fc_f, (const uint32_t*)d_fc_c, 16, nullptr, nullptr, nullptr};
    bool ec = true, ef = true;
    for (int t = 0; t < 6; ++t) {
        for (int s = 0; s < 2; ++s) {
            check(vfx_collide_stream(&lf, VFX_K17, ef ? 1 : 0, VFX_TEMPLATE_DEFAULT, nullptr), "fine");
            ef = !ef;
        }
        check(vfx_collide_stream(&lc, VFX_K17, ec ? 1 : 0, VFX_TEMPLATE_DEFAULT, nullptr), "coarse");
        ec = !ec;
        check(vfx_scale_fc(&lc, &lf, &fc, ec ? 1 : 0, 1, nullptr), "fc");
        check(vfx_scale_cf(&lc, &lf, &cf, ec ? 1 : 0, 1, nullptr), "cf");
    }
    check(vfx_synchronize(nullptr), "sync");
    std::vector<double> a((size_t)N * 27), b((size_t)N * 27);
    for (int lvl = 0; lvl < 2; ++lvl) {
        check(vfx_copy_to_host(a.data(), f[lvl], a.size() * 8), "d2h");
        check(vfx_copy_to_host(b.data(), f[lvl + 2], b.size() * 8), "d2h");
        EXPECT(std::memcmp(a.data(), b.data(), a.size() * 8) == 0);
    }
    EXPECT(c->isEvenTimestep == ec && fi->isEvenTimestep == ef);
    for (void* d : {d_cf_c, d_cf_f, d_fc_c, d_fc_f, fields[0], fields[1], fields[2], fields[3], f[0], f[1], f[2], f[3]})
        vfx_device_free(d);
}

// The AMD model through UpdateGrid27 + TurbulenceModelFactory (collision with
// the turbulent viscosity, swap, calcTurbulentViscosityAMD) equals the same
// launches through the C ABI directly, populations and viscosity bit for bit.
static void gpu_turbulence_tests()
{
    const uint32_t nx = 10, ny = 8, nz = 6, N = nx * ny * nz;
    std::vector<double> drho(N), ux(N), uy(N), uz(N), zero(N, 0.0);
    for (uint32_t k = 0; k < N; ++k) {
        drho[k] = 1e-3 * std::sin(0.21 * k);
        ux[k] = 0.04 * std::sin(0.17 * k);
        uy[k] = 0.03 * std::cos(0.09 * k);
        uz[k] = 0.02 * std::sin(0.07 * k + 0.3);
    }
    auto up = [&](const void* src, size_t bytes) {
        void* d = nullptr;
        check(vfx_device_alloc(bytes, &d), "alloc");
        check(vfx_copy_to_device(d, src, bytes), "h2d");
        return d;
    };
    void* fields[4] = {up(drho.data(), N * 8), up(ux.data(), N * 8), up(uy.data(), N * 8), up(uz.data(), N * 8)};
    // per run: populations, turbulent viscosity, rho/vx/vy/vz
    void* arr[2][6];
    for (auto& run : arr) {
        check(vfx_device_alloc((uint64_t)N * 27 * 8, &run[0]), "alloc");
        for (int i = 1; i < 6; ++i) run[i] = up(zero.data(), N * 8);
    }
    auto lattice = [&](void* const* run) {
        vfx_lattice l{};
        l.storage = VFX_FP64;
        l.n_nodes = N;
        l.direct = 1;
        l.dims[0] = nx;
        l.dims[1] = ny;
        l.dims[2] = nz;
        l.omega = 1.95;
        l.quadric_limiter[0] = l.quadric_limiter[1] = l.quadric_limiter[2] = 0.01;
        l.f = run[0];
        l.turbulent_viscosity = run[1];
        l.rho = run[2];
        l.vx = run[3];
        l.vy = run[4];
        l.vz = run[5];
        check(vfx_init_equilibrium(&l, 1, (double*)fields[0], (double*)fields[1], (double*)fields[2],
                                   (double*)fields[3], nullptr),
              "init");
        return l;
    };
    // (a) host layer
    auto para = std::make_shared<Parameter>(0);
    para->getParD(0)->lattice = lattice(arr[0]);
    auto tm = std::make_shared<TurbulenceModelFactory>(para);
    tm->setTurbulenceModel(TurbulenceModel::AMD);
    tm->setModelConstant(0.33);
    KernelFactoryImp factory;
    UpdateGrid27 ug(para, factory.makeKernels(para), BoundaryConditionFactory(), nullptr, tm);
    for (unsigned t = 0; t < 10; ++t) ug.updateGrid(0, t);
    // (b) C ABI
    vfx_lattice lb = lattice(arr[1]);
    lb.turbulence_model = VFX_TURB_AMD;
    lb.sgs_constant = 0.33;
    bool even = true;
    for (int t = 0; t < 10; ++t) {
        check(vfx_k17_step(&lb, even ? 1 : 0, VFX_TEMPLATE_DEFAULT, nullptr), "step");
        even = !even;
        check(vfx_turbulent_viscosity_amd(&lb, nullptr), "amd");
    }
    check(vfx_synchronize(nullptr), "sync");
    std::vector<double> a((size_t)N * 27), b((size_t)N * 27), na(N), nb(N);
    check(vfx_copy_to_host(a.data(), arr[0][0], a.size() * 8), "d2h");
    check(vfx_copy_to_host(b.data(), arr[1][0], b.size() * 8), "d2h");
    check(vfx_copy_to_host(na.data(), arr[0][1], N * 8), "d2h");
    check(vfx_copy_to_host(nb.data(), arr[1][1], N * 8), "d2h");
    EXPECT(std::memcmp(a.data(), b.data(), a.size() * 8) == 0);
    EXPECT(std::memcmp(na.data(), nb.data(), N * 8) == 0);
    double nmax = 0;
    for (double v : na) nmax = std::fmax(nmax, v);
    EXPECT(nmax > 0.0);
    for (auto& run : arr)
        for (void* d : run) vfx_device_free(d);
    for (void* d : fields) vfx_device_free(d);
}

// Restart files: write the reference unit test's data (RestartObjectTests.cpp,
// saveAndLoad) plus values that exercise the text format, and read back the
// files the Python writer made for the same data (<prefix>_py.bin/.txt).
static void restart_tests(const std::string& prefix)
{
    const std::vector<std::vector<float>> data = {{1, 2, 3}, {4, 5, 6}, {0.123456789f, -1e-7f, 12345678.f, -0.f}};
    BinaryRestartObjectT<float> b;
    b.fs = data;
    b.serialize_internal(prefix + "_cpp");
    ASCIIRestartObjectT<float> a;
    a.fs = data;
    a.serialize_internal(prefix + "_cpp");
    BinaryRestartObjectT<double> d;
    d.fs = {{1, 2, 3}, {0.1, -2.5e-300}};
    d.serialize_internal(prefix + "_cpp64");
    BinaryRestartObjectT<float> rb;
    rb.deserialize_internal(prefix + "_py");
    EXPECT(rb.fs == data);
    ASCIIRestartObjectT<float> ra;
    ra.deserialize_internal(prefix + "_py");
    EXPECT(ra.fs.size() == data.size());
    for (size_t i = 0; i < ra.fs.size() && i < data.size(); ++i) {
        EXPECT(ra.fs[i].size() == data[i].size());
        for (size_t j = 0; j < ra.fs[i].size() && j < data[i].size(); ++j)
            EXPECT(std::fabs(ra.fs[i][j] - data[i][j]) <= 5e-6f * std::fabs(data[i][j])); // six digits
    }
    BinaryRestartObjectT<double> rd;
    rd.deserialize_internal(prefix + "_py64");
    EXPECT(rd.fs == d.fs);
    b.serialize_internal(prefix + "_del");
    EXPECT(b.delete_restart_file(prefix + "_del"));
    EXPECT(!b.delete_restart_file(prefix + "_does_not_exist"));
}

// FileWriter through the host layer: the written point data of an output step
// equal vfx_macroscopic_quantities of the level (press and rho / 3).
static void gpu_output_tests()
{
    const uint32_t nx = 6, ny = 5, nz = 4, N = nx * ny * nz;
    std::vector<double> drho(N), ux(N), uy(N), uz(N);
    for (uint32_t k = 0; k < N; ++k) {
        drho[k] = 1e-3 * std::sin(0.4 * k);
        ux[k] = 0.03 * std::cos(0.2 * k);
        uy[k] = 0.02 * std::sin(0.3 * k);
        uz[k] = 0.01 * std::cos(0.1 * k + 0.2);
    }
    auto up = [&](const void* src, size_t bytes) {
        void* d = nullptr;
        check(vfx_device_alloc(bytes, &d), "alloc");
        check(vfx_copy_to_device(d, src, bytes), "h2d");
        return d;
    };
    void* fields[4] = {up(drho.data(), N * 8), up(ux.data(), N * 8), up(uy.data(), N * 8), up(uz.data(), N * 8)};
    void* f = nullptr;
    check(vfx_device_alloc((uint64_t)N * 27 * 8, &f), "alloc");
    auto para = std::make_shared<Parameter>(0);
    LBMSimulationParameter* p = para->getParD(0);
    vfx_lattice& lat = p->lattice;
    lat.f = f;
    lat.storage = VFX_FP64;
    lat.n_nodes = N;
    lat.direct = 1;
    lat.dims[0] = nx;
    lat.dims[1] = ny;
    lat.dims[2] = nz;
    lat.omega = 1.9;
    check(vfx_init_equilibrium(&lat, 1, (double*)fields[0], (double*)fields[1], (double*)fields[2], (double*)fields[3],
                               nullptr),
          "init");
    HostGrid& g = p->host;
    for (uint32_t k = 0; k < N; ++k) {
        const uint32_t x = k % nx, y = (k / nx) % ny, z = k / (nx * ny);
        g.coordinateX.push_back(x);
        g.coordinateY.push_back(y);
        g.coordinateZ.push_back(z);
        g.typeOfGridNode.push_back(19);
        g.neighborX.push_back((x + 1) % nx + nx * (y + ny * z));
        g.neighborY.push_back(x + nx * ((y + 1) % ny + ny * z));
        g.neighborZ.push_back(x + nx * (y + ny * ((z + 1) % nz)));
    }
    const std::string prefix = "host_adapters_output/run";
    FileWriter w(prefix);
    w.writeTimestep(para, 0, true);
    // expected fields
    void* q[5];
    for (auto& a : q) check(vfx_device_alloc(N * 8, &a), "alloc");
    check(vfx_macroscopic_quantities(&lat, 1, q[0], q[1], q[2], q[3], q[4], nullptr), "macro");
    std::vector<std::vector<double>> want(5, std::vector<double>(N));
    for (int i = 0; i < 5; ++i) check(vfx_copy_to_host(want[i].data(), q[i], N * 8), "d2h");
    std::ifstream in(prefix + "_bin_lev_0_ID_0_Part_1_t_0.vtk.bin.vtu", std::ios::binary);
    std::string raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const size_t at = raw.find("<AppendedData encoding=\"raw\">\n_");
    EXPECT(at != std::string::npos);
    const size_t m = (nx - 1) * (ny - 1) * (nz - 1);
    size_t pos = at + std::strlen("<AppendedData encoding=\"raw\">\n_") + (4 + 12 * N) + (4 + 32 * m) + (4 + 4 * m) +
                 (4 + m);
    const int order[5] = {4, 0, 1, 2, 3}; // press, rho, vx1, vx2, vx3
    for (int s = 0; s < 5 && at != std::string::npos; ++s) {
        std::vector<double> got(N);
        std::memcpy(got.data(), raw.data() + pos + 4, N * 8);
        pos += 4 + 8 * N;
        bool same = true;
        for (uint32_t k = 0; k < N; ++k) {
            const double e = s < 2 ? want[order[s]][k] / 3.0 : want[order[s]][k];
            same = same && got[k] == e;
        }
        EXPECT(same);
    }
    EXPECT(std::ifstream(prefix + "_bin_ID_0_t_0.vtk.pvtu").good());
    EXPECT(std::ifstream(prefix + "_bin_ID_0.pvd").good());
    // time means: two calculateMean calls on the same state, calcMean over 2
    // steps gives the state's own density and velocity; the _bin_mean files
    para->setCalcMean(true);
    allocMean(para.get());
    calculateMean(para.get(), 0);
    calculateMean(para.get(), 0);
    calcMean(para.get(), 2);
    bool mean_ok = p->meanDensityOut.size() == N;
    for (uint32_t k = 0; k < N && mean_ok; ++k)
        mean_ok = std::abs(p->meanDensityOut[k] - want[0][k]) <= 1e-15 &&
                  std::abs(p->meanVelocityInXdirectionOut[k] - want[1][k]) <= 1e-15;
    EXPECT(mean_ok);
    w.writeTimestep(para, 2, false);
    EXPECT(std::ifstream(prefix + "_bin_mean_lev_0_ID_0_Part_1_t_2.vtk.bin.vtu").good());
    EXPECT(std::ifstream(prefix + "_bin_mean_ID_0_t_2.vtk.pvtu").good());
    resetMean(para.get());
    calcMean(para.get(), 1);
    bool zero = true;
    for (uint32_t k = 0; k < N; ++k) zero = zero && p->meanPressureOut[k] == 0.0;
    EXPECT(zero);
    freeMean(para.get());
    std::filesystem::remove_all("host_adapters_output");
    for (void* d : q) vfx_device_free(d);
    for (void* d : fields) vfx_device_free(d);
    vfx_device_free(f);
}

// Output: the FilePartCalculator / WriterUtilities answers of the reference's
// unit tests (FilePartCalculatorTest.cpp, WriterUtilitiesTest.cpp)
static void output_unit_tests()
{
    const uint32_t L = FilePartCalculator::limitOfNodesForVTK;
    EXPECT(FilePartCalculator::calculateNumberOfParts(L * 2) == 3);
    EXPECT(FilePartCalculator::calculateNumberOfParts(L) == 2);
    EXPECT(FilePartCalculator::calculateNumberOfParts(L - 1) == 1);
    EXPECT(FilePartCalculator::calculateNumberOfParts(1) == 1);
    EXPECT(FilePartCalculator::calculateNumberOfNodesInPart(L + 13, 0) == L);
    EXPECT(FilePartCalculator::calculateNumberOfNodesInPart(L + 13, 1) == 13);
    bool threw = false;
    try {
        FilePartCalculator::calculateNumberOfNodesInPart(L + 13, 2);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    EXPECT(FilePartCalculator::calculateStartingPostionOfPart(0) == 0);
    EXPECT(FilePartCalculator::calculateStartingPostionOfPart(1) == L);
    HostGrid g;
    g.gridSpacing = 1.0;
    g.coordinateX = g.coordinateY = g.coordinateZ = {0.0, 1.0};
    EXPECT(!WriterUtilities::isPeriodicCell(g, 0, 1) && !WriterUtilities::isPeriodicCell(g, 1, 0));
    for (int axes = 1; axes < 8; ++axes) {
        HostGrid h = g;
        if (axes & 1) h.coordinateX[1] = 2.0;
        if (axes & 2) h.coordinateY[1] = 2.0;
        if (axes & 4) h.coordinateZ[1] = 2.0;
        EXPECT(WriterUtilities::isPeriodicCell(h, 0, 1) && WriterUtilities::isPeriodicCell(h, 1, 0));
    }
    // an oct of the node indices 7..0 (reversed), neighbours as in the reference's fixture
    HostGrid o;
    o.neighborX.assign(8, 0);
    o.neighborY.assign(8, 0);
    o.neighborZ.assign(8, 0);
    const std::array<uint32_t, 8> idx = {7, 6, 5, 4, 3, 2, 1, 0};
    for (uint32_t i = 0; i < 8; i++) {
        const uint32_t c = idx[i];
        o.neighborZ[c] = i < 4 ? idx[i + 4] : 99;
        o.neighborY[c] = (i == 0 || i == 4) ? idx[i + 3] : (i == 1 || i == 5) ? idx[i + 1] : 999;
        o.neighborX[c] = (i == 0 || i == 4) ? idx[i + 1] : (i == 3 || i == 7) ? idx[i - 1] : 9999;
    }
    std::array<uint32_t, 8> got;
    WriterUtilities::getIndicesOfAllNodesInOct(got, idx[0], o);
    EXPECT(got == idx);
    std::array<uint32_t, 8> rel;
    WriterUtilities::calculateRelativeNodeIndexInPart(rel, got, 0);
    EXPECT(rel == idx);
    WriterUtilities::calculateRelativeNodeIndexInPart(rel, got, 2);
    EXPECT(rel[0] == 5 && rel[7] == (uint32_t)-2);
}

// vtu <input> <prefix>: the parts, .pvtu and .pvd of one level from a binary
// input the Python side wrote (the two writers must give identical files)
static int vtu_mode(const std::string& input, const std::string& prefix)
{
    std::ifstream in(input, std::ios::binary);
    if (!in) return 2;
    auto rd = [&](void* p, size_t n) { in.read((char*)p, (std::streamsize)n); };
    int32_t n = 0, nf = 0, t = 0;
    rd(&n, 4);
    HostGrid g;
    for (auto* v : {&g.coordinateX, &g.coordinateY, &g.coordinateZ}) {
        v->resize(n);
        rd(v->data(), 8 * (size_t)n);
    }
    for (auto* v : {&g.typeOfGridNode, &g.neighborX, &g.neighborY, &g.neighborZ}) {
        v->resize(n);
        rd(v->data(), 4 * (size_t)n);
    }
    rd(&g.gridSpacing, 8);
    rd(&nf, 4);
    std::vector<std::vector<double>> data(nf, std::vector<double>(n));
    for (auto& d : data) rd(d.data(), 8 * (size_t)n);
    rd(&t, 4);
    if (!in) return 3;
    std::vector<std::string> names = {"press", "rho", "vx1", "vx2", "vx3", "geo"};
    if (nf == 7) names.push_back("nut");
    FileWriter w(prefix, 0);
    std::vector<std::string> files = w.writeUnstructuredGridLT(g, names, data, 0, (unsigned)t);
    for (auto& f : files) f = f.substr(f.find_last_of('/') + 1);
    WbWriterVtkXmlBinary::writeParallelFile(prefix + "_bin_ID_0_t_" + std::to_string(t) + ".vtk", files, names, {});
    std::map<uint32_t, std::vector<std::string>> series = {{(uint32_t)t, files}, {1000000u, files}};
    WbWriterVtkXmlBinary::writeCollectionForTimeSeries(prefix + "_bin_ID_0", series, false);
    return 0;
}

// the _bin_mean files of one level from host means: grid as in vtu_mode, then
// five double arrays (vx, vy, vz, rho, press means), int32 t, double density
// and velocity ratio
static int vtu_mean_mode(const std::string& input, const std::string& prefix)
{
    std::ifstream in(input, std::ios::binary);
    if (!in) return 2;
    auto rd = [&](void* p, size_t n) { in.read((char*)p, (std::streamsize)n); };
    int32_t n = 0, t = 0;
    rd(&n, 4);
    HostGrid g;
    for (auto* v : {&g.coordinateX, &g.coordinateY, &g.coordinateZ}) {
        v->resize(n);
        rd(v->data(), 8 * (size_t)n);
    }
    for (auto* v : {&g.typeOfGridNode, &g.neighborX, &g.neighborY, &g.neighborZ}) {
        v->resize(n);
        rd(v->data(), 4 * (size_t)n);
    }
    rd(&g.gridSpacing, 8);
    LBMSimulationParameter lev;
    for (auto* v : {&lev.meanVelocityInXdirectionOut, &lev.meanVelocityInYdirectionOut,
                    &lev.meanVelocityInZdirectionOut, &lev.meanDensityOut, &lev.meanPressureOut}) {
        v->resize(n);
        rd(v->data(), 8 * (size_t)n);
    }
    rd(&t, 4);
    FileWriter w(prefix, 0);
    rd(&w.densityRatio, 8);
    rd(&w.velocityRatio, 8);
    if (!in) return 3;
    std::vector<std::string> files = w.writeUnstructuredGridMeanLT(g, lev, 0, (unsigned)t);
    for (auto& f : files) f = f.substr(f.find_last_of('/') + 1);
    WbWriterVtkXmlBinary::writeParallelFile(prefix + "_bin_mean_ID_0_t_" + std::to_string(t) + ".vtk", files,
                                            FileWriter::getMeanNodeDataNames(), {});
    return 0;
}

// precursor <prefix> <n> y_0 .. y_n-1 z_0 .. z_n-1 t_0 ..: VTKReader over the
// files of <prefix> (level 0): the stencil lists, then the plane at each time
static int precursor_mode(int argc, char** argv)
{
    const std::string prefix = argv[2];
    const int n = std::atoi(argv[3]);
    if (argc < 4 + 2 * n) return 2;
    std::vector<real> y(n), z(n);
    for (int i = 0; i < n; i++) {
        y[i] = (real)std::strtod(argv[4 + i], nullptr);
        z[i] = (real)std::strtod(argv[4 + n + i], nullptr);
    }
    auto reader = createReaderForCollection(createFileCollection(prefix, TransientBCFileType::VTK), 0);
    reader->fillArrays(y, z);
    const uint m = reader->getNPointsRead();
    const size_t q = reader->getNumberOfQuantities();
    std::printf("points %u quantities %zu\n", m, q);
    std::vector<uint> nb[4];
    std::vector<real> w[4];
    for (int s = 0; s < 4; s++) {
        nb[s].resize(n);
        w[s].resize(n);
    }
    reader->getNeighbors(nb[0].data(), nb[1].data(), nb[2].data(), nb[3].data());
    reader->getWeights(w[0].data(), w[1].data(), w[2].data(), w[3].data());
    for (int s = 0; s < 4; s++) {
        std::printf("neighbors");
        for (int i = 0; i < n; i++) std::printf(" %u", nb[s][i]);
        std::printf("\nweights");
        for (int i = 0; i < n; i++) std::printf(" %.17g", (double)w[s][i]);
        std::printf("\n");
    }
    QforPrecursorBoundaryConditions pbc;
    setPrecursorFileReader(pbc, reader);
    std::vector<real> plane(m * q);
    for (int a = 4 + 2 * n; a < argc; a++) {
        pbc.getNextData(plane.data(), pbc.numberOfPrecursorNodes, std::strtod(argv[a], nullptr));
        std::printf("plane");
        for (real v : plane) std::printf(" %.17g", (double)v);
        std::printf("\n");
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc > 3 && std::strcmp(argv[1], "precursor") == 0) {
        try {
            return precursor_mode(argc, argv);
        } catch (const std::exception& e) {
            std::printf("exception: %s\n", e.what());
            return 3;
        }
    }
    if (argc > 3 && std::strcmp(argv[1], "vtu") == 0) return vtu_mode(argv[2], argv[3]);
    if (argc > 3 && std::strcmp(argv[1], "vtumean") == 0) return vtu_mean_mode(argv[2], argv[3]);
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    if (argc > 2 && std::strcmp(argv[1], "restart") == 0) {
        try {
            restart_tests(argv[2]);
        } catch (const std::exception& e) {
            std::printf("exception: %s\n", e.what());
            return 2;
        }
        std::printf("restart file tests: %d failure(s)\n", failures);
        return failures ? 1 : 0;
    }
    try {
        if (gpu) {
            gpu_tests();
            gpu_refinement_tests();
            gpu_turbulence_tests();
            gpu_output_tests();
        }
        else {
            cpu_tests();
            output_unit_tests();
        }
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s host adapter tests: %d failure(s)\n", gpu ? "gpu" : "cpu", failures);
    return failures ? 1 : 0;
}
