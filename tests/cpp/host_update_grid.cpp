// Runs a case through the C++ host layer the way VirtualFluids' Simulation
// drives a level: KernelFactoryImp::makeKernels, BoundaryConditionFactory,
// UpdateGrid27::updateGrid(0, t) for t = 1..steps.  Built twice by the tests,
// with `real` = float and (VF_DOUBLE_ACCURACY) double; the boundary tables are
// held as `real` arrays in QforBoundaryConditions / QforDirectionalBoundaryCondition
// / QforPrecursorBoundaryConditions, exactly as VirtualFluids' grid provider
// fills them, and go to the C ABI as they are.
//
//   host_update_grid <case_dir>
// reads <case_dir>/case.txt (one item per line, arrays in raw files next to it)
//   dims nx ny nz | storage s | omega w | steps n | fp16_scale s | solid file nwords
//   populations file                         node-major 27*N doubles, even parity
//   bc velocity|noslip|slip <kind> n k q vx vy vz    (file names, "-" = none)
//   bc pressure <kind> n k kn rho direction
//   outflow_correction c
//   precursor <kind> n k q nprec nquant T velocity_ratio time_ratio vx vy vz pn0..pn3 w0..w3 planes nplanes
// (arrays: .i32 / .u32 raw 32-bit integers, .f64 raw doubles; planes = nplanes
// planes of nquant*nprec values, plane m is the precursor data at time m*T)
// and writes <case_dir>/out.f64 (node-major populations after the steps, read
// at the level's parity) and <case_dir>/load_times.txt (getNextData's times).
#include "../../virtualfluids_amd/csrc/host/vfx_host.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace vfx::host;

static std::string dir;

template <typename T> static std::vector<T> read_raw(const std::string& name)
{
    std::ifstream f(dir + "/" + name, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + name);
    f.seekg(0, std::ios::end);
    const size_t bytes = (size_t)f.tellg();
    f.seekg(0);
    std::vector<T> v(bytes / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)bytes);
    return v;
}

static void* upload(const void* src, size_t bytes)
{
    void* d = nullptr;
    check(vfx_device_alloc(bytes, &d), "alloc");
    if (bytes) check(vfx_copy_to_device(d, src, bytes), "h2d");
    return d;
}

// a file of doubles as a device array of `real` (the caller's own data type)
static real* real_array(const std::string& name)
{
    if (name == "-") return nullptr;
    const std::vector<double> d = read_raw<double>(name);
    std::vector<real> r(d.begin(), d.end());
    return (real*)upload(r.data(), r.size() * sizeof(real));
}
static int* int_array(const std::string& name) { auto v = read_raw<int32_t>(name); return (int*)upload(v.data(), v.size() * 4); }
static uint* uint_array(const std::string& name) { auto v = read_raw<uint32_t>(name); return (uint*)upload(v.data(), v.size() * 4); }

static int32_t bc_kind(const std::string& s) { return std::stoi(s); }

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: host_update_grid <case_dir>\n");
        return 2;
    }
    dir = argv[1];
    try {
        auto para = std::make_shared<Parameter>(0);
        LBMSimulationParameter* p = para->getParD(0);
        vfx_lattice& lat = p->lattice;
        lat.direct = 1;
        lat.quadric_limiter[0] = lat.quadric_limiter[1] = lat.quadric_limiter[2] = 0.01;
        BoundaryConditionFactory bcf;
        int steps = 0;
        bool prime_outflow = false;
        std::vector<double> F;
        std::vector<double> load_times;
        std::vector<double> planes;
        std::ifstream mf(dir + "/case.txt");
        if (!mf) throw std::runtime_error("no case.txt");
        std::string line;
        while (std::getline(mf, line)) {
            std::istringstream is(line);
            std::string key;
            is >> key;
            if (key == "dims") {
                is >> lat.dims[0] >> lat.dims[1] >> lat.dims[2];
                lat.n_nodes = (uint64_t)lat.dims[0] * lat.dims[1] * lat.dims[2];
            } else if (key == "storage") {
                is >> lat.storage;
            } else if (key == "omega") {
                is >> lat.omega;
            } else if (key == "steps") {
                is >> steps;
            } else if (key == "fp16_scale") {
                is >> lat.fp16_scale;
            } else if (key == "solid") {
                std::string f;
                is >> f;
                lat.solid_mask = uint_array(f);
            } else if (key == "populations") {
                std::string f;
                is >> f;
                F = read_raw<double>(f);
            } else if (key == "outflow_correction") {
                double c;
                is >> c;
                para->setOutflowPressureCorrectionFactor(c);
            } else if (key == "bc") {
                std::string group, kind, k, q;
                unsigned n;
                is >> group >> kind >> n >> k;
                if (group == "pressure") {
                    std::string kn, rho;
                    size_t direction;
                    is >> kn >> rho >> direction;
                    QforDirectionalBoundaryCondition b;
                    b.k = int_array(k);
                    b.kN = int_array(kn);
                    b.RhoBC = real_array(rho);
                    b.numberOfBCnodes = n;
                    b.direction = direction;
                    p->pressureBCDirectional.push_back(b);
                    const int32_t kd = bc_kind(kind);
                    prime_outflow = kd != VFX_BC_PRESSURE_NEQ;
                    bcf.setPressureBoundaryCondition(
                        kd == VFX_BC_PRESSURE_NEQ ? BoundaryConditionFactory::PressureBC::PressureNonEquilibriumCompressible
                        : kd == VFX_BC_OUTFLOW_NON_REFLECTING ? BoundaryConditionFactory::PressureBC::OutflowNonReflective
                                                              : BoundaryConditionFactory::PressureBC::OutflowNonReflectivePressureCorrection);
                    continue;
                }
                std::string vx, vy, vz;
                is >> q >> vx >> vy >> vz;
                QforBoundaryConditions* b = group == "velocity" ? &p->velocityBC : group == "noslip" ? &p->noSlipBC : &p->slipBC;
                b->k = int_array(k);
                b->numberOfBCnodes = n;
                // q27[d] = q27[0] + d * numberOfBCnodes, as the grid provider sets the pointers
                b->q27[0] = real_array(q);
                for (int d = 1; d < 27; ++d) b->q27[d] = b->q27[0] + (size_t)d * n;
                b->Vx = real_array(vx);
                b->Vy = real_array(vy);
                b->Vz = real_array(vz);
                const int32_t kd = bc_kind(kind);
                using F_ = BoundaryConditionFactory;
                if (group == "velocity")
                    bcf.setVelocityBoundaryCondition(kd == VFX_BC_VELOCITY_BOUNCE_BACK ? F_::VelocityBC::VelocityBounceBack
                                                     : kd == VFX_BC_VELOCITY_INTERPOLATED
                                                         ? F_::VelocityBC::VelocityInterpolatedCompressible
                                                         : F_::VelocityBC::VelocityWithPressureInterpolatedCompressible);
                else if (group == "noslip")
                    bcf.setNoSlipBoundaryCondition(kd == VFX_BC_VELOCITY_BOUNCE_BACK ? F_::NoSlipBC::NoSlipBounceBack
                                                                                    : F_::NoSlipBC::NoSlipInterpolatedCompressible);
                else
                    bcf.setSlipBoundaryCondition(F_::SlipBC::SlipCompressible);
            } else if (key == "precursor") {
                std::string kind, k, q, vxs, vys, vzs, pn[4], w[4], pf;
                unsigned n, nprec, nquant, T, nplanes;
                double vratio, tratio;
                is >> kind >> n >> k >> q >> nprec >> nquant >> T >> vratio >> tratio >> vxs >> vys >> vzs;
                for (auto& s : pn) is >> s;
                for (auto& s : w) is >> s;
                is >> pf >> nplanes;
                QforPrecursorBoundaryConditions& b = p->precursorBC;
                b.k = int_array(k);
                b.numberOfBCnodes = (int)n;
                b.numberOfPrecursorNodes = (int)nprec;
                b.numberOfQuantities = nquant;
                b.timeStepsBetweenReads = T;
                if (q != "-") {
                    b.q27[0] = real_array(q);
                    for (int d = 1; d < 27; ++d) b.q27[d] = b.q27[0] + (size_t)d * n;
                }
                b.velocityX = (real)std::stod(vxs);
                b.velocityY = (real)std::stod(vys);
                b.velocityZ = (real)std::stod(vzs);
                b.planeNeighbor0PP = uint_array(pn[0]);
                b.planeNeighbor0PM = uint_array(pn[1]);
                b.planeNeighbor0MP = uint_array(pn[2]);
                b.planeNeighbor0MM = uint_array(pn[3]);
                b.weights0PP = real_array(w[0]);
                b.weights0PM = real_array(w[1]);
                b.weights0MP = real_array(w[2]);
                b.weights0MM = real_array(w[3]);
                planes = read_raw<double>(pf);
                const size_t psz = (size_t)nquant * nprec;
                if (planes.size() != (size_t)nplanes * psz) throw std::runtime_error("plane file size");
                // GridGenerator.cpp:436-478: planes at 0, T and 2T are loaded before the run
                auto plane_dev = [&](size_t m) {
                    std::vector<real> r(planes.begin() + m * psz, planes.begin() + (m + 1) * psz);
                    return (real*)upload(r.data(), psz * sizeof(real));
                };
                b.last = plane_dev(0);
                b.current = plane_dev(1);
                b.next = plane_dev(2);
                b.nPrecursorReads = 2;
                b.getNextData = [&planes, &load_times, psz, T, nplanes](real* data, int, double loadTime) {
                    load_times.push_back(loadTime);
                    const size_t m = (size_t)std::llround(loadTime / T);
                    if (m >= nplanes) throw std::runtime_error("precursor plane beyond the file");
                    for (size_t i = 0; i < psz; ++i) data[i] = (real)planes[m * psz + i];
                };
                para->setVelocityRatio(vratio);
                (void)tratio;
                const int32_t kd = bc_kind(kind);
                bcf.setPrecursorBoundaryCondition(kd == VFX_BC_PRECURSOR_VELOCITY
                                                      ? BoundaryConditionFactory::PrecursorBC::PrecursorNonReflectiveCompressible
                                                      : BoundaryConditionFactory::PrecursorBC::PrecursorDistributions);
            } else if (!key.empty()) {
                throw std::runtime_error("unknown case key " + key);
            }
        }
        const size_t esz = lat.storage == VFX_FP64 ? 8 : lat.storage == VFX_FP32 ? 4 : 2;
        check(vfx_device_alloc(27 * lat.n_nodes * esz, &lat.f), "lattice alloc");
        void* dF = upload(F.data(), F.size() * 8);
        check(vfx_write_populations(&lat, 1, (const double*)dF, nullptr), "write populations");
        p->isEvenTimestep = true;
        // the outflow boundaries write what the collision after next reads:
        // apply them once with the starting parity, as the tests of the C ABI do
        if (prime_outflow)
            for (auto& q : p->pressureBCDirectional)
                std::get<boundaryConditionDirectional>(bcf.getPressureBoundaryConditionPre())(p, &q);
        KernelFactoryImp kf;
        UpdateGrid27 ug(para, kf.makeKernels(para), bcf);
        for (int t = 1; t <= steps; ++t) ug.updateGrid(0, (unsigned)t);
        check(vfx_synchronize(nullptr), "sync");
        check(vfx_read_populations(&lat, p->isEvenTimestep ? 1 : 0, (double*)dF, nullptr), "read populations");
        check(vfx_synchronize(nullptr), "sync");
        check(vfx_copy_to_host(F.data(), dF, F.size() * 8), "d2h");
        std::ofstream out(dir + "/out.f64", std::ios::binary);
        out.write(reinterpret_cast<const char*>(F.data()), (std::streamsize)(F.size() * 8));
        std::ofstream lt(dir + "/load_times.txt");
        for (double t : load_times) lt << t << "\n";
        std::ofstream par(dir + "/parity.txt");
        par << (p->isEvenTimestep ? 1 : 0) << "\n" << sizeof(real) << "\n";
    } catch (const std::exception& e) {
        std::fprintf(stderr, "host_update_grid: %s\n", e.what());
        return 1;
    }
    return 0;
}
