// vfx_host.cpp -- implementation of the C++ host layer (see vfx_host.hpp).
#include "vfx_host.hpp"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

VFX_HOST_BEGIN

void check(vfx_status s, const char* what)
{
    if (s != VFX_OK) throw std::runtime_error(std::string(what) + ": " + vfx_last_error());
}

static int32_t vfx_template(CollisionTemplate t)
{
    switch (t) {
    case CollisionTemplate::Default: return VFX_TEMPLATE_DEFAULT;
    case CollisionTemplate::WriteMacroVars: return VFX_TEMPLATE_WRITE_MACRO;
    case CollisionTemplate::ApplyBodyForce: return VFX_TEMPLATE_APPLY_BODY_FORCE;
    // K17CompressibleNavierStokes.cu:126-131: the subdomain border runs with all features
    case CollisionTemplate::SubDomainBorder:
    case CollisionTemplate::AllFeatures: return VFX_TEMPLATE_ALL_FEATURES;
    default: throw std::runtime_error("unknown collision template");
    }
}

int32_t vfx_turbulence_model_of(TurbulenceModel m)
{
    switch (m) {
    case TurbulenceModel::Smagorinsky: return VFX_TURB_SMAGORINSKY;
    case TurbulenceModel::AMD: return VFX_TURB_AMD;
    case TurbulenceModel::QR: return VFX_TURB_QR;
    default: return VFX_TURB_NONE;
    }
}

// the collision parameters the kernels take from Parameter
static void set_collision_parameters(vfx_lattice& lat, const Parameter& para)
{
    const double* lim = para.getQuadricLimiters();
    for (int i = 0; i < 3; ++i) lat.quadric_limiter[i] = lim[i];
    lat.turbulence_model = vfx_turbulence_model_of(para.getTurbulenceModel());
    lat.sgs_constant = para.getSGSConstant();
}

void TurbulenceModelFactory::setTurbulenceModel(TurbulenceModel m)
{
    turbulenceModel = m;
    para->setTurbulenceModel(m);
    if (m != TurbulenceModel::None) para->setUseTurbulentViscosity(true);
}

void TurbulenceModelFactory::readConfig(const std::map<std::string, std::string>& config)
{
    auto it = config.find("TurbulenceModel");
    if (it != config.end()) {
        const std::string& v = it->second;
        if (v == "Smagorinsky") setTurbulenceModel(TurbulenceModel::Smagorinsky);
        else if (v == "AMD") setTurbulenceModel(TurbulenceModel::AMD);
        else if (v == "QR") setTurbulenceModel(TurbulenceModel::QR);
        else if (v == "None") setTurbulenceModel(TurbulenceModel::None);
        else throw std::runtime_error("TurbulenceModelFactory: Invalid turbulence model!");
    }
    it = config.find("SGSconstant");
    if (it != config.end()) para->setSGSConstant(std::stod(it->second));
}

// calcTurbulentViscosityAMD (TurbulentViscosityKernels.cu:115-130) for the AMD
// model; the other models need no kernel between steps
void TurbulenceModelFactory::runTurbulenceModelKernel(int level) const
{
    if (turbulenceModel != TurbulenceModel::AMD) return;
    LBMSimulationParameter* p = para->getParD(level);
    vfx_lattice lat = p->lattice;
    set_collision_parameters(lat, *para);
    check(vfx_turbulent_viscosity_amd(&lat, p->stream), "calcTurbulentViscosityAMD");
}

void KernelImp::run()
{
    LBMSimulationParameter* p = para->getParD(level);
    vfx_lattice lat = p->lattice;
    set_collision_parameters(lat, *para);
    // the wall models sample the velocities of this step's collision
    const int32_t tmpl = p->stressBC.numberOfBCnodes ? VFX_TEMPLATE_WRITE_MACRO : VFX_TEMPLATE_DEFAULT;
    check(vfx_collide_stream(&lat, collision, p->isEvenTimestep ? 1 : 0, tmpl, p->stream), "Kernel::run");
}

// K17CompressibleNavierStokes.cu:102-145: one launch over a list of fluid nodes
void KernelImp::runOnIndices(const unsigned int* indices, unsigned int size_indices, CollisionTemplate collisionTemplate,
                             CudaStreamIndex)
{
    LBMSimulationParameter* p = para->getParD(level);
    vfx_lattice lat = p->lattice;
    if (lat.direct) throw std::runtime_error("runOnIndices needs the sparse (neighbour array) layout");
    // a regular-wavefront table describes the level's own fluid-node list only
    if (indices != lat.fluid_nodes || size_indices != lat.n_fluid) lat.regular_groups = nullptr;
    lat.fluid_nodes = indices;
    lat.n_fluid = size_indices;
    set_collision_parameters(lat, *para);
    int32_t tmpl = vfx_template(collisionTemplate);
    // a level without body-force arrays runs its subdomain border with the macroscopic write only
    if (collisionTemplate == CollisionTemplate::SubDomainBorder && !lat.body_force[0]) tmpl = VFX_TEMPLATE_WRITE_MACRO;
    check(vfx_collide_stream(&lat, collision, p->isEvenTimestep ? 1 : 0, tmpl, p->stream), "Kernel::runOnIndices");
}

std::vector<std::shared_ptr<Kernel>> KernelFactoryImp::makeKernels(std::shared_ptr<Parameter> para)
{
    std::vector<std::shared_ptr<Kernel>> kernels;
    for (int level = 0; level <= para->getMaxLevel(); level++) kernels.push_back(makeKernel(para, para->getMainKernel(), level));
    return kernels;
}

std::shared_ptr<Kernel> KernelFactoryImp::makeKernel(std::shared_ptr<Parameter> para, const std::string& kernel, int level)
{
    if (kernel == "BGKCompSP27") return B92CompressibleNavierStokes::getNewInstance(para, level);
    if (kernel == "K17CompressibleNavierStokes") return K17CompressibleNavierStokes::getNewInstance(para, level);
    throw std::runtime_error("KernelFactory does not know the KernelType.");
}

// The lattice storage must carry the host layer's `real`: its boundary and
// interface tables go to the C ABI as they are.
static void check_real(const vfx_lattice& lat, const char* what)
{
    if ((lat.storage == VFX_FP64) != (sizeof(real) == 8))
        throw std::runtime_error(std::string(what) +
                                 (sizeof(real) == 8 ? ": the host layer built with VF_DOUBLE_ACCURACY needs VFX_FP64 storage"
                                                    : ": the host layer built with float real needs VFX_FP32 or "
                                                      "VFX_FP16_SCALED storage"));
}

// VelocityBounceBack / NoSlip* / Slip* launchers (e.g. Velocity.cu, NoSlip.cu):
// QforBoundaryConditions k, q27[0], Vx/Vy/Vz as they are
static boundaryCondition make_bc(int32_t kind)
{
    return [kind](LBMSimulationParameter* p, QforBoundaryConditions* q) {
        check_real(p->lattice, "boundary condition");
        vfx_bc bc{};
        bc.kind = kind;
        bc.n = q->numberOfBCnodes;
        bc.k = reinterpret_cast<const uint32_t*>(q->k);
        bc.q27 = q->q27[0];
        bc.vx = q->Vx;
        bc.vy = q->Vy;
        bc.vz = q->Vz;
        check(vfx_bc_apply(&p->lattice, &bc, p->isEvenTimestep ? 1 : 0, p->stream), "boundary condition");
    };
}

// Pressure.cu:63-80, Outflow.cu:44-90: QforDirectionalBoundaryCondition k, kN,
// RhoBC and direction (the outflow kinds also the level's correction factor)
static boundaryConditionDirectional make_directional(int32_t kind)
{
    return [kind](LBMSimulationParameter* p, QforDirectionalBoundaryCondition* q) {
        check_real(p->lattice, "pressure boundary condition");
        vfx_bc bc{};
        bc.kind = kind;
        bc.n = q->numberOfBCnodes;
        bc.k = reinterpret_cast<const uint32_t*>(q->k);
        bc.q27 = q->q27[0];
        bc.rho = q->RhoBC;
        bc.k_neighbor = reinterpret_cast<const uint32_t*>(q->kN);
        bc.normal = (int32_t)q->direction;
        bc.pressure_correction_factor = p->outflowPressureCorrectionFactor;
        check(vfx_bc_apply(&p->lattice, &bc, p->isEvenTimestep ? 1 : 0, p->stream), "pressure boundary condition");
    };
}

// Stress.cu:43-80: the boundary's k, kN (sampling node), q27[0], normals,
// Vx/Vy/Vz and Vx1/Vy1/Vz1 (filtered velocities, updated) and the level's
// WallModelParameters (monitors written when hasMonitor)
static boundaryConditionWithParameter make_stress(int32_t kind)
{
    return [kind](Parameter* para, QforBoundaryConditions* q, const int level) {
        LBMSimulationParameter* p = para->getParD(level);
        check_real(p->lattice, "stress boundary condition");
        vfx_bc bc{};
        bc.kind = kind;
        bc.n = q->numberOfBCnodes;
        bc.k = reinterpret_cast<const uint32_t*>(q->k);
        bc.q27 = q->q27[0];
        bc.k_neighbor = reinterpret_cast<const uint32_t*>(q->kN);
        bc.wm_normal[0] = q->normalX;
        bc.wm_normal[1] = q->normalY;
        bc.wm_normal[2] = q->normalZ;
        bc.wm_exchange_velocity[0] = q->Vx;
        bc.wm_exchange_velocity[1] = q->Vy;
        bc.wm_exchange_velocity[2] = q->Vz;
        bc.wm_wall_velocity[0] = q->Vx1;
        bc.wm_wall_velocity[1] = q->Vy1;
        bc.wm_wall_velocity[2] = q->Vz1;
        const WallModelParameters& wm = p->wallModel;
        bc.wm_sampling_offset = wm.samplingOffset;
        bc.wm_z0 = wm.z0;
        if (wm.hasMonitor) {
            bc.wm_u_star = wm.u_star;
            bc.wm_force[0] = wm.Fx;
            bc.wm_force[1] = wm.Fy;
            bc.wm_force[2] = wm.Fz;
        }
        check(vfx_bc_apply(&p->lattice, &bc, p->isEvenTimestep ? 1 : 0, p->stream), "stress boundary condition");
    };
}

// Precursor.cu:44-117 through vfx_bc_apply kinds 12 / 13 (velocityX/Y/Z are
// the reference's scalars, handed over as they are)
precursorBoundaryConditionFunc BoundaryConditionFactory::getPrecursorBoundaryConditionPost() const
{
    int32_t kind;
    switch (precursor) {
    case PrecursorBC::PrecursorNonReflectiveCompressible: kind = VFX_BC_PRECURSOR_VELOCITY; break;
    case PrecursorBC::PrecursorDistributions: kind = VFX_BC_PRECURSOR_DISTRIBUTIONS; break;
    default: return nullptr;
    }
    return [kind](LBMSimulationParameter* p, QforPrecursorBoundaryConditions* q, real tRatio, real velocityRatio) {
        check_real(p->lattice, "precursor boundary condition");
        vfx_bc bc{};
        bc.kind = kind;
        bc.n = (uint32_t)q->numberOfBCnodes;
        bc.k = reinterpret_cast<const uint32_t*>(q->k);
        bc.q27 = q->q27[0];
        bc.precursor_velocity[0] = q->velocityX;
        bc.precursor_velocity[1] = q->velocityY;
        bc.precursor_velocity[2] = q->velocityZ;
        bc.n_precursor_nodes = (uint32_t)q->numberOfPrecursorNodes;
        bc.plane_neighbor[0] = q->planeNeighbor0PP;
        bc.plane_neighbor[1] = q->planeNeighbor0PM;
        bc.plane_neighbor[2] = q->planeNeighbor0MP;
        bc.plane_neighbor[3] = q->planeNeighbor0MM;
        bc.plane_weight[0] = q->weights0PP;
        bc.plane_weight[1] = q->weights0PM;
        bc.plane_weight[2] = q->weights0MP;
        bc.plane_weight[3] = q->weights0MM;
        bc.precursor_last = q->last;
        bc.precursor_current = q->current;
        bc.time_ratio = tRatio;
        bc.velocity_ratio = velocityRatio;
        check(vfx_bc_apply(&p->lattice, &bc, p->isEvenTimestep ? 1 : 0, p->stream), "precursor boundary condition");
    };
}

boundaryCondition BoundaryConditionFactory::getVelocityBoundaryConditionPost() const
{
    switch (velocity) {
    case VelocityBC::VelocityBounceBack: return make_bc(VFX_BC_VELOCITY_BOUNCE_BACK);
    case VelocityBC::VelocityInterpolatedCompressible: return make_bc(VFX_BC_VELOCITY_INTERPOLATED);
    case VelocityBC::VelocityWithPressureInterpolatedCompressible: return make_bc(VFX_BC_VELOCITY_PRESSURE_INTERPOLATED);
    default: return nullptr;
    }
}
boundaryCondition BoundaryConditionFactory::getNoSlipBoundaryConditionPost() const
{
    switch (noSlip) {
    case NoSlipBC::NoSlipBounceBack: return make_bc(VFX_BC_VELOCITY_BOUNCE_BACK); // zero wall velocity
    case NoSlipBC::NoSlipInterpolatedCompressible: return make_bc(VFX_BC_NOSLIP_INTERPOLATED);
    default: return nullptr;
    }
}
boundaryCondition BoundaryConditionFactory::getSlipBoundaryConditionPost() const
{
    switch (slip) {
    case SlipBC::SlipCompressible: return make_bc(VFX_BC_SLIP);
    case SlipBC::SlipTurbulentViscosityCompressible: return make_bc(VFX_BC_SLIP_TURBULENT_VISCOSITY);
    default: return nullptr;
    }
}
boundaryConditionWithParameter BoundaryConditionFactory::getStressBoundaryConditionPost() const
{
    switch (stress) {
    case StressBC::StressCompressible: return make_stress(VFX_BC_STRESS_INTERPOLATED);
    case StressBC::StressBounceBackCompressible: return make_stress(VFX_BC_STRESS_BOUNCE_BACK);
    case StressBC::StressBounceBackPressureCompressible: return make_stress(VFX_BC_STRESS_BOUNCE_BACK_PRESSURE);
    default: return nullptr;
    }
}

std::variant<boundaryCondition, boundaryConditionDirectional> BoundaryConditionFactory::getPressureBoundaryConditionPre() const
{
    switch (pressure) {
    case PressureBC::PressureNonEquilibriumCompressible: return make_directional(VFX_BC_PRESSURE_NEQ);
    case PressureBC::OutflowNonReflective: return make_directional(VFX_BC_OUTFLOW_NON_REFLECTING);
    case PressureBC::OutflowNonReflectivePressureCorrection:
        return make_directional(VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION);
    default: return boundaryCondition(nullptr);
    }
}
bool BoundaryConditionFactory::hasDirectionalPressureBoundaryCondition() const
{
    return std::holds_alternative<boundaryConditionDirectional>(getPressureBoundaryConditionPre());
}

// BoundaryConditionKernelManager.cpp:53-83: boundary nodes without a selected
// boundary condition are an error
static void require(const boundaryCondition& f, const QforBoundaryConditions& q, const char* name)
{
    if (q.numberOfBCnodes > 0 && !f)
        throw std::runtime_error(std::string(name) + " boundary nodes exist but no boundary condition was selected");
}

// Scaling.cu:105-188: the interface kernels take the coarse level's parity and
// treat the fine level as even (both levels are back at their starting parity
// when the interpolation runs)
static gridScaling make_scaling(bool coarse_to_fine)
{
    return [coarse_to_fine](LBMSimulationParameter* c, LBMSimulationParameter* f, ICells* cells, ICellNeigh& nb,
                            vfx_stream stream) {
        vfx_icells ic{};
        ic.fine_cells = cells->fineCellIndices;
        ic.coarse_cells = cells->coarseCellIndices;
        ic.n_cells = cells->numberOfCells;
        check_real(c->lattice, "grid scaling");
        ic.offset_x = nb.x;
        ic.offset_y = nb.y;
        ic.offset_z = nb.z;
        const int32_t even_c = c->isEvenTimestep ? 1 : 0;
        if (coarse_to_fine) check(vfx_scale_cf(&c->lattice, &f->lattice, &ic, even_c, 1, stream), "scaleCoarseToFine");
        else check(vfx_scale_fc(&c->lattice, &f->lattice, &ic, even_c, 1, stream), "scaleFineToCoarse");
    };
}

// GridScalingFactory.cpp: ScaleCompressible selects the compressible scaling;
// hasTurbulentViscosity selects the per-node rates, which the C ABI takes from
// the levels' turbulence_model / turbulent_viscosity fields
gridScaling GridScalingFactory::getGridScalingFC(bool) const
{
    if (type != GridScaling::ScaleCompressible) return nullptr;
    return make_scaling(false);
}
gridScaling GridScalingFactory::getGridScalingCF(bool) const
{
    if (type != GridScaling::ScaleCompressible) return nullptr;
    return make_scaling(true);
}

UpdateGrid27::UpdateGrid27(std::shared_ptr<Parameter> p, std::vector<std::shared_ptr<Kernel>> k,
                           const BoundaryConditionFactory& f, const GridScalingFactory* s,
                           std::shared_ptr<TurbulenceModelFactory> tm)
    : tmFactory(std::move(tm)), para(std::move(p)), kernels(std::move(k)), velocityBC(f.getVelocityBoundaryConditionPost()),
      noSlipBC(f.getNoSlipBoundaryConditionPost()), slipBC(f.getSlipBoundaryConditionPost()),
      stressBC(f.getStressBoundaryConditionPost()), precursorBC(f.getPrecursorBoundaryConditionPost())
{
    if (f.hasDirectionalPressureBoundaryCondition())
        directionalPressureBC = std::get<boundaryConditionDirectional>(f.getPressureBoundaryConditionPre());
    if ((int)kernels.size() != para->getMaxLevel() + 1) throw std::runtime_error("one kernel per level expected");
    // GridScalingKernelManager.cpp:45-60
    if (para->getMaxLevel() != 0) {
        if (!s) throw std::runtime_error("There is more than one level, but no scalingFactory was provided.");
        scalingFineToCoarse = s->getGridScalingFC(para->getUseTurbulentViscosity());
        scalingCoarseToFine = s->getGridScalingCF(para->getUseTurbulentViscosity());
        for (int l = 0; l < para->getMaxLevel(); ++l) {
            if (para->getParD(l)->fineToCoarse.numberOfCells && !scalingFineToCoarse)
                throw std::runtime_error("scalingFineToCoarse: interface cells exist but no scaling function was selected");
            if (para->getParD(l)->coarseToFine.numberOfCells && !scalingCoarseToFine)
                throw std::runtime_error("scalingCoarseToFine: interface cells exist but no scaling function was selected");
        }
    }
    for (int l = 0; l <= para->getMaxLevel(); ++l) {
        LBMSimulationParameter* d = para->getParD(l);
        require(velocityBC, d->velocityBC, "velocity");
        require(noSlipBC, d->noSlipBC, "no-slip");
        require(slipBC, d->slipBC, "slip");
        for (const auto& q : d->pressureBCDirectional)
            if (q.numberOfBCnodes > 0 && !directionalPressureBC)
                throw std::runtime_error("pressure boundary nodes exist but no boundary condition was selected");
        if (d->stressBC.numberOfBCnodes > 0 && !stressBC)
            throw std::runtime_error("stress boundary nodes exist but no boundary condition was selected");
        if (d->precursorBC.numberOfBCnodes && !precursorBC)
            throw std::runtime_error("precursor boundary nodes exist but no boundary condition was selected");
        if (d->precursorBC.numberOfBCnodes && l != 0)
            throw std::runtime_error("precursor boundary conditions are supported on level 0");
    }
}

// BoundaryConditionKernelManager.cpp:346-382 on level 0 (t_level = t): once
// the step reaches the time of `current`, rotate last <- current <- next, read
// the following plane (getNextData fills a host array, copied to next as
// cudaCopyPrecursorData does) and pass (t - lastTime) / timeStepsBetweenReads
void UpdateGrid27::runPrecursorBCKernelPost(int level, unsigned int t)
{
    LBMSimulationParameter* p = para->getParD(level);
    QforPrecursorBoundaryConditions& q = p->precursorBC;
    if (q.numberOfBCnodes == 0) return;
    const uint32_t T = q.timeStepsBetweenReads;
    uint32_t lastTime = (q.nPrecursorReads - 2) * T;
    const uint32_t currentTime = (q.nPrecursorReads - 1) * T;
    uint32_t nextTime = q.nPrecursorReads * T;
    if (t >= currentTime) {
        lastTime = currentTime;
        nextTime += T;
        real* tmp = q.last;
        q.last = q.current;
        q.current = q.next;
        q.next = tmp;
        if (q.getNextData) {
            const size_t n = (size_t)q.numberOfPrecursorNodes * q.numberOfQuantities;
            std::vector<real> plane(n);
            q.getNextData(plane.data(), q.numberOfPrecursorNodes, nextTime * std::exp2(-level) * para->getTimeRatio());
            // the kernels of this level read the planes on p->stream: finish them before overwriting `next`
            check(vfx_synchronize(p->stream), "precursor plane read");
            check(vfx_copy_to_device(q.next, plane.data(), n * sizeof(real)), "precursor plane copy");
        }
        q.nPrecursorReads++;
    }
    const real tRatio = real(t - lastTime) / real(T);
    precursorBC(p, &q, tRatio, (real)para->getVelocityRatio());
}

void UpdateGrid27::postCollisionBC(int level, unsigned int timestep)
{
    LBMSimulationParameter* p = para->getParD(level);
    if (p->velocityBC.numberOfBCnodes) velocityBC(p, &p->velocityBC);
    if (p->noSlipBC.numberOfBCnodes) noSlipBC(p, &p->noSlipBC);
    if (p->slipBC.numberOfBCnodes) slipBC(p, &p->slipBC);
    if (p->stressBC.numberOfBCnodes) stressBC(para.get(), &p->stressBC, level); // UpdateGrid27.cpp:241-243
    runPrecursorBCKernelPost(level, timestep);                                    // :252
}

// BoundaryConditionKernelManager::runPressureBCKernelPre (:321-328): every
// directional pressure/outflow boundary in turn
void UpdateGrid27::preCollisionBC(int level)
{
    LBMSimulationParameter* p = para->getParD(level);
    for (auto& q : p->pressureBCDirectional)
        if (q.numberOfBCnodes) directionalPressureBC(p, &q);
}

void UpdateGrid27::refinement(int level)
{
    LBMSimulationParameter* c = para->getParD(level);
    LBMSimulationParameter* f = para->getParD(level + 1);
    if (scalingFineToCoarse && c->fineToCoarse.numberOfCells)
        scalingFineToCoarse(c, f, &c->fineToCoarse, c->neighborFineToCoarse, c->stream);
    if (scalingCoarseToFine && c->coarseToFine.numberOfCells)
        scalingCoarseToFine(c, f, &c->coarseToFine, c->neighborCoarseToFine, c->stream);
}

void UpdateGrid27::updateGrid(int level, unsigned int timestep)
{
    const bool finest = level == para->getMaxLevel();
    if (!finest) {
        updateGrid(level + 1, timestep);
        updateGrid(level + 1, timestep);
    }
    LBMSimulationParameter* p = para->getParD(level);
    // the level's lattice carries the model for the interface interpolation
    // (per-node rates with a turbulent viscosity)
    set_collision_parameters(p->lattice, *para);
    auto* k = dynamic_cast<KernelImp*>(kernels.at(level).get());
    if (k && k->getKernelUsesFluidNodeIndices() && !p->lattice.direct) {
        for (CollisionTemplate tag : p->allocatedBulkFluidNodeTags) {
            auto it = p->taggedFluidNodeIndices.find(tag);
            if (it == p->taggedFluidNodeIndices.end() || it->second.second == 0) continue;
            kernels[level]->runOnIndices(it->second.first, it->second.second, tag);
        }
    } else {
        kernels[level]->run();
    }
    postCollisionBC(level, timestep);
    p->isEvenTimestep = !p->isEvenTimestep; // UpdateGrid27.cpp:264-268
    if (tmFactory) tmFactory->runTurbulenceModelKernel(level); // :81
    preCollisionBC(level);
    if (!finest) refinement(level);
}

// -------------------------------------------------------------- restart ----
template <typename Real>
void RestartObjectT<Real>::serialize(const std::string& filename, const std::shared_ptr<Parameter>& para)
{
    fs.clear();
    constexpr int32_t precision = sizeof(Real) == 8 ? VFX_FP64 : VFX_FP32;
    for (int level = 0; level <= para->getMaxLevel(); ++level) {
        LBMSimulationParameter* p = para->getParD(level);
        const uint64_t n = 27 * p->lattice.n_nodes;
        void* dev = nullptr;
        check(vfx_device_alloc(n * sizeof(Real), &dev), "RestartObject::serialize alloc");
        std::vector<Real> host(n);
        vfx_status st = vfx_export_distributions(&p->lattice, p->isEvenTimestep ? 1 : 0, dev, precision, p->stream);
        if (st == VFX_OK) st = vfx_synchronize(p->stream);
        if (st == VFX_OK) st = vfx_copy_to_host(host.data(), dev, n * sizeof(Real));
        vfx_device_free(dev);
        check(st, "RestartObject::serialize");
        fs.push_back(std::move(host));  // NaN already written as 0 by the export
    }
    serialize_internal(filename);
}

template <typename Real>
void RestartObjectT<Real>::deserialize(const std::string& filename, std::shared_ptr<Parameter>& para)
{
    fs.clear();
    deserialize_internal(filename);
    if ((int)fs.size() != para->getMaxLevel() + 1)
        throw std::runtime_error("restart file " + filename + " holds " + std::to_string(fs.size()) + " levels, the parameter " +
                                 std::to_string(para->getMaxLevel() + 1));
    constexpr int32_t precision = sizeof(Real) == 8 ? VFX_FP64 : VFX_FP32;
    for (int level = 0; level <= para->getMaxLevel(); ++level) {
        LBMSimulationParameter* p = para->getParD(level);
        const uint64_t n = 27 * p->lattice.n_nodes;
        if (fs[level].size() != n)
            throw std::runtime_error("restart file " + filename + ": level " + std::to_string(level) + " holds " +
                                     std::to_string(fs[level].size()) + " values, the lattice needs " + std::to_string(n));
        void* dev = nullptr;
        check(vfx_device_alloc(n * sizeof(Real), &dev), "RestartObject::deserialize alloc");
        vfx_status st = vfx_copy_to_device(dev, fs[level].data(), n * sizeof(Real));
        if (st == VFX_OK)
            st = vfx_import_distributions(&p->lattice, p->isEvenTimestep ? 1 : 0, dev, precision, p->stream);
        if (st == VFX_OK) st = vfx_synchronize(p->stream);
        vfx_device_free(dev);
        check(st, "RestartObject::deserialize");
    }
}

template <typename Real> bool RestartObjectT<Real>::delete_restart_file(const std::string& filename)
{
    const std::string file = filename + getFileExtension();
    if (std::remove(file.c_str()) != 0) {
        std::fprintf(stdout, "[warning] Could not delete the restart object file \"%s\".\n", file.c_str());
        return false;
    }
    return true;
}

// one line per level, each value followed by a space, default stream format
template <typename Real> void ASCIIRestartObjectT<Real>::serialize_internal(const std::string& filename)
{
    std::ofstream stream(filename + getFileExtension());
    if (!stream) throw std::runtime_error("couldn't open file " + filename);
    for (const auto& level : this->fs) {
        for (Real v : level) stream << v << " ";
        stream << "\n";
    }
}

template <typename Real> void ASCIIRestartObjectT<Real>::deserialize_internal(const std::string& filename)
{
    std::ifstream stream(filename + getFileExtension(), std::ios_base::in);
    if (!stream.is_open()) throw std::runtime_error("couldn't open check point file " + filename);
    std::string line;
    while (std::getline(stream, line)) {
        std::vector<Real> values;
        std::stringstream ls(line);
        Real v;
        while (ls >> v) values.push_back(v);
        this->fs.push_back(std::move(values));
    }
}

// int32 level count, then per level an int32 count and the raw values
template <typename Real> void BinaryRestartObjectT<Real>::serialize_internal(const std::string& filename)
{
    std::ofstream stream(filename + getFileExtension(), std::ios_base::binary);
    if (!stream) throw std::runtime_error("couldn't open file " + filename);
    const int32_t levels = (int32_t)this->fs.size();
    stream.write(reinterpret_cast<const char*>(&levels), sizeof(levels));
    for (const auto& v : this->fs) {
        const int32_t size = (int32_t)v.size();
        stream.write(reinterpret_cast<const char*>(&size), sizeof(size));
        if (size) stream.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(Real)));
    }
}

template <typename Real> void BinaryRestartObjectT<Real>::deserialize_internal(const std::string& filename)
{
    std::ifstream stream(filename + getFileExtension(), std::ios_base::in | std::ios_base::binary);
    if (!stream.is_open()) throw std::runtime_error("couldn't open check point file " + filename);
    int32_t levels = 0;
    stream.read(reinterpret_cast<char*>(&levels), sizeof(levels));
    if (!stream || levels < 0) throw std::runtime_error("check point file " + filename + " is not a restart object");
    this->fs.assign((size_t)levels, {});
    for (int32_t n = 0; n < levels; ++n) {
        int32_t size = 0;
        stream.read(reinterpret_cast<char*>(&size), sizeof(size));
        if (!stream || size < 0) throw std::runtime_error("check point file " + filename + " is truncated");
        this->fs[n].resize((size_t)size);
        if (size) stream.read(reinterpret_cast<char*>(this->fs[n].data()), (std::streamsize)(size * sizeof(Real)));
        if (!stream) throw std::runtime_error("check point file " + filename + " is truncated");
    }
}

template class RestartObjectT<float>;
template class RestartObjectT<double>;
template class ASCIIRestartObjectT<float>;
template class ASCIIRestartObjectT<double>;
template class BinaryRestartObjectT<float>;
template class BinaryRestartObjectT<double>;

VFX_HOST_END

// ---------------------------------------------------------------- output ----
VFX_HOST_BEGIN

uint32_t FilePartCalculator::calculateNumberOfNodesInPart(uint32_t numberOfNodes, uint32_t indexOfFilePart)
{
    const uint32_t last = calculateNumberOfParts(numberOfNodes) - 1;
    if (indexOfFilePart > last)
        throw std::runtime_error("The number of nodes for a non-existing part can not be calculated");
    if (indexOfFilePart == last) return numberOfNodes - indexOfFilePart * limitOfNodesForVTK;
    return limitOfNodesForVTK;
}
uint32_t FilePartCalculator::calculateNumberOfParts(uint32_t numberOfNodes)
{
    return numberOfNodes / limitOfNodesForVTK + 1;
}
uint32_t FilePartCalculator::calculateStartingPostionOfPart(uint32_t indexOfPart) { return indexOfPart * limitOfNodesForVTK; }

std::string WriterUtilities::makePartFileNameEnding(uint32_t level, int ID, int part, int timestep)
{
    return "_lev_" + std::to_string((int)level) + "_ID_" + std::to_string(ID) + "_Part_" + std::to_string(part) +
           "_t_" + std::to_string(timestep) + ".vtk";
}

bool WriterUtilities::isPeriodicCell(const HostGrid& g, uint32_t base, uint32_t other)
{
    const double distance = std::sqrt(std::pow(g.coordinateX[other] - g.coordinateX[base], 2.) +
                                      std::pow(g.coordinateY[other] - g.coordinateY[base], 2.) +
                                      std::pow(g.coordinateZ[other] - g.coordinateZ[base], 2.));
    return distance > 1.01 * std::sqrt(3 * std::pow(g.gridSpacing, 2.));
}

void WriterUtilities::getIndicesOfAllNodesInOct(std::array<uint32_t, 8>& n, uint32_t base, const HostGrid& g)
{
    n[0] = base;
    n[1] = g.neighborX[n[0]];
    n[2] = g.neighborY[n[1]];
    n[3] = g.neighborY[n[0]];
    n[4] = g.neighborZ[n[0]];
    n[5] = g.neighborZ[n[1]];
    n[6] = g.neighborZ[n[2]];
    n[7] = g.neighborZ[n[3]];
}

void WriterUtilities::calculateRelativeNodeIndexInPart(std::array<uint32_t, 8>& rel,
                                                       const std::array<uint32_t, 8>& octs, uint32_t start)
{
    for (size_t i = 0; i < rel.size(); i++) rel[i] = octs[i] - start;
}

bool WriterUtilities::areAllNodesInOctValidForWriting(const std::array<uint32_t, 8>& octs, const HostGrid& g,
                                                      uint32_t end)
{
    bool fluid = true, out = false;
    for (uint32_t i : octs) {
        fluid = fluid && g.typeOfGridNode[i] == 19; // GEO_FLUID
        out = out || i > end;
    }
    return fluid && !out;
}

static std::ofstream open_output(const std::string& path)
{
    std::ofstream out(path, std::ios::out | std::ios::binary);
    if (!out) {
        // createFileStream: make the directory and try again
        const std::filesystem::path dir = std::filesystem::path(path).parent_path();
        if (!dir.empty()) {
            std::error_code ec;
            std::filesystem::create_directories(dir, ec);
            out.open(path, std::ios::out | std::ios::binary);
        }
        if (!out) throw std::runtime_error("couldn't open file " + path);
    }
    return out;
}

std::string WbWriterVtkXmlBinary::writeOctsWithNodeData(const std::string& filename,
                                                        const std::vector<std::array<float, 3>>& nodes,
                                                        const std::vector<std::array<uint32_t, 8>>& cells,
                                                        const std::vector<std::string>& names,
                                                        const std::vector<std::vector<double>>& data)
{
    const std::string path = filename + ".bin.vtu";
    std::ofstream out = open_output(path);
    const int n = (int)nodes.size(), m = (int)cells.size();
    const int bPoints = 3 * n * (int)sizeof(float), bConn = 8 * m * (int)sizeof(int), bOff = m * (int)sizeof(int);
    const int bTypes = m, bScalar = n * (int)sizeof(double);
    unsigned long long offset = 0;
    out << "<?xml version=\"2.0\"?>\n";
    out << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" >\n";
    out << "   <UnstructuredGrid>\n";
    out << "      <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << m << "\">\n";
    out << "         <Points>\n";
    out << "            <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << offset
        << "\"  />\n";
    out << "         </Points>\n";
    offset += 4 + bPoints;
    out << "         <Cells>\n";
    out << "            <DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" << offset << "\" />\n";
    offset += 4 + bConn;
    out << "            <DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" << offset << "\" />\n";
    offset += 4 + bOff;
    out << "            <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"" << offset << "\" />\n ";
    offset += 4 + bTypes;
    out << "         </Cells>\n";
    out << "         <PointData>\n";
    for (const std::string& s : names) {
        out << "            <DataArray type=\"Float64\" Name=\"" << s << "\" format=\"appended\" offset=\"" << offset
            << "\" /> \n";
        offset += 4 + bScalar;
    }
    out << "         </PointData>\n";
    out << "      </Piece>\n";
    out << "   </UnstructuredGrid>\n";
    out << "   <AppendedData encoding=\"raw\">\n";
    out << "_";
    out.write((const char*)&bPoints, 4);
    for (const auto& p : nodes) out.write((const char*)p.data(), 3 * sizeof(float));
    out.write((const char*)&bConn, 4);
    static const int order[8] = {0, 1, 3, 2, 4, 5, 7, 6}; // VTK_VOXEL corner order
    for (const auto& c : cells)
        for (int i : order) out.write((const char*)&c[i], 4);
    out.write((const char*)&bOff, 4);
    for (int c = 1; c <= m; c++) {
        const int v = 8 * c;
        out.write((const char*)&v, 4);
    }
    out.write((const char*)&bTypes, 4);
    const unsigned char voxel = 11;
    for (int c = 0; c < m; c++) out.write((const char*)&voxel, 1);
    for (const auto& d : data) {
        out.write((const char*)&bScalar, 4);
        out.write((const char*)d.data(), (std::streamsize)(d.size() * sizeof(double)));
    }
    out << "\n</AppendedData>\n";
    out << "</VTKFile>";
    out << std::endl;
    return path;
}

std::string WbWriterVtkXmlBinary::writeParallelFile(const std::string& filename, const std::vector<std::string>& pieces,
                                                    const std::vector<std::string>& pointNames,
                                                    const std::vector<std::string>& cellNames)
{
    const std::string path = filename + ".pvtu";
    std::ofstream out = open_output(path);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <PUnstructuredGrid GhostLevel=\"1\">\n";
    out << "    <PPointData>\n";
    for (const auto& s : pointNames) out << "      <PDataArray type=\"Float64\" Name=\"" << s << "\"/>\n";
    out << "    </PPointData>\n";
    if (!cellNames.empty()) {
        out << "    <PCellData>\n";
        for (const auto& s : cellNames) out << "      <PDataArray type=\"Float64\" Name=\"" << s << "\"/>\n";
        out << "    </PCellData>\n";
    }
    out << "    <PPoints>\n";
    out << "      <PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n";
    out << "    </PPoints>\n";
    for (const auto& s : pieces) out << "    <Piece Source=\"" << s << "\"/>\n";
    out << "  </PUnstructuredGrid>\n";
    out << "</VTKFile>";
    out << std::endl;
    return path;
}

std::string WbWriterVtkXmlBinary::writeCollectionForTimeSeries(const std::string& filename,
                                                               const std::map<uint32_t, std::vector<std::string>>& files,
                                                               bool separateGroups)
{
    const std::string path = filename + ".pvd";
    std::ofstream out = open_output(path);
    out << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\" >" << std::endl;
    out << "   <Collection>" << std::endl;
    for (const auto& [step, names] : files) {
        int group = 0, part = 0;
        for (const auto& f : names) {
            out << "       <DataSet timestep=\"" << (double)step << "\" group=\"" << group << "\" part=\"" << part
                << "\" file=\"" << f << "\"/>" << std::endl;
            if (separateGroups) group++;
            else part++;
        }
    }
    out << "   </Collection>\n</VTKFile>";
    return path;
}

std::vector<std::string> FileWriter::getNodeDataNames(const Parameter& para) const
{
    std::vector<std::string> names = {"press", "rho", "vx1", "vx2", "vx3", "geo"};
    if (para.getUseTurbulentViscosity()) names.push_back("nut");
    return names;
}

std::vector<std::string> FileWriter::writeUnstructuredGridLT(const HostGrid& g, const std::vector<std::string>& names,
                                                             const std::vector<std::vector<double>>& nodeData,
                                                             int level, unsigned int timestep)
{
    const uint32_t numberOfNodes = (uint32_t)g.typeOfGridNode.size();
    const uint32_t parts = FilePartCalculator::calculateNumberOfParts(numberOfNodes);
    std::vector<std::array<float, 3>> nodes;
    std::vector<std::array<uint32_t, 8>> cells;
    std::vector<std::vector<double>> data(names.size());
    std::vector<std::string> outNames;
    for (uint32_t part = 0; part < parts; part++) {
        const uint32_t size = FilePartCalculator::calculateNumberOfNodesInPart(numberOfNodes, part);
        const uint32_t start = FilePartCalculator::calculateStartingPostionOfPart(part);
        const uint32_t end = start + size;
        cells.clear();
        nodes.resize(size); // keeps what a previous part left, as the reference
        for (auto& d : data) d.resize(size);
        for (uint32_t pos = start; pos < end; pos++) {
            if (g.typeOfGridNode[pos] != 19) continue; // GEO_FLUID
            const uint32_t rel = pos - start;
            nodes[rel] = {(float)g.coordinateX[pos], (float)g.coordinateY[pos], (float)g.coordinateZ[pos]};
            for (size_t i = 0; i < data.size(); ++i) data[i][rel] = nodeData[i][pos];
            std::array<uint32_t, 8> octs, relPos;
            WriterUtilities::getIndicesOfAllNodesInOct(octs, pos, g);
            if (WriterUtilities::isPeriodicCell(g, octs[0], octs[6])) continue;
            if (WriterUtilities::areAllNodesInOctValidForWriting(octs, g, end)) {
                WriterUtilities::calculateRelativeNodeIndexInPart(relPos, octs, start);
                cells.push_back(relPos);
            }
        }
        const std::string fname =
            prefix + "_bin" + WriterUtilities::makePartFileNameEnding((uint32_t)level, ID, (int)part + 1, (int)timestep);
        outNames.push_back(WbWriterVtkXmlBinary::writeOctsWithNodeData(fname, nodes, cells, names, data));
    }
    return outNames;
}

// cudaCopyPrint + writeUnstructuredGridLT of one level
void FileWriter::writeLevel(const std::shared_ptr<Parameter>& para, unsigned int timestep, int level)
{
    LBMSimulationParameter* p = para->getParD(level);
    const HostGrid& g = p->host;
    const uint64_t n = p->lattice.n_nodes;
    if (g.typeOfGridNode.size() != n || g.coordinateX.size() != n || g.neighborX.size() != n)
        throw std::runtime_error("FileWriter: the level's host grid does not match its lattice");
    const size_t rbytes = p->lattice.storage == VFX_FP64 ? 8 : 4;
    void* dev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<std::vector<double>> fields(5, std::vector<double>(n));
    try {
        for (auto& d : dev) check(vfx_device_alloc(n * rbytes, &d), "FileWriter alloc");
        check(vfx_macroscopic_quantities(&p->lattice, p->isEvenTimestep ? 1 : 0, dev[0], dev[1], dev[2], dev[3], dev[4],
                                         p->stream),
              "calculateMacroscopicQuantities");
        check(vfx_synchronize(p->stream), "FileWriter sync");
        std::vector<char> buf(n * rbytes);
        for (int i = 0; i < 5; ++i) {
            check(vfx_copy_to_host(buf.data(), dev[i], n * rbytes), "FileWriter d2h");
            for (uint64_t k = 0; k < n; ++k)
                fields[i][k] = rbytes == 8 ? ((const double*)buf.data())[k] : (double)((const float*)buf.data())[k];
        }
    } catch (...) {
        for (void* d : dev) vfx_device_free(d);
        throw;
    }
    for (void* d : dev) vfx_device_free(d);
    // FileWriter.cpp:268-287: press, rho, vx1, vx2, vx3, geo [, nut]
    std::vector<std::vector<double>> data;
    for (int i : {4, 0}) {
        std::vector<double> v(n);
        for (uint64_t k = 0; k < n; ++k)
            v[k] = fields[i][k] / (double)3.0 * (double)(real)densityRatio * (double)(real)velocityRatio *
                   (double)(real)velocityRatio;
        data.push_back(std::move(v));
    }
    for (int i : {1, 2, 3}) {
        std::vector<double> v(n);
        for (uint64_t k = 0; k < n; ++k) v[k] = fields[i][k] * (double)(real)velocityRatio;
        data.push_back(std::move(v));
    }
    std::vector<double> geo(n);
    for (uint64_t k = 0; k < n; ++k) geo[k] = (double)g.typeOfGridNode[k];
    data.push_back(std::move(geo));
    if (para->getUseTurbulentViscosity()) {
        if (!p->lattice.turbulent_viscosity) throw std::runtime_error("FileWriter: no turbulent viscosity array");
        std::vector<char> buf(n * rbytes);
        check(vfx_copy_to_host(buf.data(), p->lattice.turbulent_viscosity, n * rbytes), "FileWriter d2h");
        std::vector<double> v(n);
        for (uint64_t k = 0; k < n; ++k)
            v[k] = (rbytes == 8 ? ((const double*)buf.data())[k] : (double)((const float*)buf.data())[k]) *
                   (double)(real)viscosityRatio;
        data.push_back(std::move(v));
    }
    for (const std::string& f : writeUnstructuredGridLT(g, getNodeDataNames(*para), data, level, timestep))
        fileNamesForCollectionFile.push_back(f.substr(f.find_last_of('/') + 1));
}

std::vector<std::string> FileWriter::getMeanNodeDataNames()
{
    return {"pressMed", "rhoMed", "vx1Med", "vx2Med", "vx3Med", "geo"};
}

std::vector<std::string> FileWriter::writeUnstructuredGridMeanLT(const HostGrid& g, const LBMSimulationParameter& p,
                                                                 int level, unsigned int timestep)
{
    const size_t n = g.typeOfGridNode.size();
    if (p.meanPressureOut.size() != n || p.meanDensityOut.size() != n || p.meanVelocityInXdirectionOut.size() != n ||
        p.meanVelocityInYdirectionOut.size() != n || p.meanVelocityInZdirectionOut.size() != n)
        throw std::runtime_error("FileWriter: the level's means do not match its grid (calcMean first)");
    // meanPressureOut / 3.0f * densityRatio * velocityRatio * velocityRatio, all real
    const real dr = (real)densityRatio, vr = (real)velocityRatio;
    std::vector<std::vector<double>> data(6, std::vector<double>(n));
    for (size_t k = 0; k < n; ++k) {
        data[0][k] = p.meanPressureOut[k] / 3.0f * dr * vr * vr;
        data[1][k] = p.meanDensityOut[k] / 3.0f * dr * vr * vr;
        data[2][k] = p.meanVelocityInXdirectionOut[k] * vr;
        data[3][k] = p.meanVelocityInYdirectionOut[k] * vr;
        data[4][k] = p.meanVelocityInZdirectionOut[k] * vr;
        data[5][k] = (double)g.typeOfGridNode[k];
    }
    // the cells of writeUnstructuredGridMeanLT are those of writeUnstructuredGridLT
    // (same oct, fluid, bound and periodic tests); only the file name differs
    const std::vector<std::string> names = getMeanNodeDataNames();
    const uint32_t numberOfNodes = (uint32_t)n;
    const uint32_t parts = FilePartCalculator::calculateNumberOfParts(numberOfNodes);
    std::vector<std::array<float, 3>> nodes;
    std::vector<std::array<uint32_t, 8>> cells;
    std::vector<std::vector<double>> part(names.size());
    std::vector<std::string> outNames;
    for (uint32_t pi = 0; pi < parts; pi++) {
        const uint32_t size = FilePartCalculator::calculateNumberOfNodesInPart(numberOfNodes, pi);
        const uint32_t start = FilePartCalculator::calculateStartingPostionOfPart(pi);
        const uint32_t end = start + size;
        cells.clear();
        nodes.resize(size);
        for (auto& d : part) d.resize(size);
        for (uint32_t pos = start; pos < end; pos++) {
            if (g.typeOfGridNode[pos] != 19) continue; // GEO_FLUID
            const uint32_t rel = pos - start;
            nodes[rel] = {(float)g.coordinateX[pos], (float)g.coordinateY[pos], (float)g.coordinateZ[pos]};
            for (size_t i = 0; i < part.size(); ++i) part[i][rel] = data[i][pos];
            std::array<uint32_t, 8> octs, relPos;
            WriterUtilities::getIndicesOfAllNodesInOct(octs, pos, g);
            if (WriterUtilities::isPeriodicCell(g, octs[0], octs[6])) continue;
            if (WriterUtilities::areAllNodesInOctValidForWriting(octs, g, end)) {
                WriterUtilities::calculateRelativeNodeIndexInPart(relPos, octs, start);
                cells.push_back(relPos);
            }
        }
        const std::string fname = prefix + "_bin_mean" +
                                  WriterUtilities::makePartFileNameEnding((uint32_t)level, ID, (int)pi + 1, (int)timestep);
        outNames.push_back(WbWriterVtkXmlBinary::writeOctsWithNodeData(fname, nodes, cells, names, part));
    }
    return outNames;
}

void FileWriter::writeInit(const std::shared_ptr<Parameter>& para, unsigned int timestep)
{
    writeTimestep(para, timestep, false);
}

void FileWriter::writeTimestep(const std::shared_ptr<Parameter>& para, unsigned int timestep, bool lastTimestep)
{
    for (int level = 0; level <= para->getMaxLevel(); level++) {
        writeLevel(para, timestep, level);
        if (para->getCalcMean()) {
            LBMSimulationParameter* p = para->getParD(level);
            for (const std::string& f : writeUnstructuredGridMeanLT(p->host, *p, level, timestep))
                fileNamesForCollectionFileMean.push_back(f.substr(f.find_last_of('/') + 1));
        }
    }
    fileNamesForCollectionFileTimeSeries[timestep] = fileNamesForCollectionFile;
    if (lastTimestep)
        WbWriterVtkXmlBinary::writeCollectionForTimeSeries(prefix + "_bin_ID_" + std::to_string(ID),
                                                           fileNamesForCollectionFileTimeSeries, false);
    const std::string ending = "_ID_" + std::to_string(ID) + "_t_" + std::to_string(timestep) + ".vtk";
    WbWriterVtkXmlBinary::writeParallelFile(prefix + "_bin" + ending, fileNamesForCollectionFile,
                                            getNodeDataNames(*para), {});
    fileNamesForCollectionFile.clear();
    if (para->getCalcMean()) {
        WbWriterVtkXmlBinary::writeParallelFile(prefix + "_bin_mean" + ending, fileNamesForCollectionFileMean,
                                                getMeanNodeDataNames(), {});
        fileNamesForCollectionFileMean.clear();
    }
}

// ------------------------------------------------------------- time means --
static std::array<void**, 5> mean_sums(LBMSimulationParameter* p)
{
    return {&p->meanVelocityInXdirection, &p->meanVelocityInYdirection, &p->meanVelocityInZdirection, &p->meanDensity,
            &p->meanPressure};
}

void allocMean(Parameter* para)
{
    for (int level = 0; level <= para->getMaxLevel(); level++) {
        LBMSimulationParameter* p = para->getParD(level);
        check_real(p->lattice, "mean values");
        const auto s = mean_sums(p);
        for (int i = 0; i < 5; ++i)
            if (!*s[i]) check(vfx_device_alloc(p->lattice.n_nodes * sizeof(real), s[i]), "allocMean");
        check(vfx_reset_mean(&p->lattice, p->meanVelocityInXdirection, p->meanVelocityInYdirection,
                             p->meanVelocityInZdirection, p->meanDensity, p->meanPressure, p->stream),
              "allocMean reset");
        for (auto* v : {&p->meanVelocityInXdirectionOut, &p->meanVelocityInYdirectionOut, &p->meanVelocityInZdirectionOut,
                        &p->meanDensityOut, &p->meanPressureOut})
            v->assign(p->lattice.n_nodes, (real)0.0);
    }
}

void calculateMean(Parameter* para, int level)
{
    LBMSimulationParameter* p = para->getParD(level);
    if (!p->meanPressure) throw std::runtime_error("calculateMean: allocMean first");
    check(vfx_calculate_mean(&p->lattice, p->isEvenTimestep ? 1 : 0, p->meanVelocityInXdirection,
                             p->meanVelocityInYdirection, p->meanVelocityInZdirection, p->meanDensity, p->meanPressure,
                             p->stream),
          "calculateMeanCompressible");
}

void calcMean(Parameter* para, uint tdiff)
{
    for (int level = 0; level <= para->getMaxLevel(); level++) {
        LBMSimulationParameter* p = para->getParD(level);
        if (!p->meanPressure) throw std::runtime_error("calcMean: allocMean first");
        check(vfx_synchronize(p->stream), "calcMean sync");
        const uint64_t n = p->lattice.n_nodes;
        std::vector<real> sum(n);
        const auto s = mean_sums(p);
        std::vector<real>* out[5] = {&p->meanVelocityInXdirectionOut, &p->meanVelocityInYdirectionOut,
                                     &p->meanVelocityInZdirectionOut, &p->meanDensityOut, &p->meanPressureOut};
        for (int i = 0; i < 5; ++i) {
            check(vfx_copy_to_host(sum.data(), *s[i], n * sizeof(real)), "calcMean d2h");
            out[i]->resize(n);
            for (uint64_t k = 0; k < n; ++k) (*out[i])[k] = sum[k] / (real)tdiff;
        }
    }
}

void resetMean(Parameter* para)
{
    for (int level = 0; level <= para->getMaxLevel(); level++) {
        LBMSimulationParameter* p = para->getParD(level);
        if (!p->meanPressure) continue;
        check(vfx_reset_mean(&p->lattice, p->meanVelocityInXdirection, p->meanVelocityInYdirection,
                             p->meanVelocityInZdirection, p->meanDensity, p->meanPressure, p->stream),
              "ResetMeanValuesSP27");
    }
}

void freeMean(Parameter* para)
{
    for (int level = 0; level <= para->getMaxLevel(); level++) {
        const auto s = mean_sums(para->getParD(level));
        for (int i = 0; i < 5; ++i) {
            vfx_device_free(*s[i]);
            *s[i] = nullptr;
        }
    }
}

// ---------------------------------------------------------------- precursor files --
namespace {

std::string readElement(const std::string& line)
{
    const size_t start = line.find('<') + 1;
    const size_t end = line.find(' ', start);
    return end == std::string::npos ? line.substr(start) : line.substr(start, end - start);
}

std::string readAttribute(const std::string& line, const std::string& name)
{
    const size_t start = line.find(name) + name.size() + 2; // skip '="'
    return line.substr(start, line.find('"', start) - start);
}

template <typename T> std::vector<T> readStringToVector(const std::string& s)
{
    // every number read as float, then converted (TransientBCSetter.cpp:71-82)
    std::vector<T> out;
    std::stringstream input(s);
    float num;
    while (input >> num) out.push_back((T)num);
    return out;
}

} // namespace

void VTKFile::readHeader()
{
    std::ifstream file(fileName);
    if (!file) throw std::runtime_error("VTKFile: cannot open " + fileName);
    std::string line;
    std::getline(file, line); // VTKFile or the xml declaration
    if (line.size() > 1 && line[1] == '?') std::getline(file, line);
    std::getline(file, line); // ImageData
    const std::vector<float> origin = readStringToVector<float>(readAttribute(line, "Origin"));
    const std::vector<float> spacing = readStringToVector<float>(readAttribute(line, "Spacing"));
    std::getline(file, line); // Piece
    const std::vector<int> extent = readStringToVector<int>(readAttribute(line, "Extent"));
    std::getline(file, line); // PointData
    std::getline(file, line);
    while (readElement(line) == "DataArray") {
        Quantity q;
        q.name = readAttribute(line, "Name");
        q.offset = std::stoi(readAttribute(line, "offset"));
        quantities.push_back(q);
        std::getline(file, line);
    }
    std::getline(file, line); // </Piece>
    std::getline(file, line); // </ImageData>
    std::getline(file, line); // AppendedData
    const int base = int(file.tellg()) + 1 + 4; // the underscore and the first array's byte count
    for (auto& q : quantities) q.offset += base;
    if (origin.size() < 3 || spacing.size() < 3 || extent.size() < 6)
        throw std::runtime_error("VTKFile: malformed header in " + fileName);
    deltaX = spacing[0];
    deltaY = spacing[1];
    deltaZ = spacing[2];
    nx = extent[1] - extent[0] + 1;
    ny = extent[3] - extent[2] + 1;
    nz = extent[5] - extent[4] + 1;
    minX = origin[0] + deltaX * extent[0];
    maxX = (nx - 1) * deltaX + minX;
    minY = origin[1] + deltaY * extent[2];
    maxY = (ny - 1) * deltaY + minY;
    minZ = origin[2] + deltaZ * extent[4];
    maxZ = (nz - 1) * deltaZ + minZ;
}

int VTKFile::getClosestIdxZ(real z) const
{
    const int i = (int)std::round((z - minZ) / deltaZ);
    return i > nz ? nz : (i < 0 ? 0 : i);
}

bool VTKFile::markNANs(const std::vector<uint>& readIndices) const
{
    // the first record's values at the read indices (the reference reads into
    // a reserved, empty vector, so its check finds nothing)
    std::ifstream buf(fileName, std::ios::in | std::ios::binary);
    std::vector<double> first(std::max(1, getNumberOfPointsInXYPlane()));
    buf.seekg(quantities.at(0).offset);
    buf.read(reinterpret_cast<char*>(first.data()), (std::streamsize)(first.size() * sizeof(double)));
    for (uint i : readIndices)
        if (i < first.size() && std::isnan(first[i])) return true;
    return false;
}

void VTKFile::loadFile()
{
    std::ifstream buf(fileName, std::ios::in | std::ios::binary);
    for (auto& q : quantities) {
        q.values.resize(getNumberOfPoints());
        buf.seekg(q.offset);
        buf.read(reinterpret_cast<char*>(q.values.data()), (std::streamsize)(getNumberOfPoints() * sizeof(double)));
    }
    if (!buf) throw std::runtime_error("VTKFile: short file " + fileName);
    loaded = true;
}

void VTKFile::unloadFile()
{
    for (auto& q : quantities) std::vector<double>().swap(q.values);
    loaded = false;
}

void VTKFile::getData(real* data, uint numberOfNodes, const std::vector<uint>& readIndices,
                      const std::vector<uint>& writeIndices, uint offsetRead, uint offsetWrite)
{
    if (!loaded) loadFile();
    for (size_t j = 0; j < quantities.size(); j++) {
        real* quant = &data[j * numberOfNodes];
        for (size_t i = 0; i < writeIndices.size(); i++)
            quant[offsetWrite + writeIndices[i]] = (real)quantities[j].values.at(readIndices[i] + offsetRead);
    }
}

std::string VTKFileCollection::makeFileName(int level, int id, int part) const
{
    return prefix + "_lev_" + std::to_string(level) + "_ID_" + std::to_string(id) + "_File_" + std::to_string(part) +
           ".bin." + suffix;
}

void VTKFileCollection::findFiles()
{
    while (true) {
        std::vector<std::vector<VTKFile>> level;
        while (true) {
            std::vector<VTKFile> parts;
            while (true) {
                const std::string name = makeFileName((int)files.size(), (int)level.size(), (int)parts.size());
                if (!std::ifstream(name).good()) break;
                parts.emplace_back(name);
            }
            if (parts.empty()) break;
            level.push_back(std::move(parts));
        }
        if (level.empty()) break;
        files.push_back(std::move(level));
    }
    if (files.empty()) throw std::runtime_error("VTKFileCollection found no files!");
}

void TransientBCInputFileReader::getNeighbors(uint* n0PP, uint* n0PM, uint* n0MP, uint* n0MM)
{
    std::copy(planeNeighbor0PP.begin(), planeNeighbor0PP.end(), &n0PP[writingOffset]);
    std::copy(planeNeighbor0PM.begin(), planeNeighbor0PM.end(), &n0PM[writingOffset]);
    std::copy(planeNeighbor0MP.begin(), planeNeighbor0MP.end(), &n0MP[writingOffset]);
    std::copy(planeNeighbor0MM.begin(), planeNeighbor0MM.end(), &n0MM[writingOffset]);
}

void TransientBCInputFileReader::getWeights(real* w0PP, real* w0PM, real* w0MP, real* w0MM)
{
    std::copy(weights0PP.begin(), weights0PP.end(), &w0PP[writingOffset]);
    std::copy(weights0PM.begin(), weights0PM.end(), &w0PM[writingOffset]);
    std::copy(weights0MP.begin(), weights0MP.end(), &w0MP[writingOffset]);
    std::copy(weights0MM.begin(), weights0MM.end(), &w0MM[writingOffset]);
}

VTKReader::VTKReader(std::shared_ptr<VTKFileCollection> collection, uint level)
    : fileCollection(std::move(collection)), readLevel(level)
{
    nQuantities = fileCollection->getNumberOfQuantities();
}

void VTKReader::initializeIndexVectors()
{
    const size_t nl = fileCollection->files.size();
    readIndices.assign(nl, {});
    writeIndices.assign(nl, {});
    nFile.assign(nl, {});
    for (size_t l = 0; l < nl; l++) {
        readIndices[l].resize(fileCollection->files[l].size());
        writeIndices[l].resize(fileCollection->files[l].size());
        nFile[l].assign(fileCollection->files[l].size(), 0);
    }
}

uint VTKReader::getWriteIndex(int level, int id, int linearIndex)
{
    std::vector<uint>& r = readIndices[level][id];
    const auto it = std::find(r.begin(), r.end(), (uint)linearIndex);
    if (it != r.end()) return writeIndices[level][id][it - r.begin()];
    writeIndices[level][id].push_back(nPointsRead);
    r.push_back((uint)linearIndex);
    return nPointsRead++;
}

void VTKReader::fillArrays(std::vector<real>& coordsY, std::vector<real>& coordsZ)
{
    nPoints = (uint)coordsY.size();
    initializeIndexVectors();
    const real maxDiff = (real)1e-4, eps = (real)1e-7;
    const uint level = readLevel;
    if (level >= fileCollection->files.size()) throw std::runtime_error("VTKReader: no files on the read level");
    auto weight = [&](const VTKFile& f, int idx, real y, real z) {
        const real dy = f.getX(idx) - y;
        const real dz = f.getY(idx) - z;
        return (real)1 / (dy * dy + dz * dz + eps);
    };
    for (uint i = 0; i < nPoints; i++) {
        const real y = coordsY[i], z = coordsZ[i];
        bool f0PP = false, f0PM = false, f0MP = false, f0MM = false, all = false;
        for (int id = 0; id < (int)fileCollection->files[level].size(); id++) {
            VTKFile& f = fileCollection->files[level][id][0];
            if (!f.inBoundingBox(y, z, (real)0)) continue;
            const int idx = f.findNeighborMMM(y, z, (real)0);
            if (idx != -1) {
                if (std::abs(y - f.getX(idx)) < maxDiff && std::abs(z - f.getY(idx)) < maxDiff) {
                    weights0PP.push_back((real)1e6f);
                    weights0PM.push_back(0);
                    weights0MP.push_back(0);
                    weights0MM.push_back(0);
                    const uint w = getWriteIndex(level, id, idx);
                    planeNeighbor0PP.push_back(w);
                    planeNeighbor0PM.push_back(w);
                    planeNeighbor0MP.push_back(w);
                    planeNeighbor0MM.push_back(w);
                    f0PP = f0PM = f0MP = f0MM = true;
                }
                if (!f0MM) {
                    f0MM = true;
                    weights0MM.push_back(weight(f, idx, y, z));
                    planeNeighbor0MM.push_back(getWriteIndex(level, id, idx));
                }
            }
            if (!f0PP) {
                const int n = f.findNeighborPPM(y, z, (real)0);
                if (n != -1) {
                    f0PP = true;
                    weights0PP.push_back(weight(f, n, y, z));
                    planeNeighbor0PP.push_back(getWriteIndex(level, id, n));
                }
            }
            if (!f0PM) {
                const int n = f.findNeighborPMM(y, z, (real)0);
                if (n != -1) {
                    f0PM = true;
                    weights0PM.push_back(weight(f, n, y, z));
                    planeNeighbor0PM.push_back(getWriteIndex(level, id, n));
                }
            }
            if (!f0MP) {
                const int n = f.findNeighborMPM(y, z, (real)0);
                if (n != -1) {
                    f0MP = true;
                    weights0MP.push_back(weight(f, n, y, z));
                    planeNeighbor0MP.push_back(getWriteIndex(level, id, n));
                }
            }
            all = f0PP && f0PM && f0MP && f0MM;
            if (all) break;
        }
        if (!all)
            throw std::runtime_error("VTKReader::fillArrays(): Did not find neighbors in the FileCollection for all points");
    }
    for (size_t l = 0; l < fileCollection->files.size(); l++)
        for (size_t id = 0; id < fileCollection->files[l].size(); id++)
            if (fileCollection->files[l][id][0].markNANs(readIndices[l][id]))
                throw std::runtime_error("Found a NAN in the precursor where a velocity is needed");
}

void VTKReader::getNextData(real* data, uint numberOfNodes, real time)
{
    const uint level = readLevel;
    for (size_t id = 0; id < fileCollection->files[level].size(); id++) {
        auto& parts = fileCollection->files[level][id];
        size_t k = nFile[level][id];
        if (!parts[k].inZBounds(time)) {
            k++;
            if (k == parts.size()) throw std::runtime_error("Not enough Precursor Files to read");
            parts[k - 1].unloadFile();
        }
        VTKFile& f = parts[k];
        const int off = f.getClosestIdxZ(time) * f.getNumberOfPointsInXYPlane();
        f.getData(data, numberOfNodes, readIndices[level][id], writeIndices[level][id], (uint)off, writingOffset);
        nFile[level][id] = k;
    }
}

std::shared_ptr<FileCollection> createFileCollection(std::string prefix, TransientBCFileType type)
{
    if (type == TransientBCFileType::VTK) return std::make_shared<VTKFileCollection>(std::move(prefix));
    throw std::runtime_error("createFileCollection: Unknown file type");
}

std::shared_ptr<TransientBCInputFileReader> createReaderForCollection(std::shared_ptr<FileCollection> collection,
                                                                      uint readLevel)
{
    if (collection->getFileType() == TransientBCFileType::VTK)
        return std::make_shared<VTKReader>(std::static_pointer_cast<VTKFileCollection>(collection), readLevel);
    throw std::runtime_error("createReaderForCollection: no reader for this file type");
}

void setPrecursorFileReader(QforPrecursorBoundaryConditions& q, std::shared_ptr<TransientBCInputFileReader> reader)
{
    q.numberOfPrecursorNodes = (int)reader->getNPointsRead();
    q.numberOfQuantities = reader->getNumberOfQuantities();
    q.getNextData = [reader](real* data, int numberOfNodes, double time) {
        reader->getNextData(data, (uint)numberOfNodes, (real)time);
    };
}

VFX_HOST_END
