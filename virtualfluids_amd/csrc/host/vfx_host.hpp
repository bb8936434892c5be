// vfx_host.hpp -- C++ host layer over the libvfx C ABI.
//
// Mirrors the plug-in interface VirtualFluids' GPU solver uses for this path,
// so a caller written against it finds the same names and behaviour:
//   Kernel / KernelImp           src/gpu/core/Kernel/Kernel.h:44-52, KernelImp.h:45-65
//   KernelFactory(Imp)           src/gpu/core/Kernel/KernelFactory/KernelFactory.h:47-55,
//                                KernelFactoryImp.cpp:61-135 (string keys of KernelTypes.h:35-46)
//   BoundaryConditionFactory     src/gpu/core/BoundaryConditions/BoundaryConditionFactory.h:50-174
//   UpdateGrid27 step order      src/gpu/core/Calculation/UpdateGrid27.cpp:54-95, :227-307
//   TurbulenceModelFactory       src/gpu/core/TurbulenceModels/TurbulenceModelFactory.cpp:44-100
// Every device operation goes through include/vfx.h; errors become
// std::runtime_error (the reference aborts the process instead).
#pragma once

#include "../../../include/vfx.h"

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

// VirtualFluids' `real` (src/basics/DataTypes.h): float, or double when built
// with VF_DOUBLE_ACCURACY.  The host layer is compiled once for each (both are
// in libvfx.so); the inline namespace keeps the two builds' symbols apart, so
// a caller sees the structs of the `real` it was compiled with.  The lattice
// storage must match: VFX_FP64 for double, VFX_FP32 or VFX_FP16_SCALED for float.
#ifdef VF_DOUBLE_ACCURACY
#define VFX_HOST_BEGIN namespace vfx::host { inline namespace real_double {
#define VFX_HOST_REAL double
#else
#define VFX_HOST_BEGIN namespace vfx::host { inline namespace real_float {
#define VFX_HOST_REAL float
#endif
#define VFX_HOST_END } }

VFX_HOST_BEGIN

using real = VFX_HOST_REAL;
using uint = unsigned int;

// src/gpu/core/Calculation/Calculation.h:55-66
enum class CollisionTemplate { Default, WriteMacroVars, ApplyBodyForce, AllFeatures, SubDomainBorder };
// src/gpu/core/Cuda/CudaStreamManager.h:41-46
enum class CudaStreamIndex { Legacy, Bulk, SubDomainBorder, Precursor };
enum class PreProcessorType { InitNavierStokesCompressible };
// vf::lbm::TurbulenceModel (src/lbm/collision/TurbulentViscosity.h:50-61)
enum class TurbulenceModel { Smagorinsky, AMD, QR, None };

// Calculation.h:105-117, field for field.  q27[d] points at the SoA block of
// direction d (q27[d] = q27[0] + d * numberOfBCnodes, as the grid provider sets
// them); the kernels receive q27[0].  The wall models read Vx/Vy/Vz (filtered
// velocity at the sampling node) and Vx1/Vy1/Vz1 (at the wall node), and the
// wall normals normalX/Y/Z; kN is the sampling (stress) or inward (pressure)
// neighbour.
struct QforBoundaryConditions {
    int* k = nullptr;
    int* kN = nullptr;
    real* q27[27] = {};
    real* q19[19] = {};
    unsigned int numberOfBCnodes = 0;
    int kArray = 0;
    real *Vx = nullptr, *Vy = nullptr, *Vz = nullptr;
    real *Vx1 = nullptr, *Vy1 = nullptr, *Vz1 = nullptr;
    real *deltaVz = nullptr, *RhoBC = nullptr;
    real *normalX = nullptr, *normalY = nullptr, *normalZ = nullptr;
};

// Calculation.h:119-127: the pressure and outflow boundaries, one struct per
// outward direction (LBMSimulationParameter::pressureBCDirectional)
struct QforDirectionalBoundaryCondition {
    int* k = nullptr;
    int* kN = nullptr;
    real* q27[27] = {};
    unsigned int numberOfBCnodes = 0;
    real* RhoBC = nullptr;
    size_t direction = 0;
};

// Calculation.h:129-145.  getNextData stands for
// TransientBCInputFileReader::getNextData (TransientBCSetter.h:174-215): it
// fills the plane array `next` (numberOfPrecursorNodes * numberOfQuantities
// values; host memory, as the reference's reader fills planeData before
// cudaCopyPrecursorData) with the plane at the physical time loadTime; the
// host layer copies it to the device array next.
struct QforPrecursorBoundaryConditions {
    int* k = nullptr;
    int numberOfBCnodes = 0;
    int sizeQ = 0;
    int numberOfPrecursorNodes = 0;
    uint streamIndex = 0;
    uint nPrecursorReads = 0;
    uint timeStepsBetweenReads = 1;
    size_t numberOfQuantities = 0;
    real* q27[27] = {};
    uint *planeNeighbor0PP = nullptr, *planeNeighbor0PM = nullptr, *planeNeighbor0MP = nullptr,
         *planeNeighbor0MM = nullptr;
    real *weights0PP = nullptr, *weights0PM = nullptr, *weights0MP = nullptr, *weights0MM = nullptr;
    real *last = nullptr, *current = nullptr, *next = nullptr;
    real velocityX = 0, velocityY = 0, velocityZ = 0;
    std::function<void(real* planeData, int numberOfPrecursorNodes, double loadTime)> getNextData;
};

// Calculation.h:162-171
struct WallModelParameters {
    real* z0 = nullptr;
    int* samplingOffset = nullptr;
    bool hasMonitor = false;
    real* u_star = nullptr;
    real* Fx = nullptr;
    real* Fy = nullptr;
    real* Fz = nullptr;
};

// Calculation.h:72-89: interface cells between a level and the next finer one
struct ICells {
    uint* fineCellIndices = nullptr;
    uint* coarseCellIndices = nullptr;
    uint numberOfCells = 0;
};
struct ICellNeigh {
    real* x = nullptr;
    real* y = nullptr;
    real* z = nullptr;
};

// Host copies of a level's grid the output needs (LBMSimulationParameter's
// coordinateX/Y/Z, typeOfGridNode, neighborX/Y/Z, gridSpacing on the host side)
struct HostGrid {
    std::vector<double> coordinateX, coordinateY, coordinateZ;
    std::vector<uint32_t> typeOfGridNode, neighborX, neighborY, neighborZ;
    double gridSpacing = 1.0;
};

// The per-level data of src/gpu/core/Parameter/Parameter.h:66-255 this path uses.
struct LBMSimulationParameter {
    vfx_lattice lattice{};      // distributions, addressing, omega, limiters, force
    bool isEvenTimestep = true;
    // interface to the next finer level (held by the coarser of the two)
    ICells fineToCoarse, coarseToFine;
    ICellNeigh neighborFineToCoarse, neighborCoarseToFine;
    // fluid node lists per collision template (Parameter::taggedFluidNodeIndices)
    std::map<CollisionTemplate, std::pair<uint32_t*, uint32_t>> taggedFluidNodeIndices;
    std::vector<CollisionTemplate> allocatedBulkFluidNodeTags{CollisionTemplate::Default};
    // Parameter.h:112-138
    QforBoundaryConditions velocityBC, noSlipBC, slipBC, stressBC;
    QforPrecursorBoundaryConditions precursorBC;
    std::vector<QforDirectionalBoundaryCondition> pressureBCDirectional;
    WallModelParameters wallModel;                // Parameter.h:152
    double outflowPressureCorrectionFactor = 0.0; // Parameter.h:158
    HostGrid host;                                // for the output (FileWriter)
    vfx_stream stream = nullptr;
    // Parameter.h:324-329: time-average sums on the device (real, numberOfNodes
    // each; allocMean) and the means of the last interval on the host (calcMean)
    void *meanVelocityInXdirection = nullptr, *meanVelocityInYdirection = nullptr,
         *meanVelocityInZdirection = nullptr, *meanDensity = nullptr, *meanPressure = nullptr;
    std::vector<real> meanVelocityInXdirectionOut, meanVelocityInYdirectionOut, meanVelocityInZdirectionOut,
        meanDensityOut, meanPressureOut;
};

class Parameter {
public:
    explicit Parameter(int maxLevel = 0) : levels(maxLevel + 1) {}
    int getMaxLevel() const { return (int)levels.size() - 1; }
    LBMSimulationParameter* getParD(int level) { return &levels.at(level); }
    void setMainKernel(const std::string& k) { mainKernel = k; }
    const std::string& getMainKernel() const { return mainKernel; }
    void setQuadricLimiters(double p, double m, double d) { limiters[0] = p; limiters[1] = m; limiters[2] = d; }
    const double* getQuadricLimiters() const { return limiters; }
    // Parameter.h:486-491, 773-779: the eddy-viscosity model of the K17 kernel;
    // each level's lattice then needs turbulent_viscosity (and, for AMD,
    // rho/vx/vy/vz; on the sparse layout also type_of_grid_node and
    // neighbor_inverse)
    void setTurbulenceModel(TurbulenceModel m) { turbulenceModel = m; }
    TurbulenceModel getTurbulenceModel() const { return turbulenceModel; }
    void setSGSConstant(double c) { sgsConstant = c; }
    double getSGSConstant() const { return sgsConstant; }
    void setUseTurbulentViscosity(bool u) { useTurbulentViscosity = u; }
    // Parameter::setOutflowPressureCorrectionFactor (Parameter.cpp:669-672), every level
    void setOutflowPressureCorrectionFactor(double f)
    {
        for (auto& l : levels) l.outflowPressureCorrectionFactor = f;
    }
    bool getUseTurbulentViscosity() const { return useTurbulentViscosity; }
    // Parameter.cpp:141-144, 526-532: time averages between TimeStartCalcMean and TimeEndCalcMean
    void setCalcMean(bool c) { calcMean = c; }
    bool getCalcMean() const { return calcMean; }
    void setTimeCalcMedStart(int t) { tCalcMedStart = t; }
    int getTimeCalcMedStart() const { return tCalcMedStart; }
    void setTimeCalcMedEnd(int t) { tCalcMedEnd = t; }
    int getTimeCalcMedEnd() const { return tCalcMedEnd; }
    // Parameter::getVelocityRatio / getTimeRatio (physical per lattice units)
    void setVelocityRatio(double v) { velocityRatio = v; }
    double getVelocityRatio() const { return velocityRatio; }
    void setTimeRatio(double t) { timeRatio = t; }
    double getTimeRatio() const { return timeRatio; }

private:
    std::vector<LBMSimulationParameter> levels;
    std::string mainKernel = "K17CompressibleNavierStokes";
    double limiters[3] = {0.01, 0.01, 0.01};
    TurbulenceModel turbulenceModel = TurbulenceModel::None;
    double sgsConstant = 0.0;
    bool useTurbulentViscosity = false;
    double velocityRatio = 1.0;
    double timeRatio = 1.0;
    bool calcMean = false;
    int tCalcMedStart = 0, tCalcMedEnd = 0;
};

// CalcMean.cpp:42-80 and the averaging block of Simulation::run
// (Simulation.cpp:448-470, 689-697): allocMean gives every level zeroed sums,
// calculateMean adds the current state of a level (vfx_calculate_mean, at
// every step of the averaging window), calcMean divides the sums by the steps
// since the previous output into the *Out arrays, resetMean zeroes the sums,
// freeMean releases them.
void allocMean(Parameter* para);
void calculateMean(Parameter* para, int level);
void calcMean(Parameter* para, uint tdiff);
void resetMean(Parameter* para);
void freeMean(Parameter* para);

int32_t vfx_turbulence_model_of(TurbulenceModel m);

// TurbulenceModelFactory (TurbulenceModelFactory.cpp:44-100): selects the model
// (Parameter), and runs the AMD viscosity kernel between steps
class TurbulenceModelFactory {
public:
    explicit TurbulenceModelFactory(std::shared_ptr<Parameter> para) : para(std::move(para)) {}
    void setTurbulenceModel(TurbulenceModel m);
    void setModelConstant(double c) { para->setSGSConstant(c); }
    // "TurbulenceModel" = Smagorinsky | AMD | QR | None, "SGSconstant" = number
    void readConfig(const std::map<std::string, std::string>& config);
    void runTurbulenceModelKernel(int level) const;

private:
    std::shared_ptr<Parameter> para;
    TurbulenceModel turbulenceModel = TurbulenceModel::None;
};

void check(vfx_status s, const char* what);

class Kernel {
public:
    virtual ~Kernel() = default;
    virtual void run() = 0;
    virtual void runOnIndices(const unsigned int* indices, unsigned int size_indices, CollisionTemplate collisionTemplate,
                              CudaStreamIndex streamIndex = CudaStreamIndex::Legacy) = 0;
    virtual std::vector<PreProcessorType> getPreProcessorTypes() = 0;
};

class KernelImp : public Kernel {
public:
    void run() override;
    void runOnIndices(const unsigned int* indices, unsigned int size_indices, CollisionTemplate collisionTemplate,
                      CudaStreamIndex streamIndex = CudaStreamIndex::Legacy) override;
    std::vector<PreProcessorType> getPreProcessorTypes() override { return {PreProcessorType::InitNavierStokesCompressible}; }
    bool getKernelUsesFluidNodeIndices() const { return kernelUsesFluidNodeIndices; }

protected:
    KernelImp(std::shared_ptr<Parameter> para, int level, int32_t collision, bool usesIndices)
        : para(std::move(para)), level(level), collision(collision), kernelUsesFluidNodeIndices(usesIndices) {}
    std::shared_ptr<Parameter> para;
    int level;
    int32_t collision;
    bool kernelUsesFluidNodeIndices;
};

// "K17CompressibleNavierStokes" (K17CompressibleNavierStokes.cu:57-156), TurbulenceModel::None
class K17CompressibleNavierStokes : public KernelImp {
public:
    static std::shared_ptr<K17CompressibleNavierStokes> getNewInstance(std::shared_ptr<Parameter> para, int level)
    {
        return std::shared_ptr<K17CompressibleNavierStokes>(new K17CompressibleNavierStokes(std::move(para), level));
    }

private:
    K17CompressibleNavierStokes(std::shared_ptr<Parameter> p, int l) : KernelImp(std::move(p), l, VFX_K17, true) {}
};

// "BGKCompSP27" (B92CompressibleNavierStokes.cu:43-57)
class B92CompressibleNavierStokes : public KernelImp {
public:
    static std::shared_ptr<B92CompressibleNavierStokes> getNewInstance(std::shared_ptr<Parameter> para, int level)
    {
        return std::shared_ptr<B92CompressibleNavierStokes>(new B92CompressibleNavierStokes(std::move(para), level));
    }
    void runOnIndices(const unsigned int*, unsigned int, CollisionTemplate, CudaStreamIndex) override
    {
        throw std::runtime_error("B92CompressibleNavierStokes runs over all non-solid nodes: use run()");
    }

private:
    B92CompressibleNavierStokes(std::shared_ptr<Parameter> p, int l) : KernelImp(std::move(p), l, VFX_BGK, false) {}
};

class KernelFactory {
public:
    virtual ~KernelFactory() = default;
    virtual std::vector<std::shared_ptr<Kernel>> makeKernels(std::shared_ptr<Parameter> para) = 0;
};

class KernelFactoryImp : public KernelFactory {
public:
    std::vector<std::shared_ptr<Kernel>> makeKernels(std::shared_ptr<Parameter> para) override;
    std::shared_ptr<Kernel> makeKernel(std::shared_ptr<Parameter> para, const std::string& kernel, int level);
};

// BoundaryConditionFactory.h:50-53
using boundaryCondition = std::function<void(LBMSimulationParameter*, QforBoundaryConditions*)>;
using boundaryConditionDirectional = std::function<void(LBMSimulationParameter*, QforDirectionalBoundaryCondition*)>;
using boundaryConditionWithParameter = std::function<void(Parameter*, QforBoundaryConditions*, const int level)>;
using precursorBoundaryConditionFunc =
    std::function<void(LBMSimulationParameter*, QforPrecursorBoundaryConditions*, real timeRatio, real velocityRatio)>;

class BoundaryConditionFactory {
public:
    // subsets of the reference enums (BoundaryConditionFactory.h:60-132) this stepper implements
    enum class VelocityBC {
        VelocityBounceBack,
        VelocityInterpolatedCompressible,
        VelocityWithPressureInterpolatedCompressible, // combine with PressureBC::OutflowNonReflective
        NotSpecified
    };
    enum class NoSlipBC { NoSlipBounceBack, NoSlipInterpolatedCompressible, NotSpecified };
    enum class SlipBC { SlipCompressible, SlipTurbulentViscosityCompressible, NotSpecified };
    enum class StressBC {
        StressCompressible,
        StressBounceBackCompressible,
        StressBounceBackPressureCompressible,
        NotSpecified
    };
    enum class PressureBC {
        PressureNonEquilibriumCompressible,
        OutflowNonReflective,
        OutflowNonReflectivePressureCorrection,
        NotSpecified
    };

    void setVelocityBoundaryCondition(VelocityBC t) { velocity = t; }
    void setNoSlipBoundaryCondition(NoSlipBC t) { noSlip = t; }
    void setSlipBoundaryCondition(SlipBC t) { slip = t; }
    void setPressureBoundaryCondition(PressureBC t) { pressure = t; }
    void setStressBoundaryCondition(StressBC t) { stress = t; }
    enum class PrecursorBC { PrecursorNonReflectiveCompressible, PrecursorDistributions, NotSpecified };
    void setPrecursorBoundaryCondition(PrecursorBC t) { precursor = t; }
    precursorBoundaryConditionFunc getPrecursorBoundaryConditionPost() const;

    boundaryCondition getVelocityBoundaryConditionPost() const;
    boundaryCondition getNoSlipBoundaryConditionPost() const;
    boundaryCondition getSlipBoundaryConditionPost() const;
    // BoundaryConditionFactory.cpp:151-170: every pressure kind of this
    // stepper is directional (QforDirectionalBoundaryCondition)
    std::variant<boundaryCondition, boundaryConditionDirectional> getPressureBoundaryConditionPre() const;
    bool hasDirectionalPressureBoundaryCondition() const;
    boundaryConditionWithParameter getStressBoundaryConditionPost() const;

private:
    VelocityBC velocity = VelocityBC::NotSpecified;
    NoSlipBC noSlip = NoSlipBC::NotSpecified;
    SlipBC slip = SlipBC::NotSpecified;
    PressureBC pressure = PressureBC::NotSpecified;
    StressBC stress = StressBC::NotSpecified;
    PrecursorBC precursor = PrecursorBC::NotSpecified;
};

// GridScalingFactory.h:46-71
using gridScaling =
    std::function<void(LBMSimulationParameter* coarse, LBMSimulationParameter* fine, ICells*, ICellNeigh&, vfx_stream)>;

class GridScalingFactory {
public:
    enum class GridScaling { ScaleCompressible, NotSpecified };
    void setScalingFactory(GridScaling t) { type = t; }
    // the turbulent-viscosity variants of the reference are not part of this stepper
    gridScaling getGridScalingFC(bool hasTurbulentViscosity) const;
    gridScaling getGridScalingCF(bool hasTurbulentViscosity) const;

private:
    GridScaling type = GridScaling::NotSpecified;
};

// UpdateGrid27::updateGrid (UpdateGrid27.cpp:54-95): for a level below the
// finest, the next finer level is updated twice first; then collision over the
// tagged fluid nodes (or the whole box), post-collision boundary conditions,
// parity swap, the turbulence model's viscosity kernel (AMD), pre-collision
// boundary conditions and, below the finest level,
// fine-to-coarse then coarse-to-fine interpolation (Refinement_noExchange,
// RefinementStrategy.cpp:169-178).
class UpdateGrid27 {
public:
    UpdateGrid27(std::shared_ptr<Parameter> para, std::vector<std::shared_ptr<Kernel>> kernels,
                 const BoundaryConditionFactory& bcFactory, const GridScalingFactory* scalingFactory = nullptr,
                 std::shared_ptr<TurbulenceModelFactory> tmFactory = nullptr);
    void updateGrid(int level, unsigned int timestep);

private:
    std::shared_ptr<TurbulenceModelFactory> tmFactory;
    void postCollisionBC(int level, unsigned int timestep);
    void runPrecursorBCKernelPost(int level, unsigned int timestep);
    void preCollisionBC(int level);
    void refinement(int level);
    std::shared_ptr<Parameter> para;
    std::vector<std::shared_ptr<Kernel>> kernels;
    boundaryCondition velocityBC, noSlipBC, slipBC;
    boundaryConditionDirectional directionalPressureBC;
    boundaryConditionWithParameter stressBC;
    precursorBoundaryConditionFunc precursorBC;
    gridScaling scalingFineToCoarse, scalingCoarseToFine;
};

// -------------------------------------------------------------- restart ----
// RestartObject / ASCIIRestartObject / BinaryRestartObject
// (src/gpu/core/Restart/RestartObject.h:40-80, RestartObject.cpp): one
// distribution array per level, coarse to fine, in VirtualFluids' layout at
// each level's current parity.  Real is VirtualFluids' `real` (float by
// default, double with VF_DOUBLE_ACCURACY): it fixes the file's value type.
template <typename Real> class RestartObjectT {
public:
    virtual ~RestartObjectT() = default;

    void deserialize(const std::string& filename, std::shared_ptr<Parameter>& para);
    void serialize(const std::string& filename, const std::shared_ptr<Parameter>& para);

    std::vector<std::vector<Real>> fs;

    virtual void serialize_internal(const std::string& filename) = 0;
    virtual void deserialize_internal(const std::string& filename) = 0;

    // returns false (and prints a warning) if there was no file to delete
    bool delete_restart_file(const std::string& filename);

protected:
    virtual std::string getFileExtension() const = 0;
};

template <typename Real> class ASCIIRestartObjectT : public RestartObjectT<Real> {
public:
    void serialize_internal(const std::string& filename) override;
    void deserialize_internal(const std::string& filename) override;

protected:
    std::string getFileExtension() const override { return ".txt"; }
};

template <typename Real> class BinaryRestartObjectT : public RestartObjectT<Real> {
public:
    void serialize_internal(const std::string& filename) override;
    void deserialize_internal(const std::string& filename) override;

protected:
    std::string getFileExtension() const override { return ".bin"; }
};

using RestartObject = RestartObjectT<real>;
using ASCIIRestartObject = ASCIIRestartObjectT<real>;
using BinaryRestartObject = BinaryRestartObjectT<real>;

// ---------------------------------------------------------------- output ----
// FilePartCalculator (src/gpu/core/Output/FilePartCalculator.cpp:36-58)
struct FilePartCalculator {
    static constexpr uint32_t limitOfNodesForVTK = 30000000; // FilePartCalculator.h:53
    static uint32_t calculateNumberOfNodesInPart(uint32_t numberOfNodes, uint32_t indexOfFilePart);
    static uint32_t calculateNumberOfParts(uint32_t numberOfNodes);
    static uint32_t calculateStartingPostionOfPart(uint32_t indexOfPart);
};

// WriterUtilities (src/gpu/core/Output/WriterUtilities.cpp:38-86)
struct WriterUtilities {
    static std::string makePartFileNameEnding(uint32_t level, int ID, int part, int timestep);
    static bool isPeriodicCell(const HostGrid& grid, uint32_t baseNodeOfCell, uint32_t otherNodeInCell);
    static void getIndicesOfAllNodesInOct(std::array<uint32_t, 8>& nodeIndices, uint32_t baseNodeOfOct,
                                          const HostGrid& grid);
    static void calculateRelativeNodeIndexInPart(std::array<uint32_t, 8>& relativePositionInPart,
                                                 const std::array<uint32_t, 8>& indicesOfOct, uint32_t startPositionOfPart);
    static bool areAllNodesInOctValidForWriting(const std::array<uint32_t, 8>& indicesOfOct, const HostGrid& grid,
                                                uint32_t endPositionOfPart);
};

// The three writers of WbWriterVtkXmlBinary the output uses
// (src/basics/writer/WbWriterVtkXmlBinary.cpp:88-177, 1306-1425)
struct WbWriterVtkXmlBinary {
    static std::string writeOctsWithNodeData(const std::string& filename, const std::vector<std::array<float, 3>>& nodes,
                                             const std::vector<std::array<uint32_t, 8>>& cells,
                                             const std::vector<std::string>& datanames,
                                             const std::vector<std::vector<double>>& nodedata);
    static std::string writeParallelFile(const std::string& filename, const std::vector<std::string>& pieceSources,
                                         const std::vector<std::string>& pointDataNames,
                                         const std::vector<std::string>& cellDataNames);
    static std::string writeCollectionForTimeSeries(const std::string& filename,
                                                    const std::map<uint32_t, std::vector<std::string>>& files,
                                                    bool separateGroups);
};

// FileWriter (src/gpu/core/Output/FileWriter.cpp:102-330): at an output step
// every level's macroscopic quantities (vfx_macroscopic_quantities, plus the
// turbulent viscosity with a turbulence model) are copied to the host and
// written as parts of at most 30 M nodes, a .pvtu per step and, at the last
// step, a .pvd of the time series.  The ratios scale to physical units as the
// reference's Parameter does (1 = lattice units).
class FileWriter {
public:
    explicit FileWriter(std::string prefix, int ID = 0) : prefix(std::move(prefix)), ID(ID) {}
    double densityRatio = 1.0, velocityRatio = 1.0, viscosityRatio = 1.0;

    void writeInit(const std::shared_ptr<Parameter>& para, unsigned int timestep);
    void writeTimestep(const std::shared_ptr<Parameter>& para, unsigned int timestep, bool lastTimestep = false);
    std::vector<std::string> getNodeDataNames(const Parameter& para) const;
    // the parts of one level from host data (already scaled), as writeUnstructuredGridLT
    std::vector<std::string> writeUnstructuredGridLT(const HostGrid& grid, const std::vector<std::string>& names,
                                                     const std::vector<std::vector<double>>& nodeData, int level,
                                                     unsigned int timestep);
    // writeUnstructuredGridMeanLT (FileWriter.cpp:332-431): the _bin_mean parts of
    // a level from its *Out means, scaled in real arithmetic
    std::vector<std::string> writeUnstructuredGridMeanLT(const HostGrid& grid, const LBMSimulationParameter& level,
                                                         int levelIndex, unsigned int timestep);
    static std::vector<std::string> getMeanNodeDataNames();

private:
    void writeLevel(const std::shared_ptr<Parameter>& para, unsigned int timestep, int level);
    std::string prefix;
    int ID;
    std::vector<std::string> fileNamesForCollectionFile, fileNamesForCollectionFileMean;
    std::map<uint32_t, std::vector<std::string>> fileNamesForCollectionFileTimeSeries;
};

// Precursor input files (src/gpu/GridGenerator/TransientBCSetter/TransientBCSetter.h:53-233):
// the planes PrecursorWriter writes (VTK image data, x and y = the inflow
// plane's y and z, z = time; <prefix>_lev_<l>_ID_<id>_File_<part>.bin.vti),
// their discovery and the reader that builds each boundary node's four-point
// stencil (fillArrays) and fills the plane array at a physical time
// (getNextData).  Same behaviour as the reference, with two changes off the
// precursor grid: the 0PM neighbour goes to the 0PM list (the reference
// appends it to 0PP, TransientBCSetter.cpp:386) and a file point is looked up
// among the read indices (getWriteIndex, :430-441, searches the write indices).
enum class TransientBCFileType { VTK };

struct Quantity {
    std::string name;
    int offset = 0;
    std::vector<double> values;
};

class VTKFile {
public:
    explicit VTKFile(std::string fileName) : fileName(std::move(fileName)) { readHeader(); }
    void getData(real* data, uint numberOfNodes, const std::vector<uint>& readIndices, const std::vector<uint>& writeIndices,
                 uint offsetRead, uint offsetWrite);
    bool markNANs(const std::vector<uint>& readIndices) const;
    bool inBoundingBox(real x, real y, real z) const { return inXBounds(x) && inYBounds(y) && inZBounds(z); }
    bool inXBounds(real x) const { return x <= maxX && x >= minX; }
    bool inYBounds(real y) const { return y <= maxY && y >= minY; }
    bool inZBounds(real z) const { return z <= maxZ && z >= minZ; }
    int findNeighborMMM(real x, real y, real z) const { return neighbor(getIdxM00(x), getIdx0M0(y), getIdx00M(z)); }
    int findNeighborMPM(real x, real y, real z) const { return neighbor(getIdxM00(x), getIdx0M0(y) + 1, getIdx00M(z)); }
    int findNeighborPMM(real x, real y, real z) const { return neighbor(getIdxM00(x) + 1, getIdx0M0(y), getIdx00M(z)); }
    int findNeighborPPM(real x, real y, real z) const { return neighbor(getIdxM00(x) + 1, getIdx0M0(y) + 1, getIdx00M(z)); }
    int getIdxX(int i) const { return i % nx; }
    int getIdxY(int i) const { return (i / nx) % ny; }
    int getIdxZ(int i) const { return i / (nx * ny); }
    real getX(int i) const { return getIdxX(i) * deltaX + minX; }
    real getY(int i) const { return getIdxY(i) * deltaY + minY; }
    real getZ(int i) const { return getIdxZ(i) * deltaZ + minZ; }
    int getIdxM00(real x) const { return (int)((x - minX) / deltaX); }
    int getIdx0M0(real y) const { return (int)((y - minY) / deltaY); }
    int getIdx00M(real z) const { return (int)((z - minZ) / deltaZ); }
    int getClosestIdxZ(real z) const;
    int getLinearIndex(int ix, int iy, int iz) const { return ix + nx * (iy + ny * iz); }
    int getNumberOfPointsInXYPlane() const { return nx * ny; }
    int getNumberOfPoints() const { return nx * ny * nz; }
    size_t getNumberOfQuantities() const { return quantities.size(); }
    void loadFile();
    void unloadFile();
    bool isLoaded() const { return loaded; }

private:
    int neighbor(int ix, int iy, int iz) const
    {
        const int i = getLinearIndex(ix, iy, iz);
        return i >= 0 && i < nx * ny * nz ? i : -1;
    }
    void readHeader();
    std::string fileName;
    real minX = 0, maxX = 0, minY = 0, maxY = 0, minZ = 0, maxZ = 0;
    real deltaX = 1, deltaY = 1, deltaZ = 1;
    int nx = 0, ny = 0, nz = 0;
    std::vector<Quantity> quantities;
    bool loaded = false;
};

class FileCollection {
public:
    explicit FileCollection(std::string prefix) : prefix(std::move(prefix)) {}
    virtual ~FileCollection() = default;
    virtual size_t getNumberOfQuantities() = 0;
    virtual TransientBCFileType getFileType() = 0;

protected:
    std::string prefix;
};

class VTKFileCollection : public FileCollection {
public:
    explicit VTKFileCollection(std::string prefix) : FileCollection(std::move(prefix)) { findFiles(); }
    TransientBCFileType getFileType() override { return TransientBCFileType::VTK; }
    size_t getNumberOfQuantities() override { return files[0][0][0].getNumberOfQuantities(); }
    static inline const std::string suffix = "vti";
    std::vector<std::vector<std::vector<VTKFile>>> files;

private:
    void findFiles();
    std::string makeFileName(int level, int id, int part) const;
};

class TransientBCInputFileReader {
public:
    virtual ~TransientBCInputFileReader() = default;
    virtual void getNextData(real* data, uint numberOfNodes, real time) = 0;
    virtual void fillArrays(std::vector<real>& coordsY, std::vector<real>& coordsZ) = 0;
    uint getNPoints() const { return nPoints; }
    uint getNPointsRead() const { return nPointsRead; }
    size_t getNumberOfQuantities() const { return nQuantities; }
    void setWritingOffset(uint offset) { writingOffset = offset; }
    void getNeighbors(uint* neighbor0PP, uint* neighbor0PM, uint* neighbor0MP, uint* neighbor0MM);
    void getWeights(real* weights0PP, real* weights0PM, real* weights0MP, real* weights0MM);

    std::vector<uint> planeNeighbor0PP, planeNeighbor0PM, planeNeighbor0MP, planeNeighbor0MM;
    std::vector<real> weights0PP, weights0PM, weights0MP, weights0MM;

protected:
    uint nPoints = 0, nPointsRead = 0, writingOffset = 0;
    size_t nQuantities = 0;
};

class VTKReader : public TransientBCInputFileReader {
public:
    VTKReader(std::shared_ptr<VTKFileCollection> fileCollection, uint readLevel);
    void getNextData(real* data, uint numberOfNodes, real time) override;
    void fillArrays(std::vector<real>& coordsY, std::vector<real>& coordsZ) override;

private:
    uint getWriteIndex(int level, int id, int linearIndex);
    void initializeIndexVectors();
    std::vector<std::vector<std::vector<uint>>> readIndices, writeIndices;
    std::vector<std::vector<size_t>> nFile;
    std::shared_ptr<VTKFileCollection> fileCollection;
    uint readLevel;
};

std::shared_ptr<FileCollection> createFileCollection(std::string prefix, TransientBCFileType type);
std::shared_ptr<TransientBCInputFileReader> createReaderForCollection(std::shared_ptr<FileCollection> fileCollection,
                                                                      uint readLevel);
// the reader as the precursor boundary's plane source (GridGenerator hands
// the reader's getNextData the array `next`, GridGenerator.cpp:436-478)
void setPrecursorFileReader(QforPrecursorBoundaryConditions& q, std::shared_ptr<TransientBCInputFileReader> reader);

VFX_HOST_END
