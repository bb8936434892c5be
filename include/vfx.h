/* vfx.h -- C ABI of the MI355X (gfx950) D3Q27 lattice-Boltzmann stepper.
 *
 * One GPU, stream ordered, no process aborts: every entry point returns a
 * vfx_status and vfx_last_error() describes the last failure of the calling
 * thread.  All array pointers are device pointers owned by the caller; the
 * library keeps none of them between calls.
 *
 * Data layout (identical to VirtualFluids' GPU path so arrays can be handed
 * over unchanged):
 *   - distributions: 27 planes of n_nodes values, plane d at d*n_nodes, holding
 *     shifted ("well-conditioned") populations f_d - w_d in VirtualFluids'
 *     direction order (src/lbm/constants/D3Q27.h:42-72);
 *   - in-place esoteric-twist streaming: a step of parity is_even_timestep
 *     reads and writes the same elements (src/gpu/core/Utilities/
 *     KernelUtilities.h:66-125, :252-379); the caller flips the parity after
 *     every step exactly as UpdateGrid27::swapBetweenEvenAndOddTimestep does
 *     (src/gpu/core/Calculation/UpdateGrid27.cpp:264-268).
 *
 * Storage forms (vfx_lattice.storage):
 *   VFX_FP64          double populations, FP64 arithmetic;
 *   VFX_FP32          float populations, FP32 arithmetic;
 *   VFX_FP16_SCALED   half populations h_d = (f_d - w_d) * scale / w_d, FP32
 *                     arithmetic (scale = vfx_lattice.fp16_scale, one per lattice).
 */
#ifndef VFX_H
#define VFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* vfx_stream; /* a hipStream_t; NULL = the null stream */

typedef enum { VFX_OK = 0, VFX_EINVAL = 1, VFX_EHIP = 2, VFX_EUNSUPPORTED = 3 } vfx_status;

typedef enum { VFX_FP64 = 0, VFX_FP32 = 1, VFX_FP16_SCALED = 2 } vfx_storage;

/* collision operators: the kernels VirtualFluids registers as
 * "K17CompressibleNavierStokes" and "BGKCompSP27" (src/gpu/core/Kernel/KernelTypes.h:35-46) */
typedef enum { VFX_K17 = 0, VFX_BGK = 1 } vfx_collision;

/* CollisionTemplate (src/gpu/core/Calculation/Calculation.h:55-66):
 * WriteMacroVars additionally stores drho, vx, vy, vz of every collided node;
 * ApplyBodyForce adds the node's body force (vfx_lattice.body_force) to the
 * global force, force = (forces + body_force[k]) / 2 * force_factor, and resets
 * body_force[k] to 0 (RunCollision.cuh:88-97); AllFeatures does both (the
 * reference's SubDomainBorder runs as AllFeatures, K17CompressibleNavierStokes.cu:126-131). */
typedef enum {
    VFX_TEMPLATE_DEFAULT = 0,
    VFX_TEMPLATE_WRITE_MACRO = 1,
    VFX_TEMPLATE_APPLY_BODY_FORCE = 2,
    VFX_TEMPLATE_ALL_FEATURES = 3
} vfx_collision_template;

typedef struct vfx_lattice {
    void* f;               /* 27 * n_nodes populations of the storage type */
    int32_t storage;       /* vfx_storage */
    float fp16_scale;      /* VFX_FP16_SCALED only: per-lattice scale (> 0) */
    uint64_t n_nodes;      /* nodes per plane */

    /* addressing: direct = 1 -> structured box dims[0]*dims[1]*dims[2], x fastest,
     * node k = x + dims[0]*(y + dims[1]*z), neighbours implicit with periodic index
     * wrap on every axis (walls are solid nodes); solid_mask bit k set = node k does
     * not collide (NULL: every node collides).
     * direct = 0 -> VirtualFluids sparse layout: neighbor_x/y/z[k] is the node in +x/+y/+z
     * (src/gpu/GridGenerator/grid/GridImp.cpp:1031-1080), fluid_nodes lists the
     * nodes that collide (src/gpu/core/Utilities/RunCollision.cuh:79-82). */
    int32_t direct;
    uint32_t dims[3];
    const uint32_t* solid_mask;
    const uint32_t* neighbor_x;
    const uint32_t* neighbor_y;
    const uint32_t* neighbor_z;
    const uint32_t* fluid_nodes;
    uint32_t n_fluid;

    /* collision parameters (vf::lbm::CollisionParameter, src/lbm/collision/CollisionParameter.h:42-50) */
    double omega;
    double quadric_limiter[3]; /* {P, M, D}; VirtualFluids' default is 0.01 */
    double force[3];           /* added to the velocity inside the collision (CollisionParameter::forceX..Z) */

    /* VFX_TEMPLATE_WRITE_MACRO output, n_nodes each, double for VFX_FP64 else float */
    void* rho;
    void* vx;
    void* vy;
    void* vz;

    /* 0: default -- FP64 keeps the reference operation order (bit-identical to the
     *    CPU path); FP32/FP16 run the reduced-operation K17 (same collision, folded
     *    algebra, hardware reciprocal);
     * 1: reference operation order in every precision ("plain" kernel). */
    int32_t kernel_variant;

    /* VFX_FP16_SCALED only: the stored value of direction d is the deviation of
     * f_d - w_d from the equilibrium at this velocity (drho = 0), scaled by
     * fp16_scale / w_d.  Choosing the mean flow (e.g. the inflow velocity)
     * keeps the stored deviations small, so FP16 rounds them less.  {0,0,0}:
     * plain shifted populations.  Must be zero for FP32/FP64 storage. */
    double fp16_reference_velocity[3];

    /* Sparse addressing only: VirtualFluids' typeOfGridNode (device, n_nodes
     * entries).  When set, every node whose type is neither GEO_SOLID (15) nor
     * GEO_VOID (16) collides (and is initialised, read and reduced), as
     * B92CompressibleNavierStokes_Device selects them
     * (src/gpu/core/Kernel/Compressible/NavierStokes/B92/B92CompressibleNavierStokes_Device.cu:58-66,
     * constants src/gpu/core/Calculation/Calculation.h:43-45); fluid_nodes is
     * then not used.  NULL: the fluid_nodes list selects the nodes. */
    const uint32_t* type_of_grid_node;

    /* Eddy-viscosity model of the K17 collision (vf::lbm::TurbulenceModel,
     * src/lbm/collision/TurbulentViscosity.h:50-61; run by runCollision,
     * src/gpu/core/Utilities/RunCollision.cuh:116-141).  Anything but
     * VFX_TURB_NONE: node k relaxes with omega / (1 + 3 omega nu_t[k])
     * (calculateOmegaWithturbulentViscosity, TurbulentViscosity.h:86-89);
     * VFX_TURB_SMAGORINSKY / VFX_TURB_QR store the model's nu_t of the new
     * velocity gradients back into turbulent_viscosity
     * (K17CompressibleNavierStokes.h:355-368); VFX_TURB_AMD takes nu_t from
     * vfx_turbulent_viscosity_amd and makes the collision write rho/vx/vy/vz,
     * which must then be set.  The reference-order K17 kernel runs (any
     * kernel_variant).  BGK ignores the model. */
    int32_t turbulence_model;     /* vfx_turbulence_model */
    double sgs_constant;          /* Parameter::SGSConstant (Smagorinsky Cs, QR / AMD constant) */
    void* turbulent_viscosity;    /* n_nodes values, double for VFX_FP64 else float (Parameter turbViscosity) */
    /* Sparse addressing with VFX_TURB_AMD: the node in -x-y-z
     * (LBMSimulationParameter::neighborInverse, src/gpu/GridGenerator/grid/GridImp.cpp:1034, 1112-1130); the
     * AMD kernel then takes GEO_FLUID (19) from type_of_grid_node, which must
     * be set.  Direct addressing: neighbours implicit, non-solid = fluid. */
    const uint32_t* neighbor_inverse;

    /* VFX_TEMPLATE_APPLY_BODY_FORCE / _ALL_FEATURES: per-node body force
     * (LBMSimulationParameter::forceX_SP/Y/Z, n_nodes values, double for VFX_FP64
     * else float; reset to 0 by the collision), the global force unscaled
     * (Parameter::forcesDouble) and the level's force factor 1 / 2^level
     * (getForceFactor, KernelUtilities.h:56-64; must be > 0).  The templates
     * without body force use `force` above, which is (forces / 2) * force_factor. */
    void* body_force[3];
    double forces[3];
    double force_factor;

    /* Launch shape of the collide-and-stream step: nodes per workgroup along
     * x (structured step: 64, 128 or 256; 0 = 256; the sparse step's
     * workgroups are 256 work items: 0 or 256) and the XCD tile run
     * (workgroups reach the 8 XCDs round robin; consecutive tiles are handed
     * to each XCD in runs of this many, a power of two up to 4096, 1 = plain
     * order; 0 = the built-in choice, 16 for 2- and 4-byte populations, 256
     * for 8-byte ones).
     * vfx_k17_autotune measures them.  Other kernels ignore both. */
    uint32_t tile_nodes;
    uint32_t xcd_run;

    /* Sparse addressing: the table vfx_sparse_regular_groups fills for these
     * index arrays and this fluid-node list (NULL: every wavefront reads its
     * indices).  It is read unchecked for every wavefront of the step, so it
     * must be the one made for these arrays and list: recompute it when they
     * change, and pass NULL with any other list (the host layer's
     * runOnIndices does). */
    const void* regular_groups;
} vfx_lattice;

/* vf::lbm::TurbulenceModel */
typedef enum {
    VFX_TURB_NONE = 0,
    VFX_TURB_SMAGORINSKY = 1,
    VFX_TURB_AMD = 2, /* Rozema et al., Phys. Fluids 27, 085107 (2015) */
    VFX_TURB_QR = 3   /* Verstappen's QR model */
} vfx_turbulence_model;

/* boundary-condition kinds (src/gpu/core/BoundaryConditions/BoundaryConditionFactory.h:60-132) */
typedef enum {
    VFX_BC_VELOCITY_BOUNCE_BACK = 0,  /* VelocityBounceBack_Device, Velocity/VelocityBounceBack.cu:44 (u = 0: NoSlipBounceBack) */
    VFX_BC_NOSLIP_INTERPOLATED = 1,   /* NoSlipInterpolatedCompressible_Device, NoSlip/NoSlipInterpolatedCompressible.cu:44 */
    VFX_BC_VELOCITY_INTERPOLATED = 2, /* VelocityInterpolatedCompressible_Device, Velocity/VelocityInterpolatedCompressible.cu:44 */
    VFX_BC_SLIP = 3,                  /* SlipCompressible_Device, Slip/SlipCompressible.cu:45 */
    VFX_BC_PRESSURE_NEQ = 4,          /* PressureNonEquilibriumCompressible_Device, Pressure/PressureNonEquilibriumCompressible.cu:45
                                         (pre-collision: apply with the parity of the step it precedes) */
    VFX_BC_VELOCITY_PRESSURE_INTERPOLATED = 5, /* VelocityWithPressureInterpolatedCompressible_Device,
                                                  Velocity/VelocityWithPressureInterpolatedCompressible.cu:44 */
    VFX_BC_OUTFLOW_NON_REFLECTING = 6, /* OutflowNonReflecting_Device, Outflow/OutflowNonReflecting.cu:50 (pre-collision;
                                          k_neighbor, normal) */
    VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION = 7, /* OutflowNonReflectingPressureCorrection_Device,
                                          Outflow/OutflowNonReflectingPressureCorrection.cu:51 (pre-collision;
                                          k_neighbor, normal, pressure_correction_factor) */
    VFX_BC_SLIP_TURBULENT_VISCOSITY = 8, /* SlipTurbulentViscosityCompressible_Device, Slip/SlipTurbulentViscosityCompressible.cu:45:
                                          slip with omega / (1 + 3 omega max(0, nu_t[k])); the lattice needs a
                                          turbulence model and its turbulent_viscosity */
    /* wall models (BoundaryConditions/Stress/, iMEM.cuh): need k_neighbor (the
       sampling node), the wm_* arrays and the lattice's vx/vy arrays holding the
       velocities of this step's collision (collide with VFX_TEMPLATE_WRITE_MACRO) */
    VFX_BC_STRESS_INTERPOLATED = 9,    /* StressCompressible_Device, Stress/StressCompressible.cu:46 */
    VFX_BC_STRESS_BOUNCE_BACK = 10,    /* StressBounceBackCompressible_Device, Stress/StressBounceBackCompressible.cu:46 */
    VFX_BC_STRESS_BOUNCE_BACK_PRESSURE = 11, /* StressBounceBackPressureCompressible_Device,
                                               Stress/StressBounceBackPressureCompressible.cu:46 */
    /* Precursor inflows (QforPrecursorBoundaryConditions, Calculation.h:120-150): the
       plane data of the precursor nodes, interpolated to each boundary node with
       four inverse-distance weights and blended in time,
       last*(1 - time_ratio) + current*time_ratio.  Post-collision, the last
       boundary condition of UpdateGrid27::postCollisionBC (UpdateGrid27.cpp:252). */
    /* PrecursorNonReflectiveCompressible_Device (Precursor/PrecursorNonReflectiveCompressible.cu:46):
       velocity inlet with pressure on the q links, velocity = (vx + blended
       plane velocity) / velocity_ratio; planes hold vx, vy, vz of every
       precursor node (3 * n_precursor_nodes).  Writes the parity's own slots, so
       a value is read by the node's collision two steps later. */
    VFX_BC_PRECURSOR_VELOCITY = 12,
    /* PrecursorDistributions_Device (Precursor/PrecursorDistributions.cu:44): sets
       the nine populations with a positive x component (dP00, dPP0, dPM0, dP0P,
       dP0M, dPPP, dPMP, dPPM, dPMM); planes hold those nine of every precursor
       node (9 * n_precursor_nodes).  No q links. */
    VFX_BC_PRECURSOR_DISTRIBUTIONS = 13
} vfx_bc_kind;

/* QforBoundaryConditions (src/gpu/core/Calculation/Calculation.h:105-117).
 * Every floating-point array is of the lattice's `real`: double for VFX_FP64
 * storage (VirtualFluids built with VF_DOUBLE_ACCURACY), float for VFX_FP32 and
 * VFX_FP16_SCALED (VirtualFluids' default `real`).  The reference's launchers
 * hand their kernels q27[0], the base of one SoA block of 27 * numberOfBCnodes
 * values (e.g. NoSlip.cu, Stress.cu:51); q27 here is that pointer. */
typedef struct vfx_bc {
    int32_t kind;              /* vfx_bc_kind */
    uint32_t n;                /* numberOfBCnodes */
    const uint32_t* k;         /* node index of each boundary node (QforBoundaryConditions::k) */
    const void* q27;           /* real[27*n] subgrid distances, SoA (q27[0]); link d of node i is a wall link
                                  iff 0 <= q27[d*n+i] <= 1 */
    const void* vx;            /* real[n] wall velocity per boundary node, Vx/Vy/Vz (NULL = 0) */
    const void* vy;
    const void* vz;
    int32_t cpu_float_velocity; /* 1: wall term of VFX_BC_VELOCITY_BOUNCE_BACK rounded as VirtualFluids' CPU
                                   BoundaryConditions does (float velocity and sum, BoundaryConditions.h:195-262) */
    const void* rho;           /* real[n], VFX_BC_PRESSURE_NEQ: prescribed density deviation per boundary node (RhoBC) */
    const uint32_t* k_neighbor;/* VFX_BC_PRESSURE_NEQ / outflow: inward neighbour node of each boundary node */
    int32_t normal;            /* VFX_BC_PRESSURE_NEQ / outflow: outward direction of the boundary (e.g. 1 = dP00);
                                  the outflow kinds act on the six face directions only */
    /* VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION: Parameter::outflowPressureCorrectionFactor
       (Parameter.h:761, default 0) */
    double pressure_correction_factor;
    /* Stress kinds: per boundary node, arrays of the lattice's real type (double
       for VFX_FP64 else float): QforBoundaryConditions normalX/Y/Z, Vx/Vy/Vz
       (filtered velocity at the sampling node, updated), Vx1/Vy1/Vz1 (filtered
       velocity at the wall node, updated), WallModelParameters samplingOffset
       and z0, and the optional monitors u_star, Fx/Fy/Fz (NULL: not written).
       As the reference's launcher (Stress.cu:57-59), the sampling node's z
       velocity is read from the lattice's vy array. */
    const void* wm_normal[3];
    void* wm_exchange_velocity[3];
    void* wm_wall_velocity[3];
    const int32_t* wm_sampling_offset;
    const void* wm_z0;
    void* wm_u_star;
    void* wm_force[3];
    /* Precursor kinds: planeNeighbor0PP/0PM/0MP/0MM (precursor node of each
       boundary node), weights0PP/... (arrays of the lattice's real type; a first
       weight >= 1e6 takes the 0PP node alone), the plane arrays last and current
       of n_precursor_nodes nodes (real type), the time ratio and
       Parameter::getVelocityRatio() (velocity kind only). */
    uint32_t n_precursor_nodes;
    const uint32_t* plane_neighbor[4];
    const void* plane_weight[4];
    const void* precursor_last;
    const void* precursor_current;
    double time_ratio;
    double velocity_ratio;
    /* VFX_BC_PRECURSOR_VELOCITY: QforPrecursorBoundaryConditions velocityX/Y/Z,
       the scalars the reference passes by value; added to vx/vy/vz[i] (0 where
       those are NULL) before the blended plane velocity */
    double precursor_velocity[3];
} vfx_bc;

/* ---------------------------------------------------------------- kernels */

/* Kernel::run / runOnIndices of "K17CompressibleNavierStokes" and "BGKCompSP27":
 * one collide-and-stream pass over every colliding node
 * (src/gpu/core/Kernel/Compressible/NavierStokes/K17/K17CompressibleNavierStokes.cu:85-145,
 *  src/gpu/core/Kernel/Compressible/NavierStokes/B92/B92CompressibleNavierStokes.cu:43-57). */
vfx_status vfx_collide_stream(const vfx_lattice* lat, int32_t collision, int32_t is_even_timestep,
                              int32_t collision_template, vfx_stream stream);
vfx_status vfx_k17_step(const vfx_lattice* lat, int32_t is_even_timestep, int32_t collision_template, vfx_stream stream);
vfx_status vfx_bgk_step(const vfx_lattice* lat, int32_t is_even_timestep, vfx_stream stream);

/* (new) Regular wavefronts of a sparse lattice: where the 64 work items of a
 * wavefront are 64 consecutive nodes (x-fastest numbering) whose eight twist
 * nodes follow one y and one z stride -- checked against neighbor_x/y/z for
 * every colliding node -- the step forms the indices by arithmetic instead of
 * reading them (same indices, same results).  `table` is a device buffer of
 * vfx_sparse_regular_groups_bytes(lat) bytes (16 per wavefront of the
 * step's grid); set lat->regular_groups to it.  Sparse lattices only. */
uint64_t vfx_sparse_regular_groups_bytes(const vfx_lattice* lat);
vfx_status vfx_sparse_regular_groups(const vfx_lattice* lat, void* table, uint64_t table_bytes, vfx_stream stream);

/* (new) Autotuner of the K17 step: times vfx_k17_step with every tile width
 * (64, 128, 256 nodes; sparse lattices: 256 only) and XCD run (1, 4, 8, 16,
 * 32, 64, 256) on `steps` (made even, at least 2) steps after two warm-up
 * steps each, from parity is_even_timestep, and returns the fastest shape
 * (*tile_nodes, *xcd_run, its ms per step in *ms) for the caller to set in
 * its vfx_lattice.  The populations advance (42 trials structured, 14
 * sparse), the parity is the same afterwards. */
vfx_status vfx_k17_autotune(const vfx_lattice* lat, int32_t is_even_timestep, int32_t collision_template,
                            int32_t steps, vfx_stream stream, uint32_t* tile_nodes, uint32_t* xcd_run, float* ms);

/* Boundary conditions fused into the collide-and-stream kernel of a
 * structured lattice, so a boundary node's populations are not read again by
 * a kernel of their own (on a face normal to x each of them sits in its own
 * cache line).  node_mask: ceil(n_nodes/32) words, bit k set iff node k is an
 * entry of bc; word_base: ceil(n_nodes/32) words, the number of entries on
 * nodes below 32*w.  bc->k must be sorted ascending, one entry per node. */
typedef struct vfx_fused_bc {
    const vfx_bc* bc;
    const uint32_t* node_mask;
    const uint32_t* word_base;
} vfx_fused_bc;

/* vfx_k17_step (VFX_TEMPLATE_DEFAULT, default kernel variant, no turbulence
 * model) of a structured lattice with fused boundary conditions:
 *  - post[0..n_post-1] (n_post <= 2): interpolated kinds 1, 2, 3, 5, applied by
 *    the thread that collided each boundary node -- the same populations as
 *    vfx_bc_apply after the step;
 *  - pre (or NULL): VFX_BC_PRESSURE_NEQ with normal dP00 (dM00) whose every
 *    entry has k_neighbor = k - 1 (k + 1) in the same wavefront of 64 nodes,
 *    i.e. x mod 64 != 0 (!= 63): the replacement vfx_bc_apply makes after the
 *    previous step's parity flip is made inside this step, from the
 *    neighbour's populations the neighbouring lane gathers (apply_pre = 1;
 *    0 on the first step of a run, before which no pressure boundary ran).
 *    Until the next step the boundary nodes then hold the un-replaced
 *    populations: vfx_bc_apply(pre->bc) with the new parity gives the state
 *    the separate launches leave, before it is read;
 *  - any_mask (or NULL): ceil(n_nodes/32) words, bit k set iff node k is an
 *    entry of some group above; loaded with the populations, it spares the
 *    other nodes the look-up in each group's mask.
 * A post group entry must be a fluid node whose every wall link (0 <= q <= 1)
 * ends in a solid node: the kernel writes the re-entering populations into
 * elements only that node would touch, and reads the boundary node's own
 * post-collision elements back, while the launch is running.
 * Results equal vfx_k17_step + vfx_bc_apply bit for bit in every storage form. */
vfx_status vfx_k17_step_fused(const vfx_lattice* lat, int32_t is_even_timestep, const vfx_fused_bc* post,
                              int32_t n_post, const vfx_fused_bc* pre, const uint32_t* any_mask, int32_t apply_pre,
                              vfx_stream stream);

/* The collide-and-stream step of one or two lattices (the levels of a nest,
 * each at its own parity and rate) in one launch: per level the plain K17 step
 * of vfx_k17_step (VFX_TEMPLATE_DEFAULT, default variant, no turbulence model,
 * no body-force arrays), structured (direct) addressing.  UpdateGrid27 runs the
 * levels' kernels one after another (src/gpu/core/Calculation/UpdateGrid27.cpp:54-95);
 * the levels' arrays are independent, so one launch gives the same values. */
typedef struct vfx_level_step {
    const vfx_lattice* lattice;
    int32_t is_even_timestep;
} vfx_level_step;
vfx_status vfx_k17_step_levels(const vfx_level_step* levels, int32_t n_levels, vfx_stream stream);

/* calcTurbulentViscosityAMD (src/gpu/core/PostProcessor/TurbulentViscosityKernels.cu:62-130):
 * the AMD eddy viscosity of every fluid node from vx/vy/vz into
 * turbulent_viscosity, with constant sgs_constant.  UpdateGrid27 runs it after
 * the parity flip and before the pre-collision boundary conditions
 * (src/gpu/core/Calculation/UpdateGrid27.cpp:77-85, TurbulenceModelFactory).
 * Sparse layout: needs type_of_grid_node and neighbor_inverse. */
vfx_status vfx_turbulent_viscosity_amd(const vfx_lattice* lat, vfx_stream stream);

/* Boundary condition on a BC node list.  Post-collision kinds run after the
 * collision of parity is_even_timestep and before the parity flip
 * (src/gpu/core/Calculation/UpdateGrid27.cpp:227-262); VFX_BC_PRESSURE_NEQ is
 * pre-collision and runs after the flip with the new parity (:288-307). */
vfx_status vfx_bc_apply(const vfx_lattice* lat, const vfx_bc* bc, int32_t is_even_timestep, vfx_stream stream);

/* Boundary groups applied by one launch, as UpdateGrid27 applies a phase's
 * boundary kernels one after another (src/gpu/core/Calculation/UpdateGrid27.cpp:227-307):
 * the groups of a set (1..32, kinds 0, 1, 2, 3, 4, 5; lattices of one storage
 * type, e.g. the coarse and the fine level of a nest) run side by side, each
 * at its own parity (bit j of even_mask for group j), with exactly the values
 * vfx_bc_apply gives.  The groups must not write elements another group of
 * the set reads (the post-collision groups of one step write wall nodes'
 * elements only; keep pre-collision pressure groups in a set of their own).
 * The set copies the lattices' and tables' pointers at creation: recreate it
 * when an array moves. */
typedef struct vfx_bc_set_s* vfx_bc_set;
vfx_status vfx_bc_set_create(const vfx_lattice* const* lattices, const vfx_bc* const* bcs, int32_t n, vfx_bc_set* out);
vfx_status vfx_bc_set_apply(vfx_bc_set set, uint32_t even_mask, vfx_stream stream);
vfx_status vfx_bc_set_destroy(vfx_bc_set set);

/* InitNavierStokesCompressible_Device (src/gpu/core/PreProcessor/PreProcessorStrategy/
 * InitNavierStokesCompressible/InitNavierStokesCompressible_Device.cu:44-213):
 * shifted compressible equilibrium of (drho, ux, uy, uz), written for parity
 * is_even_timestep.  Fields are device arrays of n_nodes doubles; solid nodes
 * (direct: solid_mask; sparse: all nodes are written) are skipped. */
vfx_status vfx_init_equilibrium(const vfx_lattice* lat, int32_t is_even_timestep, const double* drho,
                                const double* ux, const double* uy, const double* uz, vfx_stream stream);

/* Copy the pre-collision populations of every node (parity is_even_timestep)
 * out to / in from a node-major double array out[node*27 + d] (device). */
vfx_status vfx_read_populations(const vfx_lattice* lat, int32_t is_even_timestep, double* out, vfx_stream stream);
vfx_status vfx_write_populations(const vfx_lattice* lat, int32_t is_even_timestep, const double* in, vfx_stream stream);

/* The distribution array in VirtualFluids' own layout -- 27 planes of n_nodes
 * shifted populations at parity is_even_timestep, as distributions.f[0] holds
 * it -- copied to / from a device array of 27*n_nodes values of `precision`
 * (VFX_FP32 for VirtualFluids' default `real`, VFX_FP64 with
 * VF_DOUBLE_ACCURACY).  This is the array RestartObject::serialize writes and
 * RestartObject::deserialize reads back (src/gpu/core/Restart/RestartObject.cpp:
 * 19-48); export writes NaN as 0 as serialize does.  FP32/FP64 storage is
 * copied, FP16 storage decoded on export and encoded on import. */
vfx_status vfx_export_distributions(const vfx_lattice* lat, int32_t is_even_timestep, void* out, int32_t precision,
                                    vfx_stream stream);
vfx_status vfx_import_distributions(const vfx_lattice* lat, int32_t is_even_timestep, const void* in,
                                    int32_t precision, vfx_stream stream);

/* --------------------------------------------------------- diagnostics */
/* calculateMacroscopicQuantitiesCompressible (src/gpu/core/PostProcessor/MacroscopicQuantities.cu:194-237,
 * run before every output, Simulation::run): rho (density deviation), vx, vy,
 * vz and pressure of every node in the lattice's real type (double for VFX_FP64,
 * float otherwise), n_nodes each, in the reference's arithmetic
 * (vf::lbm::getDensity, getCompressibleVelocityX1..3, getPressure,
 * src/lbm/MacroscopicQuantities.h:55-149).  Every node is zeroed first; the
 * fluid nodes -- direct: not solid; sparse: isValidFluidNode of
 * type_of_grid_node (types 19, 5, 6, 7), or the fluid_nodes list when no types
 * are given -- get their values. */
vfx_status vfx_macroscopic_quantities(const vfx_lattice* lat, int32_t is_even_timestep, void* rho, void* vx, void* vy,
                                      void* vz, void* pressure, vfx_stream stream);

/* calculateMeanCompressible (src/gpu/core/PostProcessor/MacroscopicQuantities.cu:390-536,
 * :975-984), called by Simulation::run at every step in [TimeStartCalcMean,
 * TimeEndCalcMean] (src/gpu/core/Calculation/Simulation.cpp:448-470): adds the
 * density deviation and velocity of every fluid node to the running sums
 * rho, vx, vy, vz (lattice real, n_nodes each), and to pressure the
 * pressure expression evaluated with the new density and velocity sums, as the
 * reference kernel does.  Nodes that are not fluid (direct: solid; sparse:
 * type_of_grid_node != GEO_FLUID) get zero sums; a sparse lattice without
 * type_of_grid_node updates the fluid_nodes only.  The mean is sum / tdiff
 * (calcMean, src/gpu/core/PostProcessor/CalcMean.cpp:57-67), taken on the host. */
vfx_status vfx_calculate_mean(const vfx_lattice* lat, int32_t is_even_timestep, void* vx, void* vy, void* vz,
                              void* rho, void* pressure, vfx_stream stream);
/* ResetMeanValuesSP27 (MacroscopicQuantities.cu:772-796): zero the five sums. */
vfx_status vfx_reset_mean(const vfx_lattice* lat, void* vx, void* vy, void* vz, void* rho, void* pressure,
                          vfx_stream stream);

/* Macroscopic fields of the pre-collision state (calculateMacroscopicQuantitiesCompressible_device,
 * src/gpu/core/PostProcessor/MacroscopicQuantities.cu:194-237), doubles, n_nodes each. */
vfx_status vfx_macroscopic(const vfx_lattice* lat, int32_t is_even_timestep, double* drho, double* vx,
                           double* vy, double* vz, vfx_stream stream);

/* FP64 reductions (deterministic two-pass: per-block partials in `workspace`,
 * then one fixed-order pass).  `workspace` needs vfx_reduce_workspace_bytes()
 * bytes of device memory; results are written to device memory. */
uint64_t vfx_reduce_workspace_bytes(const vfx_lattice* lat);

/* out4 = {sum drho, sum j_x, sum j_y, sum j_z} over the colliding nodes of the
 * pre-collision state: the global mass deviation and momentum.  The
 * reference reduces with thrust (KineticEnergyAnalyzer.cu:99-100); this is the
 * FP64 mass/momentum conservation monitor of the stepper. */
vfx_status vfx_reduce_mass_momentum(const vfx_lattice* lat, int32_t is_even_timestep, double* workspace, double* out4,
                                    vfx_stream stream);

/* The same reduction without a workspace and without any allocation: the
 * per-block partials live in a 2 MiB device array of the library (one per
 * GPU, held until the process ends), so the call suits callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference, which allocates through thrust.  At most
 * vfx_reduce_noalloc_max_items() work items (colliding-node candidates: n_nodes
 * for direct addressing or a type_of_grid_node selection, else n_fluid);
 * larger lattices get VFX_EINVAL.  Calls that overlap on one GPU (two streams or
 * two host threads) share the array and race.  Bit-identical to
 * vfx_reduce_mass_momentum. */
uint64_t vfx_reduce_noalloc_max_items(void);
vfx_status vfx_reduce_mass_momentum_noalloc(const vfx_lattice* lat, int32_t is_even_timestep, double* out4,
                                            vfx_stream stream);

/* KineticEnergyAnalyzer::run (src/gpu/core/PostProcessor/KineticEnergyAnalyzer.cu:66-154) and
 * EnstrophyAnalyzer::run (EnstrophyAnalyzer.cu:64-336) on macroscopic fields from
 * vfx_macroscopic: out2 = {sum, number of fluid nodes with a fluid 7-point stencil};
 * the analyzers' value is out2[0] / out2[1].  Structured (direct) lattices only. */
vfx_status vfx_kinetic_energy(const vfx_lattice* lat, const double* drho, const double* vx, const double* vy,
                              const double* vz, double* workspace, double* out2, vfx_stream stream);
vfx_status vfx_enstrophy(const vfx_lattice* lat, const double* drho, const double* vx, const double* vy,
                         const double* vz, double* workspace, double* out2, vfx_stream stream);

/* CalcTurbulenceIntensity (src/gpu/core/PostProcessor/TurbulenceIntensity.cu:45-85):
 * adds the current velocity to FP64 running sums, sums9 = 9 planes of n_nodes:
 * u_x, u_y, u_z, u_x^2, u_y^2, u_z^2, u_x u_y, u_x u_z, u_y u_z.  After T samples the
 * time-averaged mean is sums/T and the Reynolds stress sums_ij/T - mean_i mean_j. */
vfx_status vfx_accumulate_stats(const vfx_lattice* lat, int32_t is_even_timestep, double* sums9, vfx_stream stream);

/* ---------------------------------------------------- batched domains */
/* ---------------------------------------------------------- refinement --
 * Interface cells between a coarse level and its 2:1 fine level: ICells and
 * ICellNeigh of src/gpu/core/Calculation/Calculation.h:72-89.  A cell is the
 * 2x2x2 block of source nodes whose corner with the smallest coordinates
 * (MMM) is listed; the other seven follow through +x, +y, +z neighbours
 * (neighbor_x/y/z, or the box for direct addressing).
 *   coarse to fine: coarse_cells = MMM corner of the coarse source cell,
 *                   fine_cells   = MMM corner of the 8 fine nodes written;
 *   fine to coarse: fine_cells   = MMM corner of the fine source cell,
 *                   coarse_cells = the coarse node written (at the cell centre).
 * offset_x/y/z (may be NULL) shift the interpolation point for cells next to
 * walls, in source-cell units. */
typedef struct {
    const uint32_t* fine_cells;   /* ICells::fineCellIndices */
    const uint32_t* coarse_cells; /* ICells::coarseCellIndices */
    uint32_t n_cells;             /* ICells::numberOfCells */
    const void* offset_x;         /* ICellNeigh x/y/z: real[n_cells] (double for VFX_FP64, float otherwise) */
    const void* offset_y;
    const void* offset_z;
} vfx_icells;

/* scaleCoarseToFineCompressible (src/gpu/core/GridScaling/Scaling.cu:105-138,
 * kernel scaleCoarseToFineCompressible.cu:305-529): coarse pre-collision
 * populations at the coarse parity -> the 8 fine nodes of every cell,
 * pre-collision at the fine parity (the reference passes the coarse level's
 * isEvenTimestep and writes the fine level as even).  Both lattices share one
 * storage form; omega of each lattice is its level's rate.  With a turbulence
 * model on both levels (hasTurbulentViscosity, GridScalingFactory) every source
 * corner and every destination node takes omega / (1 + 3 omega nu_t) with nu_t
 * of its own node and level (scaleCoarseToFineCompressible.cu:94, 369). */
vfx_status vfx_scale_cf(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells,
                        int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream);
/* scaleFineToCoarseCompressible (Scaling.cu:155-188, kernel
 * scaleFineToCoarseCompressible.cu:76-124): fine pre-collision populations at
 * the fine parity (the reference reads the fine level as even) -> the coarse
 * node of every cell, pre-collision at the coarse parity.  Turbulence models:
 * per-node omega as in vfx_scale_cf (scaleFineToCoarseCompressible.cu:64). */
vfx_status vfx_scale_fc(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells,
                        int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream);

/* vfx_scale_fc then vfx_scale_cf of one nest (RefinementStrategy.cpp:169-178)
 * in one launch.  Needs the nest layout's separation: no coarse node the
 * fine-to-coarse cells write is a corner of a coarse-to-fine cell, and no fine
 * node the coarse-to-fine cells write is a corner of a fine-to-coarse cell;
 * then the values equal the two launches in order. */
vfx_status vfx_scale_fc_cf(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* fc,
                           const vfx_icells* cf, int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream);

/* An ensemble of independent box domains (for example many lid-driven
 * cavities) advanced by one launch from a per-domain work list.  Every
 * domain is a box of dims[0]*dims[1]*dims[2] nodes with its own 27 planes at
 * `f_offset` elements into a caller-owned pool; the outer layer is wall and
 * every wall link is closed by moving-wall halfway bounce-back
 * (VelocityBounceBack_Device, Velocity/VelocityBounceBack.cu:159-184) fused into
 * the collide-stream kernel; links into the z = dims[2]-1 plane move with
 * lid_velocity.  Each domain has its own relaxation rate.  No counterpart
 * exists in the reference, which launches one kernel per level and block. */
typedef struct vfx_domain {
    uint64_t f_offset;
    uint32_t dims[3];
    double omega;
    double lid_velocity[3];
} vfx_domain;

typedef struct vfx_batch_s* vfx_batch;
typedef struct vfx_graph_s* vfx_graph;

/* pool holds pool_elements populations of the storage type; every domain's 27
 * planes [f_offset, f_offset + 27*nx*ny*nz) must lie inside it and must not
 * overlap another domain's */
vfx_status vfx_batch_create(void* pool, uint64_t pool_elements, int32_t storage, float fp16_scale,
                            const vfx_domain* domains, uint32_t n_domains, const double* quadric_limiter,
                            int32_t kernel_variant, vfx_batch* out);
vfx_status vfx_batch_destroy(vfx_batch batch);
/* zero the pool (rest state f - w = 0 in every storage form) */
vfx_status vfx_batch_init_rest(vfx_batch batch, vfx_stream stream);
/* launch shape: block 64/128/256 nodes per workgroup; grid 0 = one workgroup per
 * tile, otherwise a persistent grid that strides over the tiles */
vfx_status vfx_batch_set_launch(vfx_batch batch, uint32_t block, uint32_t grid);
vfx_status vfx_batch_get_launch(vfx_batch batch, uint32_t* block, uint32_t* grid);
/* tile order: workgroup b starts at tile xcd_block(b), which gives each of the 8
 * XCDs runs of `run` consecutive tiles (tiles next to each other share cache
 * lines, and every XCD has its own L2); 1 = plain order; 0 = slabs: every
 * domain's tiles split into 8 contiguous slabs, slab x on XCD x (the tiles a
 * node's y and z neighbours sit in then share its L2).  The per-domain step
 * always runs its tiles in plain order. */
vfx_status vfx_batch_set_xcd_run(vfx_batch batch, uint32_t run);
vfx_status vfx_batch_get_xcd_run(vfx_batch batch, uint32_t* run);
/* times every launch shape (block, grid, XCD run of 1 / 4 / 16 tiles or slabs) on
 * `steps` (even) steps from parity is_even_timestep and keeps the fastest; the
 * state advances, its parity does not change */
vfx_status vfx_batch_autotune(vfx_batch batch, int32_t is_even_timestep, int32_t steps, vfx_stream stream,
                              uint32_t* block, uint32_t* grid, float* ms_per_step);
/* one step of every domain in one launch / of one domain (the per-domain baseline) */
vfx_status vfx_batch_step(vfx_batch batch, int32_t is_even_timestep, vfx_stream stream);
vfx_status vfx_batch_step_domain(vfx_batch batch, uint32_t domain, int32_t is_even_timestep, vfx_stream stream);
/* a vfx_lattice view of one domain (direct addressing over its box) for the
 * population and diagnostic entry points */
vfx_status vfx_batch_lattice(vfx_batch batch, uint32_t domain, vfx_lattice* out);
/* n_steps steps captured into a HIP graph (per_domain = 1: as one launch per domain) */
vfx_status vfx_batch_capture(vfx_batch batch, int32_t n_steps, int32_t is_even_start, int32_t per_domain, vfx_graph* out);
vfx_status vfx_graph_launch(vfx_graph graph, vfx_stream stream);
vfx_status vfx_graph_destroy(vfx_graph graph);

/* General capture: every libvfx call issued on `stream` between begin and end
 * is recorded (not executed) into a graph, e.g. one refined coarse step (two
 * fine steps, the coarse step, fine-to-coarse, coarse-to-fine).  The stream
 * must not be the legacy NULL stream; vfx_stream_create gives a non-blocking
 * stream of default priority.  No libvfx entry point records a memset node. */
vfx_status vfx_stream_create(vfx_stream* out);
vfx_status vfx_stream_destroy(vfx_stream stream);
vfx_status vfx_capture_begin(vfx_stream stream);
vfx_status vfx_capture_end(vfx_stream stream, vfx_graph* out);

/* -------------------------------------------------------------- misc */
/* Device memory and synchronisation for C / C++ callers without their own HIP
 * runtime code (the host layer of virtualfluids_amd uses these). */
vfx_status vfx_device_alloc(uint64_t bytes, void** ptr);
vfx_status vfx_device_free(void* ptr);
vfx_status vfx_copy_to_device(void* dst, const void* src, uint64_t bytes);
vfx_status vfx_copy_to_host(void* dst, const void* src, uint64_t bytes);
vfx_status vfx_synchronize(vfx_stream stream);

const char* vfx_last_error(void);
int32_t vfx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VFX_H */
