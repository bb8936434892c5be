// api.hip -- the extern "C" surface declared in include/vfx.h.
//
// Validates every argument on the host before anything is launched (a kernel
// is never started with shapes it does not expect), translates vfx_lattice into
// the device view and dispatches on the storage type.  Errors are returned as
// vfx_status; the message is kept per host thread for vfx_last_error().
#include "kernels.hpp"

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <mutex>
#include <vector>

namespace vfx {
// the kernels live in step_fp64.hip / step_fp32.hip (built with their own FP flags)
VFX_EXTERN_LAUNCHERS(double)
VFX_EXTERN_LAUNCHERS(float)
VFX_EXTERN_LAUNCHERS(_Float16)
}

namespace {

thread_local std::string g_err;

vfx_status fail(vfx_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
vfx_status fail(vfx_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

vfx_status hip_status(hipError_t e, const char* what)
{
    if (e == hipSuccess) return VFX_OK;
    return fail(VFX_EHIP, "%s: %s", what, hipGetErrorString(e));
}

constexpr double kW[4] = {8. / 27., 2. / 27., 1. / 54., 1. / 216.};

vfx_status make_lat(const vfx_lattice* in, vfx::Lat& L)
{
    if (!in) return fail(VFX_EINVAL, "lattice is NULL");
    if (!in->f) return fail(VFX_EINVAL, "lattice.f is NULL");
    if (in->storage < VFX_FP64 || in->storage > VFX_FP16_SCALED)
        return fail(VFX_EINVAL, "unknown storage %d", in->storage);
    if (in->n_nodes == 0) return fail(VFX_EINVAL, "n_nodes is 0");
    if (in->n_nodes >= (1ull << 32)) return fail(VFX_EUNSUPPORTED, "n_nodes %llu needs 64-bit node indices",
                                                 (unsigned long long)in->n_nodes);
    if (!(in->omega > 0.0 && in->omega < 2.0)) return fail(VFX_EINVAL, "omega %g outside (0,2)", in->omega);
    const uint64_t esz = in->storage == VFX_FP64 ? 8 : in->storage == VFX_FP32 ? 4 : 2;
    if (in->n_nodes * esz >= (1ull << 32))
        return fail(VFX_EUNSUPPORTED, "a plane of %llu nodes exceeds 4 GiB (32-bit in-plane offsets)",
                    (unsigned long long)in->n_nodes);
    std::memset(&L, 0, sizeof(L));
    L.f = in->f;
    L.N = in->n_nodes;
    if (in->xcd_run > 4096 || (in->xcd_run & (in->xcd_run - 1)))
        return fail(VFX_EINVAL, "xcd_run must be 0 or a power of two up to 4096");
    L.xrun_lg = in->xcd_run ? 1u + (uint32_t)__builtin_ctz(in->xcd_run) : 0u;
    if (in->direct) {
        const uint64_t nx = in->dims[0], ny = in->dims[1], nz = in->dims[2];
        if (nx == 0 || ny == 0 || nz == 0) return fail(VFX_EINVAL, "direct addressing needs dims > 0");
        if (nx * ny * nz != in->n_nodes)
            return fail(VFX_EINVAL, "dims %llu*%llu*%llu != n_nodes %llu", (unsigned long long)nx,
                        (unsigned long long)ny, (unsigned long long)nz, (unsigned long long)in->n_nodes);
        const uint64_t tile = in->tile_nodes ? in->tile_nodes : 256;
        if (tile != 64 && tile != 128 && tile != 256) return fail(VFX_EINVAL, "tile_nodes must be 0, 64, 128 or 256");
        if (ny * nz >= (1ull << 31) / ((nx + tile - 1) / tile)) return fail(VFX_EUNSUPPORTED, "grid too large");
        L.nx = (uint32_t)nx;
        L.ny = (uint32_t)ny;
        L.nz = (uint32_t)nz;
        L.chunks = (uint32_t)((nx + tile - 1) / tile);
        L.tile = (uint32_t)tile;
        L.solid = in->solid_mask;
    } else {
        if (!in->neighbor_x || !in->neighbor_y || !in->neighbor_z)
            return fail(VFX_EINVAL, "sparse addressing needs neighbor_x/y/z");
        if (in->tile_nodes != 0 && in->tile_nodes != 256)
            return fail(VFX_EINVAL, "tile_nodes of the sparse step must be 0 or 256");
        if (in->n_fluid && !in->fluid_nodes && !in->type_of_grid_node) return fail(VFX_EINVAL, "fluid_nodes is NULL");
        L.nbx = in->neighbor_x;
        L.nby = in->neighbor_y;
        L.nbz = in->neighbor_z;
        L.fluid = in->fluid_nodes;
        L.nfluid = in->n_fluid;
        L.types = in->type_of_grid_node;
        L.rgroups = (const uint4*)in->regular_groups;
    }
    if (in->direct && in->regular_groups)
        return fail(VFX_EINVAL, "regular_groups belongs to the sparse layout");
    if (in->direct && in->type_of_grid_node)
        return fail(VFX_EINVAL, "type_of_grid_node selects nodes of the sparse layout; direct addressing uses solid_mask");
    if (in->storage == VFX_FP16_SCALED) {
        if (!(in->fp16_scale > 0.f)) return fail(VFX_EINVAL, "fp16_scale must be > 0");
        for (int c = 0; c < 4; ++c) {
            L.ld[c] = (float)(kW[c] / (double)in->fp16_scale);
            L.st[c] = (float)((double)in->fp16_scale / kW[c]);
        }
    } else {
        for (int c = 0; c < 4; ++c) L.ld[c] = L.st[c] = 1.f;
    }
    const double* ur = in->fp16_reference_velocity;
    const bool shifted = ur[0] != 0.0 || ur[1] != 0.0 || ur[2] != 0.0;
    if (shifted && in->storage != VFX_FP16_SCALED)
        return fail(VFX_EINVAL, "fp16_reference_velocity needs VFX_FP16_SCALED storage");
    if (!(std::fabs(ur[0]) < 0.5 && std::fabs(ur[1]) < 0.5 && std::fabs(ur[2]) < 0.5))
        return fail(VFX_EINVAL, "fp16_reference_velocity out of range");
    for (int d = 0; d < 27; ++d) {
        // shifted compressible equilibrium at drho = 0 (the offset only has to
        // be a fixed per-direction constant; this one is close to the flow)
        const double cu = vfx::cx(d) * ur[0] + vfx::cy(d) * ur[1] + vfx::cz(d) * ur[2];
        const double u2 = ur[0] * ur[0] + ur[1] * ur[1] + ur[2] * ur[2];
        const double sh = shifted ? kW[vfx::weight_class(d)] * (3.0 * cu + 4.5 * cu * cu - 1.5 * u2) : 0.0;
        L.sh[d] = (float)sh;
        L.ss[d] = (float)(-sh * (double)L.st[vfx::weight_class(d)]);
    }
    L.rho = in->rho;
    L.vx = in->vx;
    L.vy = in->vy;
    L.vz = in->vz;
    if (in->turbulence_model < VFX_TURB_NONE || in->turbulence_model > VFX_TURB_QR)
        return fail(VFX_EINVAL, "unknown turbulence_model %d", in->turbulence_model);
    // the turbulent viscosity is read where a model is active (collision,
    // refinement interpolation)
    if (in->turbulence_model != VFX_TURB_NONE && !in->turbulent_viscosity)
        return fail(VFX_EINVAL, "turbulence_model %d needs turbulent_viscosity", in->turbulence_model);
    L.nut = in->turbulence_model != VFX_TURB_NONE ? in->turbulent_viscosity : nullptr;
    L.nbinv = in->direct ? nullptr : in->neighbor_inverse;
    return VFX_OK;
}

} // namespace

extern "C" {

const char* vfx_last_error(void) { return g_err.c_str(); }

vfx_status vfx_device_alloc(uint64_t bytes, void** ptr)
{
    if (!ptr) return fail(VFX_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return VFX_OK;
    return hip_status(hipMalloc(ptr, bytes), "hipMalloc");
}

vfx_status vfx_device_free(void* ptr) { return hip_status(hipFree(ptr), "hipFree"); }

vfx_status vfx_copy_to_device(void* dst, const void* src, uint64_t bytes)
{
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

vfx_status vfx_copy_to_host(void* dst, const void* src, uint64_t bytes)
{
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

vfx_status vfx_synchronize(vfx_stream stream) { return hip_status(hipStreamSynchronize((hipStream_t)stream), "sync"); }

int32_t vfx_version(void) { return 1; }

vfx_status vfx_collide_stream(const vfx_lattice* lat, int32_t collision, int32_t is_even, int32_t tmpl, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (collision != VFX_K17 && collision != VFX_BGK) return fail(VFX_EINVAL, "unknown collision %d", collision);
    if (lat->kernel_variant != 0 && lat->kernel_variant != 1)
        return fail(VFX_EINVAL, "unknown kernel_variant %d", lat->kernel_variant);
    if (tmpl < VFX_TEMPLATE_DEFAULT || tmpl > VFX_TEMPLATE_ALL_FEATURES)
        return fail(VFX_EUNSUPPORTED, "collision template %d not supported", tmpl);
    const bool body_force = tmpl == VFX_TEMPLATE_APPLY_BODY_FORCE || tmpl == VFX_TEMPLATE_ALL_FEATURES;
    if (body_force) {
        if (collision != VFX_K17) return fail(VFX_EUNSUPPORTED, "body force templates belong to the K17 kernel");
        if (!lat->body_force[0] || !lat->body_force[1] || !lat->body_force[2])
            return fail(VFX_EINVAL, "ApplyBodyForce needs body_force");
        if (!(lat->force_factor > 0.0)) return fail(VFX_EINVAL, "ApplyBodyForce needs force_factor > 0");
        for (int a = 0; a < 3; ++a) {
            L.bf[a] = lat->body_force[a];
            L.gforce[a] = lat->forces[a];
        }
        L.force_factor = lat->force_factor;
    }
    // the eddy-viscosity models belong to the K17 collision (BGK has none)
    const int turb = collision == VFX_K17 ? lat->turbulence_model : VFX_TURB_NONE;
    // runCollision writes the macroscopic values for the AMD model (RunCollision.cuh:133-138)
    if (tmpl == VFX_TEMPLATE_WRITE_MACRO || tmpl == VFX_TEMPLATE_ALL_FEATURES || turb == VFX_TURB_AMD) {
        if (!lat->rho || !lat->vx || !lat->vy || !lat->vz)
            return fail(VFX_EINVAL, turb == VFX_TURB_AMD ? "the AMD model needs rho/vx/vy/vz"
                                                         : "WriteMacroVars needs rho/vx/vy/vz");
        L.write_macro = 1;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    const bool even = is_even != 0;
    const int coll = collision | (body_force ? vfx::kCollBodyForce : 0);
    switch (lat->storage) {
    case VFX_FP64:
        e = vfx::launch_step<double>(L, lat->direct, coll, even, lat->kernel_variant, lat->omega,
                                     lat->quadric_limiter, lat->force, turb, lat->sgs_constant, st);
        break;
    case VFX_FP32:
        e = vfx::launch_step<float>(L, lat->direct, coll, even, lat->kernel_variant, lat->omega,
                                    lat->quadric_limiter, lat->force, turb, lat->sgs_constant, st);
        break;
    default:
        e = vfx::launch_step<_Float16>(L, lat->direct, coll, even, lat->kernel_variant, lat->omega,
                                       lat->quadric_limiter, lat->force, turb, lat->sgs_constant, st);
        break;
    }
    return hip_status(e, "collide_stream launch");
}

vfx_status vfx_turbulent_viscosity_amd(const vfx_lattice* lat, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!lat->turbulent_viscosity) return fail(VFX_EINVAL, "the AMD model needs turbulent_viscosity");
    if (!lat->vx || !lat->vy || !lat->vz) return fail(VFX_EINVAL, "the AMD model needs vx/vy/vz");
    L.nut = lat->turbulent_viscosity;
    if (!lat->direct && (!lat->type_of_grid_node || !lat->neighbor_inverse))
        return fail(VFX_EINVAL, "the AMD model on the sparse layout needs type_of_grid_node and neighbor_inverse");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_turb_amd<double>(L, lat->direct, lat->sgs_constant, st); break;
    case VFX_FP32: e = vfx::launch_turb_amd<float>(L, lat->direct, lat->sgs_constant, st); break;
    default: e = vfx::launch_turb_amd<_Float16>(L, lat->direct, lat->sgs_constant, st); break;
    }
    return hip_status(e, "turbulent_viscosity_amd launch");
}

vfx_status vfx_k17_step(const vfx_lattice* lat, int32_t is_even, int32_t tmpl, vfx_stream stream)
{
    return vfx_collide_stream(lat, VFX_K17, is_even, tmpl, stream);
}

vfx_status vfx_bgk_step(const vfx_lattice* lat, int32_t is_even, vfx_stream stream)
{
    return vfx_collide_stream(lat, VFX_BGK, is_even, VFX_TEMPLATE_DEFAULT, stream);
}

uint64_t vfx_sparse_regular_groups_bytes(const vfx_lattice* lat)
{
    if (!lat || lat->direct) return 0;
    vfx::Lat L;
    std::memset(&L, 0, sizeof(L));
    L.N = lat->n_nodes;
    L.nfluid = lat->n_fluid;
    L.types = lat->type_of_grid_node;
    return vfx::sparse_groups(L) * 16u;
}

vfx_status vfx_sparse_regular_groups(const vfx_lattice* lat, void* table, uint64_t table_bytes, vfx_stream stream)
{
    if (!lat || !table) return fail(VFX_EINVAL, "NULL argument");
    if (lat->direct) return fail(VFX_EINVAL, "regular groups belong to the sparse layout");
    vfx_lattice plain = *lat;
    plain.regular_groups = nullptr;
    vfx::Lat L;
    if (vfx_status s = make_lat(&plain, L)) return s;
    const uint64_t groups = vfx::sparse_groups(L);
    if (table_bytes < groups * 16u)
        return fail(VFX_EINVAL, "table holds %llu bytes, %llu needed", (unsigned long long)table_bytes,
                    (unsigned long long)(groups * 16u));
    if (groups == 0) return VFX_OK;
    if (groups >= (1ull << 32)) return fail(VFX_EUNSUPPORTED, "too many wavefronts");
    return hip_status(vfx::launch_sparse_groups(L, (uint4*)table, (uint32_t)groups, (hipStream_t)stream),
                      "sparse groups launch");
}

vfx_status vfx_k17_autotune(const vfx_lattice* lat, int32_t is_even, int32_t tmpl, int32_t steps, vfx_stream stream,
                            uint32_t* tile_out, uint32_t* run_out, float* ms_out)
{
    if (!lat || !tile_out || !run_out) return fail(VFX_EINVAL, "NULL argument");
    {
        vfx::Lat L;
        if (vfx_status s = make_lat(lat, L)) return s;
    }
    if (steps < 2) steps = 2;
    steps += steps & 1;
    const bool even0 = is_even != 0;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(VFX_EHIP, "event create");
    }
    vfx_lattice trial = *lat;
    float best = 1e30f;
    uint32_t bt = 256, br = 0;
    vfx_status r = VFX_OK;
    // the sparse step's workgroups are 256 work items; only its XCD run varies
    const uint32_t tiles[3] = {256, 128, 64};
    const int ntiles = lat->direct ? 3 : 1;
    const uint32_t runs[7] = {1, 4, 8, 16, 32, 64, 256};
    for (int ti = 0; ti < ntiles; ++ti) {
        const uint32_t tile = tiles[ti];
        for (uint32_t run : runs) {
            trial.tile_nodes = tile;
            trial.xcd_run = run;
            for (int s = 0; s < 2 && r == VFX_OK; ++s) r = vfx_k17_step(&trial, (s % 2 == 0) == even0, tmpl, stream);
            if (r == VFX_OK) r = hip_status(hipEventRecord(e0, st), "event record");
            for (int s = 0; s < steps && r == VFX_OK; ++s) r = vfx_k17_step(&trial, (s % 2 == 0) == even0, tmpl, stream);
            if (r == VFX_OK) r = hip_status(hipEventRecord(e1, st), "event record");
            if (r == VFX_OK) r = hip_status(hipEventSynchronize(e1), "event synchronize");
            float ms = 0;
            if (r == VFX_OK) r = hip_status(hipEventElapsedTime(&ms, e0, e1), "event elapsed");
            if (r != VFX_OK) break;
            ms /= (float)steps;
            if (ms < best) {
                best = ms;
                bt = tile;
                br = run;
            }
        }
        if (r != VFX_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != VFX_OK) return r;
    *tile_out = bt;
    *run_out = br;
    if (ms_out) *ms_out = best;
    return VFX_OK;
}

// argument checks of a boundary table and its device view
static vfx_status make_bcdev(const vfx_lattice* lat, const vfx::Lat& L, const vfx_bc* bc, vfx::BCDev& B)
{
    if (!bc) return fail(VFX_EINVAL, "bc is NULL");
    if (bc->kind < VFX_BC_VELOCITY_BOUNCE_BACK || bc->kind > VFX_BC_PRECURSOR_DISTRIBUTIONS)
        return fail(VFX_EUNSUPPORTED, "bc kind %d not supported", bc->kind);
    if (bc->kind == VFX_BC_SLIP_TURBULENT_VISCOSITY && !L.nut)
        return fail(VFX_EINVAL, "slip with turbulent viscosity needs a turbulence model and turbulent_viscosity");
    const bool outflow = bc->kind == VFX_BC_OUTFLOW_NON_REFLECTING ||
                         bc->kind == VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION;
    if (bc->kind == VFX_BC_PRESSURE_NEQ) {
        if (bc->n && (!bc->k || !bc->k_neighbor || !bc->rho)) return fail(VFX_EINVAL, "pressure bc needs k, k_neighbor, rho");
        if (bc->normal < 1 || bc->normal > 26) return fail(VFX_EINVAL, "pressure bc normal %d invalid", bc->normal);
    } else if (bc->kind == VFX_BC_PRECURSOR_VELOCITY || bc->kind == VFX_BC_PRECURSOR_DISTRIBUTIONS) {
        if (bc->n && (!bc->k || (bc->kind == VFX_BC_PRECURSOR_VELOCITY && !bc->q27)))
            return fail(VFX_EINVAL, "precursor bc needs k%s", bc->kind == VFX_BC_PRECURSOR_VELOCITY ? " and q27" : "");
        if (bc->n && (!bc->plane_neighbor[0] || !bc->plane_neighbor[1] || !bc->plane_neighbor[2] ||
                      !bc->plane_neighbor[3] || !bc->plane_weight[0] || !bc->plane_weight[1] || !bc->plane_weight[2] ||
                      !bc->plane_weight[3] || !bc->precursor_last || !bc->precursor_current || !bc->n_precursor_nodes))
            return fail(VFX_EINVAL, "precursor bc needs plane neighbours, weights, last and current planes");
        if (bc->kind == VFX_BC_PRECURSOR_VELOCITY && !(bc->velocity_ratio != 0.0))
            return fail(VFX_EINVAL, "precursor velocity bc needs a non-zero velocity_ratio");
    } else if (bc->kind >= VFX_BC_STRESS_INTERPOLATED) {
        if (bc->n && (!bc->k || !bc->q27 || !bc->k_neighbor)) return fail(VFX_EINVAL, "stress bc needs k, q27, k_neighbor");
        if (bc->n && (!bc->wm_normal[0] || !bc->wm_normal[1] || !bc->wm_normal[2] || !bc->wm_exchange_velocity[0] ||
                      !bc->wm_exchange_velocity[1] || !bc->wm_exchange_velocity[2] || !bc->wm_wall_velocity[0] ||
                      !bc->wm_wall_velocity[1] || !bc->wm_wall_velocity[2] || !bc->wm_sampling_offset || !bc->wm_z0))
            return fail(VFX_EINVAL, "stress bc needs the wall-model arrays");
        if (bc->wm_force[0] && (!bc->wm_force[1] || !bc->wm_force[2]))
            return fail(VFX_EINVAL, "stress bc monitors need all three force arrays");
        if (!lat->vx || !lat->vy) return fail(VFX_EINVAL, "stress bc needs the lattice's velocity arrays");
    } else if (outflow) {
        if (bc->n && (!bc->k || !bc->k_neighbor)) return fail(VFX_EINVAL, "outflow bc needs k, k_neighbor");
        if (bc->normal < 0 || bc->normal > 26) return fail(VFX_EINVAL, "outflow bc normal %d invalid", bc->normal);
    } else if (bc->n && (!bc->k || !bc->q27)) {
        return fail(VFX_EINVAL, "bc.k / bc.q27 is NULL");
    }
    B = vfx::BCDev{bc->n, bc->k, bc->q27, bc->vx, bc->vy, bc->vz, bc->cpu_float_velocity, lat->direct,
                 lat->omega, bc->rho, bc->k_neighbor, bc->normal, bc->pressure_correction_factor,
                 {bc->wm_normal[0], bc->wm_normal[1], bc->wm_normal[2]},
                 {bc->wm_exchange_velocity[0], bc->wm_exchange_velocity[1], bc->wm_exchange_velocity[2]},
                 {bc->wm_wall_velocity[0], bc->wm_wall_velocity[1], bc->wm_wall_velocity[2]},
                 bc->wm_sampling_offset, bc->wm_z0, bc->wm_u_star,
                 {bc->wm_force[0], bc->wm_force[1], bc->wm_force[2]},
                 bc->n_precursor_nodes,
                 {bc->plane_neighbor[0], bc->plane_neighbor[1], bc->plane_neighbor[2], bc->plane_neighbor[3]},
                 {bc->plane_weight[0], bc->plane_weight[1], bc->plane_weight[2], bc->plane_weight[3]},
                 bc->precursor_last, bc->precursor_current, bc->time_ratio, bc->velocity_ratio,
                 {bc->precursor_velocity[0], bc->precursor_velocity[1], bc->precursor_velocity[2]}};
    return VFX_OK;
}

vfx_status vfx_bc_apply(const vfx_lattice* lat, const vfx_bc* bc, int32_t is_even, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    vfx::BCDev B;
    if (vfx_status s = make_bcdev(lat, L, bc, B)) return s;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_bc<double>(L, B, bc->kind, is_even != 0, st); break;
    case VFX_FP32: e = vfx::launch_bc<float>(L, B, bc->kind, is_even != 0, st); break;
    default: e = vfx::launch_bc<_Float16>(L, B, bc->kind, is_even != 0, st); break;
    }
    return hip_status(e, "bc launch");
}

} // extern "C"

struct vfx_bc_set_s {
    int32_t storage = 0;
    uint32_t n = 0, blocks = 0;
    vfx::BCJob* dev = nullptr;
};

extern "C" {

vfx_status vfx_bc_set_create(const vfx_lattice* const* lattices, const vfx_bc* const* bcs, int32_t n, vfx_bc_set* out)
{
    if (!out) return fail(VFX_EINVAL, "out is NULL");
    *out = nullptr;
    if (n < 1 || n > vfx::kMaxBCJobs || !lattices || !bcs)
        return fail(VFX_EINVAL, "a boundary set holds 1..%d groups", vfx::kMaxBCJobs);
    std::vector<vfx::BCJob> jobs(n);
    uint32_t blocks = 0;
    for (int j = 0; j < n; ++j) {
        if (!lattices[j] || !bcs[j]) return fail(VFX_EINVAL, "group %d: NULL lattice or boundary", j);
        if (lattices[j]->storage != lattices[0]->storage)
            return fail(VFX_EINVAL, "group %d: every lattice of a set has one storage type", j);
        const int32_t kind = bcs[j]->kind;
        if (kind != VFX_BC_VELOCITY_BOUNCE_BACK && kind != VFX_BC_NOSLIP_INTERPOLATED &&
            kind != VFX_BC_VELOCITY_INTERPOLATED && kind != VFX_BC_SLIP && kind != VFX_BC_PRESSURE_NEQ &&
            kind != VFX_BC_VELOCITY_PRESSURE_INTERPOLATED)
            return fail(VFX_EUNSUPPORTED, "group %d: kind %d in a set (kinds 0-5 only)", j, kind);
        vfx::BCJob& J = jobs[j];
        std::memset(&J, 0, sizeof(J));
        if (vfx_status s = make_lat(lattices[j], J.L)) return s;
        if (vfx_status s = make_bcdev(lattices[j], J.L, bcs[j], J.B)) return s;
        J.kind = kind;
        J.first_block = blocks;
        blocks += (bcs[j]->n + vfx::kBlock - 1) / vfx::kBlock;
    }
    vfx_bc_set_s* set = new vfx_bc_set_s;
    set->storage = lattices[0]->storage;
    set->n = (uint32_t)n;
    set->blocks = blocks;
    hipError_t e = hipMalloc((void**)&set->dev, jobs.size() * sizeof(vfx::BCJob));
    if (e == hipSuccess) e = hipMemcpy(set->dev, jobs.data(), jobs.size() * sizeof(vfx::BCJob), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (set->dev) (void)hipFree(set->dev);
        delete set;
        return hip_status(e, "boundary set tables");
    }
    *out = set;
    return VFX_OK;
}

vfx_status vfx_bc_set_apply(vfx_bc_set set, uint32_t even_mask, vfx_stream stream)
{
    if (!set) return fail(VFX_EINVAL, "set is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (set->storage) {
    case VFX_FP64: e = vfx::launch_bc_set<double>(set->dev, set->n, set->blocks, even_mask, st); break;
    case VFX_FP32: e = vfx::launch_bc_set<float>(set->dev, set->n, set->blocks, even_mask, st); break;
    default: e = vfx::launch_bc_set<_Float16>(set->dev, set->n, set->blocks, even_mask, st); break;
    }
    return hip_status(e, "boundary set launch");
}

vfx_status vfx_bc_set_destroy(vfx_bc_set set)
{
    if (!set) return VFX_OK;
    if (set->dev) (void)hipFree(set->dev);
    delete set;
    return VFX_OK;
}

} // extern "C"

namespace {

// the lattice and fused groups of one fused step, validated
vfx_status fused_inputs(const vfx_lattice* lat, const vfx_fused_bc* post, int32_t n_post, const vfx_fused_bc* pre,
                        vfx::Lat& L, vfx::FusedGroup (&g)[3])
{
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!lat->direct) return fail(VFX_EUNSUPPORTED, "fused boundary conditions need structured (direct) addressing");
    if (lat->kernel_variant != 0) return fail(VFX_EUNSUPPORTED, "fused boundary conditions run the default K17 variant");
    if (lat->turbulence_model != VFX_TURB_NONE)
        return fail(VFX_EUNSUPPORTED, "fused boundary conditions run K17 without a turbulence model");
    if (n_post < 0 || n_post > 2 || (n_post > 0 && !post)) return fail(VFX_EINVAL, "n_post must be 0, 1 or 2");
    std::memset(g, 0, sizeof(g));
    for (int j = 0; j < n_post + 1; ++j) {
        const vfx_fused_bc* f = j < n_post ? &post[j] : pre;
        if (!f) continue;
        if (!f->bc || !f->node_mask || !f->word_base) return fail(VFX_EINVAL, "fused bc needs bc, node_mask, word_base");
        const int32_t kind = f->bc->kind;
        if (j < n_post && kind != VFX_BC_NOSLIP_INTERPOLATED && kind != VFX_BC_VELOCITY_INTERPOLATED &&
            kind != VFX_BC_SLIP && kind != VFX_BC_VELOCITY_PRESSURE_INTERPOLATED)
            return fail(VFX_EUNSUPPORTED, "fused post-collision kind %d (kinds 1, 2, 3, 5 only)", kind);
        if (j == n_post && (kind != VFX_BC_PRESSURE_NEQ || (f->bc->normal != 1 && f->bc->normal != 2)))
            return fail(VFX_EUNSUPPORTED, "the fused pre-collision group is a pressure boundary with normal dP00 or dM00");
        vfx::FusedGroup& G = g[j < n_post ? j : 2];
        if (vfx_status s = make_bcdev(lat, L, f->bc, G.B)) return s;
        if (f->bc->n == 0) continue; // an empty group selects nothing
        G.mask = f->node_mask;
        G.base = f->word_base;
        G.kind = kind;
    }
    return VFX_OK;
}

} // namespace

extern "C" {

vfx_status vfx_k17_step_levels(const vfx_level_step* levels, int32_t n_levels, vfx_stream stream)
{
    if (!levels || n_levels < 1 || n_levels > 2) return fail(VFX_EINVAL, "one or two levels per launch");
    vfx::LevelInput in[2];
    for (int j = 0; j < n_levels; ++j) {
        const vfx_level_step& v = levels[j];
        const vfx_lattice* lat = v.lattice;
        if (!lat) return fail(VFX_EINVAL, "level %d: lattice is NULL", j);
        if (lat->storage != levels[0].lattice->storage) return fail(VFX_EINVAL, "levels differ in storage");
        if (vfx_status s = make_lat(lat, in[j].L)) return s;
        if (!lat->direct) return fail(VFX_EUNSUPPORTED, "level %d: one launch for levels needs structured addressing", j);
        if (lat->kernel_variant != 0 || lat->turbulence_model != VFX_TURB_NONE || lat->body_force[0] ||
            lat->body_force[1] || lat->body_force[2])
            return fail(VFX_EUNSUPPORTED,
                        "level %d: one launch for levels runs the default K17 without turbulence model or body-force arrays",
                        j);
        in[j].omega = lat->omega;
        in[j].lim = lat->quadric_limiter;
        in[j].force = lat->force;
        in[j].even = v.is_even_timestep != 0;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (levels[0].lattice->storage) {
    case VFX_FP64: e = vfx::launch_step_levels<double>(in, n_levels, st); break;
    case VFX_FP32: e = vfx::launch_step_levels<float>(in, n_levels, st); break;
    default: e = vfx::launch_step_levels<_Float16>(in, n_levels, st); break;
    }
    return hip_status(e, "level step launch");
}

vfx_status vfx_k17_step_fused(const vfx_lattice* lat, int32_t is_even, const vfx_fused_bc* post, int32_t n_post,
                              const vfx_fused_bc* pre, const uint32_t* any_mask, int32_t apply_pre, vfx_stream stream)
{
    vfx::Lat L;
    vfx::FusedGroup g[3];
    if (vfx_status s = fused_inputs(lat, post, n_post, pre, L, g)) return s;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    const int ap = apply_pre != 0;
    const double* lim = lat->quadric_limiter;
    switch (lat->storage) {
    case VFX_FP64:
        e = vfx::launch_step_fused<double>(L, is_even != 0, lat->omega, lim, lat->force, g[0], g[1], g[2], any_mask, ap, st);
        break;
    case VFX_FP32:
        e = vfx::launch_step_fused<float>(L, is_even != 0, lat->omega, lim, lat->force, g[0], g[1], g[2], any_mask, ap, st);
        break;
    default:
        e = vfx::launch_step_fused<_Float16>(L, is_even != 0, lat->omega, lim, lat->force, g[0], g[1], g[2], any_mask, ap,
                                             st);
        break;
    }
    return hip_status(e, "fused collide-stream launch");
}

vfx_status vfx_init_equilibrium(const vfx_lattice* lat, int32_t is_even, const double* drho, const double* ux,
                                const double* uy, const double* uz, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!drho || !ux || !uy || !uz) return fail(VFX_EINVAL, "init fields are NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_init<double>(L, lat->direct, is_even != 0, drho, ux, uy, uz, st); break;
    case VFX_FP32: e = vfx::launch_init<float>(L, lat->direct, is_even != 0, drho, ux, uy, uz, st); break;
    default: e = vfx::launch_init<_Float16>(L, lat->direct, is_even != 0, drho, ux, uy, uz, st); break;
    }
    return hip_status(e, "init launch");
}

static vfx_status rw(const vfx_lattice* lat, int32_t is_even, double* out, const double* in, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!out && !in) return fail(VFX_EINVAL, "population buffer is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_rw<double>(L, lat->direct, is_even != 0, out, in, st); break;
    case VFX_FP32: e = vfx::launch_rw<float>(L, lat->direct, is_even != 0, out, in, st); break;
    default: e = vfx::launch_rw<_Float16>(L, lat->direct, is_even != 0, out, in, st); break;
    }
    return hip_status(e, "population copy launch");
}

vfx_status vfx_read_populations(const vfx_lattice* lat, int32_t is_even, double* out, vfx_stream stream)
{
    if (!out) return fail(VFX_EINVAL, "out is NULL");
    return rw(lat, is_even, out, nullptr, stream);
}

vfx_status vfx_write_populations(const vfx_lattice* lat, int32_t is_even, const double* in, vfx_stream stream)
{
    if (!in) return fail(VFX_EINVAL, "in is NULL");
    return rw(lat, is_even, nullptr, in, stream);
}

static vfx_status vf_array(const vfx_lattice* lat, int32_t is_even, void* out, const void* in, int32_t precision,
                           vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (precision != VFX_FP32 && precision != VFX_FP64)
        return fail(VFX_EINVAL, "precision %d: the VirtualFluids array is float or double", precision);
    hipStream_t st = (hipStream_t)stream;
    const bool fp64 = precision == VFX_FP64;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_vf_array<double>(L, is_even != 0, fp64, out, in, st); break;
    case VFX_FP32: e = vfx::launch_vf_array<float>(L, is_even != 0, fp64, out, in, st); break;
    default: e = vfx::launch_vf_array<_Float16>(L, is_even != 0, fp64, out, in, st); break;
    }
    return hip_status(e, "distribution array copy launch");
}

vfx_status vfx_export_distributions(const vfx_lattice* lat, int32_t is_even, void* out, int32_t precision,
                                    vfx_stream stream)
{
    if (!out) return fail(VFX_EINVAL, "out is NULL");
    return vf_array(lat, is_even, out, nullptr, precision, stream);
}

vfx_status vfx_import_distributions(const vfx_lattice* lat, int32_t is_even, const void* in, int32_t precision,
                                    vfx_stream stream)
{
    if (!in) return fail(VFX_EINVAL, "in is NULL");
    return vf_array(lat, is_even, nullptr, in, precision, stream);
}

vfx_status vfx_macroscopic(const vfx_lattice* lat, int32_t is_even, double* drho, double* vx, double* vy, double* vz,
                           vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!drho || !vx || !vy || !vz) return fail(VFX_EINVAL, "macroscopic outputs are NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_macro<double>(L, lat->direct, is_even != 0, drho, vx, vy, vz, st); break;
    case VFX_FP32: e = vfx::launch_macro<float>(L, lat->direct, is_even != 0, drho, vx, vy, vz, st); break;
    default: e = vfx::launch_macro<_Float16>(L, lat->direct, is_even != 0, drho, vx, vy, vz, st); break;
    }
    return hip_status(e, "macroscopic launch");
}

vfx_status vfx_macroscopic_quantities(const vfx_lattice* lat, int32_t is_even, void* rho, void* vx, void* vy,
                                      void* vz, void* pressure, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!rho || !vx || !vy || !vz || !pressure) return fail(VFX_EINVAL, "macroscopic outputs are NULL");
    if (!lat->direct && !lat->type_of_grid_node && lat->n_fluid && !lat->fluid_nodes)
        return fail(VFX_EINVAL, "fluid_nodes is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    const bool even = is_even != 0;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_macro_vf<double>(L, lat->direct, even, rho, vx, vy, vz, pressure, st); break;
    case VFX_FP32: e = vfx::launch_macro_vf<float>(L, lat->direct, even, rho, vx, vy, vz, pressure, st); break;
    default: e = vfx::launch_macro_vf<_Float16>(L, lat->direct, even, rho, vx, vy, vz, pressure, st); break;
    }
    return hip_status(e, "macroscopic quantities launch");
}

vfx_status vfx_calculate_mean(const vfx_lattice* lat, int32_t is_even, void* vx, void* vy, void* vz, void* rho,
                              void* pressure, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!rho || !vx || !vy || !vz || !pressure) return fail(VFX_EINVAL, "mean sums are NULL");
    if (!lat->direct && !lat->type_of_grid_node && lat->n_fluid && !lat->fluid_nodes)
        return fail(VFX_EINVAL, "fluid_nodes is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    const bool even = is_even != 0;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_mean_vf<double>(L, lat->direct, even, vx, vy, vz, rho, pressure, st); break;
    case VFX_FP32: e = vfx::launch_mean_vf<float>(L, lat->direct, even, vx, vy, vz, rho, pressure, st); break;
    default: e = vfx::launch_mean_vf<_Float16>(L, lat->direct, even, vx, vy, vz, rho, pressure, st); break;
    }
    return hip_status(e, "mean launch");
}

vfx_status vfx_reset_mean(const vfx_lattice* lat, void* vx, void* vy, void* vz, void* rho, void* pressure,
                          vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!rho || !vx || !vy || !vz || !pressure) return fail(VFX_EINVAL, "mean sums are NULL");
    hipStream_t st = (hipStream_t)stream;
    const hipError_t e = lat->storage == VFX_FP64 ? vfx::launch_zero5<double>(L.N, vx, vy, vz, rho, pressure, st)
                                                  : vfx::launch_zero5<float>(L.N, vx, vy, vz, rho, pressure, st);
    return hip_status(e, "mean reset launch");
}

uint64_t vfx_reduce_workspace_bytes(const vfx_lattice* lat)
{
    if (!lat) return 0;
    const uint64_t n = lat->direct ? lat->n_nodes : lat->n_fluid;
    return (uint64_t)vfx::reduce_blocks(n ? n : 1) * 4 * sizeof(double);
}

vfx_status vfx_reduce_mass_momentum(const vfx_lattice* lat, int32_t is_even, double* workspace, double* out4,
                                    vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!workspace || !out4) return fail(VFX_EINVAL, "workspace / out4 is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_reduce_mass<double>(L, lat->direct, is_even != 0, workspace, out4, st); break;
    case VFX_FP32: e = vfx::launch_reduce_mass<float>(L, lat->direct, is_even != 0, workspace, out4, st); break;
    default: e = vfx::launch_reduce_mass<_Float16>(L, lat->direct, is_even != 0, workspace, out4, st); break;
    }
    return hip_status(e, "mass/momentum reduction launch");
}

}

// ------------------------------------------- reduction without allocation ----
// vfx_reduce_mass_momentum_noalloc keeps the per-block partial sums in this
// file-scope device array instead of a caller workspace: 65536 blocks x 4
// doubles = 2 MiB per GPU, held until the process ends.  In host code the
// array's name is not a device address, so each device's address is looked up
// once (hipGetSymbolAddress, with that device current) and only that address
// is handed to kernels.  Every call on one GPU shares the array: calls that
// overlap on one GPU (two streams, two host threads) race.  Nothing in it has
// to start at zero (each reduction block writes its own four partials before
// the final pass reads them).
namespace {
constexpr uint32_t kNoAllocMaxBlocks = 65536;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};
} // namespace

static __device__ __align__(256) double scratch[kNoAllocMaxBlocks * 4];

static vfx_status scratch_for_current_device(double** out)
{
    int dev = -1;
    if (hipError_t e = hipGetDevice(&dev)) return hip_status(e, "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(VFX_EUNSUPPORTED, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return hip_status(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(VFX_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return VFX_OK;
}

extern "C" {

uint64_t vfx_reduce_noalloc_max_items(void) { return (uint64_t)kNoAllocMaxBlocks * vfx::kRedChunk; }

vfx_status vfx_reduce_mass_momentum_noalloc(const vfx_lattice* lat, int32_t is_even, double* out4, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!out4) return fail(VFX_EINVAL, "out4 is NULL");
    const uint64_t n = vfx::node_items(L, lat->direct);
    if (vfx::reduce_blocks(n ? n : 1) > kNoAllocMaxBlocks)
        return fail(VFX_EINVAL, "%llu work items exceed the no-alloc reduction's maximum of %llu",
                    (unsigned long long)n, (unsigned long long)vfx_reduce_noalloc_max_items());
    double* ws = nullptr;
    if (vfx_status s = scratch_for_current_device(&ws)) return s;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_reduce_mass<double>(L, lat->direct, is_even != 0, ws, out4, st); break;
    case VFX_FP32: e = vfx::launch_reduce_mass<float>(L, lat->direct, is_even != 0, ws, out4, st); break;
    default: e = vfx::launch_reduce_mass<_Float16>(L, lat->direct, is_even != 0, ws, out4, st); break;
    }
    return hip_status(e, "mass/momentum reduction launch (no-alloc)");
}

static vfx_status energy(const vfx_lattice* lat, bool enstrophy, const double* drho, const double* vx, const double* vy,
                         const double* vz, double* workspace, double* out2, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!lat->direct) return fail(VFX_EUNSUPPORTED, "energy diagnostics need structured (direct) addressing");
    if (!drho || !vx || !vy || !vz || !workspace || !out2) return fail(VFX_EINVAL, "NULL field/workspace/output");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t nb = vfx::reduce_blocks(L.N);
    if (enstrophy)
        hipLaunchKernelGGL((vfx::k_reduce_energy<true>), dim3(nb), dim3(vfx::kRedBlock), 0, st, L, drho, vx, vy, vz, workspace);
    else
        hipLaunchKernelGGL((vfx::k_reduce_energy<false>), dim3(nb), dim3(vfx::kRedBlock), 0, st, L, drho, vx, vy, vz, workspace);
    hipLaunchKernelGGL((vfx::k_reduce_finish<2>), dim3(1), dim3(vfx::kRedBlock), 0, st, workspace, nb, out2);
    return hip_status(hipGetLastError(), enstrophy ? "enstrophy launch" : "kinetic energy launch");
}

vfx_status vfx_kinetic_energy(const vfx_lattice* lat, const double* drho, const double* vx, const double* vy,
                              const double* vz, double* workspace, double* out2, vfx_stream stream)
{
    return energy(lat, false, drho, vx, vy, vz, workspace, out2, stream);
}

vfx_status vfx_enstrophy(const vfx_lattice* lat, const double* drho, const double* vx, const double* vy, const double* vz,
                         double* workspace, double* out2, vfx_stream stream)
{
    return energy(lat, true, drho, vx, vy, vz, workspace, out2, stream);
}

vfx_status vfx_accumulate_stats(const vfx_lattice* lat, int32_t is_even, double* sums9, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!sums9) return fail(VFX_EINVAL, "sums9 is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_stats<double>(L, lat->direct, is_even != 0, sums9, st); break;
    case VFX_FP32: e = vfx::launch_stats<float>(L, lat->direct, is_even != 0, sums9, st); break;
    default: e = vfx::launch_stats<_Float16>(L, lat->direct, is_even != 0, sums9, st); break;
    }
    return hip_status(e, "statistics launch");
}

} // extern "C"

// --------------------------------------------------------------- refinement --
namespace {

vfx_status check_levels(const vfx_lattice* coarse, const vfx_lattice* fine, vfx::Lat& C, vfx::Lat& F)
{
    if (vfx_status s = make_lat(coarse, C)) return s;
    if (vfx_status s = make_lat(fine, F)) return s;
    if (coarse->storage != fine->storage) return fail(VFX_EINVAL, "coarse and fine storage differ");
    // Parameter::getUseTurbulentViscosity is one switch for every level
    if ((coarse->turbulence_model != VFX_TURB_NONE) != (fine->turbulence_model != VFX_TURB_NONE))
        return fail(VFX_EINVAL, "coarse and fine levels differ in using a turbulence model");
    return VFX_OK;
}

vfx_status make_cells(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells, vfx::ICellsDev& I)
{
    std::memset(&I, 0, sizeof(I));
    if (!cells) return fail(VFX_EINVAL, "cells is NULL");
    if (cells->n_cells == 0) return VFX_OK;
    if (!cells->fine_cells || !cells->coarse_cells) return fail(VFX_EINVAL, "cell index arrays are NULL");
    I.coarse = cells->coarse_cells;
    I.fine = cells->fine_cells;
    I.ox = cells->offset_x;
    I.oy = cells->offset_y;
    I.oz = cells->offset_z;
    I.n = cells->n_cells;
    I.direct_c = coarse->direct;
    I.direct_f = fine->direct;
    return VFX_OK;
}

vfx_status scale(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells, int32_t even_c,
                 int32_t even_f, vfx_stream stream, bool coarse_to_fine)
{
    vfx::Lat C, F;
    if (vfx_status s = check_levels(coarse, fine, C, F)) return s;
    vfx::ICellsDev I;
    if (vfx_status s = make_cells(coarse, fine, cells, I)) return s;
    if (I.n == 0) return VFX_OK;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (coarse->storage) {
    case VFX_FP64:
        e = vfx::launch_scale<double>(C, F, I, coarse_to_fine, even_c != 0, even_f != 0, coarse->omega, fine->omega, st);
        break;
    case VFX_FP32:
        e = vfx::launch_scale<float>(C, F, I, coarse_to_fine, even_c != 0, even_f != 0, coarse->omega, fine->omega, st);
        break;
    default:
        e = vfx::launch_scale<_Float16>(C, F, I, coarse_to_fine, even_c != 0, even_f != 0, coarse->omega, fine->omega,
                                        st);
        break;
    }
    return hip_status(e, coarse_to_fine ? "scale coarse to fine" : "scale fine to coarse");
}

} // namespace

extern "C" {

vfx_status vfx_scale_cf(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells,
                        int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream)
{
    return scale(coarse, fine, cells, coarse_is_even, fine_is_even, stream, true);
}

vfx_status vfx_scale_fc(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* cells,
                        int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream)
{
    return scale(coarse, fine, cells, coarse_is_even, fine_is_even, stream, false);
}

vfx_status vfx_scale_fc_cf(const vfx_lattice* coarse, const vfx_lattice* fine, const vfx_icells* fc,
                           const vfx_icells* cf, int32_t coarse_is_even, int32_t fine_is_even, vfx_stream stream)
{
    vfx::Lat C, F;
    if (vfx_status s = check_levels(coarse, fine, C, F)) return s;
    vfx::ICellsDev A, B;
    if (vfx_status s = make_cells(coarse, fine, fc, A)) return s;
    if (vfx_status s = make_cells(coarse, fine, cf, B)) return s;
    hipStream_t st = (hipStream_t)stream;
    const bool ec = coarse_is_even != 0, ef = fine_is_even != 0;
    hipError_t e;
    switch (coarse->storage) {
    case VFX_FP64: e = vfx::launch_scale_both<double>(C, F, A, B, ec, ef, coarse->omega, fine->omega, st); break;
    case VFX_FP32: e = vfx::launch_scale_both<float>(C, F, A, B, ec, ef, coarse->omega, fine->omega, st); break;
    default: e = vfx::launch_scale_both<_Float16>(C, F, A, B, ec, ef, coarse->omega, fine->omega, st); break;
    }
    return hip_status(e, "scale fine to coarse and coarse to fine");
}

} // extern "C"

// ------------------------------------------------------------------ batches --
struct vfx_batch_s {
    int32_t storage;
    int32_t variant;
    void* pool;
    uint64_t pool_elements;
    uint32_t n;
    std::vector<vfx_domain> domains;
    void* dev_domains = nullptr;               // DomainDev<T>[n]
    uint2* dev_tiles[3] = {nullptr, nullptr, nullptr}; // tile tables for blocks 64/128/256
    uint2* dev_slabs[3] = {nullptr, nullptr, nullptr}; // the same tiles in XCD-slab order (xcd_run 0)
    std::vector<uint32_t> first_tile[3];       // per domain, n+1 entries
    uint32_t ntiles[3] = {0, 0, 0};
    unsigned block = 256, grid = 0, xcd_run = 1;
    float ld[4], st[4];
};

struct vfx_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int block_index(unsigned b) { return b == 64 ? 0 : b == 128 ? 1 : b == 256 ? 2 : -1; }

template <typename T> void fill_domains(const vfx_batch_s& b, const double lim[3], std::vector<char>& out)
{
    out.resize(sizeof(vfx::DomainDev<T>) * b.n);
    auto* d = (vfx::DomainDev<T>*)out.data();
    const double zero[3] = {0, 0, 0};
    for (uint32_t i = 0; i < b.n; ++i) {
        const vfx_domain& s = b.domains[i];
        vfx::DomainDev<T> e{};
        e.f_off = s.f_offset;
        e.nx = s.dims[0];
        e.ny = s.dims[1];
        e.nz = s.dims[2];
        e.N = s.dims[0] * s.dims[1] * s.dims[2];
        e.nfx = s.dims[0] - 2;
        e.nfy = s.dims[1] - 2;
        e.nfluid = (s.dims[0] - 2) * (s.dims[1] - 2) * (s.dims[2] - 2);
        e.div_x = vfx::make_fastdiv(e.nfx);
        e.div_y = vfx::make_fastdiv(e.nfy);
        // wall term of VelocityBounceBack.cu:159-184 in the kernel's arithmetic
        // type, component order of proj() (first non-zero component with its
        // sign, the others added or subtracted in x, y, z order)
        const T u[3] = {(T)s.lid_velocity[0], (T)s.lid_velocity[1], (T)s.lid_velocity[2]};
        for (int dir = 0; dir < 27; ++dir) {
            const int c[3] = {vfx::cx(dir), vfx::cy(dir), vfx::cz(dir)};
            const int wc = vfx::weight_class(dir);
            const T coef = wc == 1 ? T(4. / 9.) : wc == 2 ? T(1. / 9.) : T(1. / 36.);
            T pr = T(0);
            bool first = true;
            for (int a = 0; a < 3; ++a) {
                if (c[a] == 0) continue;
                if (first) pr = c[a] > 0 ? u[a] : -u[a];
                else pr = c[a] > 0 ? pr + u[a] : pr - u[a];
                first = false;
            }
            e.lid_term[dir] = dir == 0 ? T(0) : coef * pr;
        }
        e.R = vfx::make_rates<T>(s.omega, lim, zero);
        d[i] = e;
    }
}

// whole: every tile of the ensemble (XCD-slab order when xcd_run is 0);
// otherwise the tiles [first, first + count) of the plain table
vfx::BatchDev batch_view(const vfx_batch_s* b, int bi, uint32_t first, uint32_t count, bool whole = false)
{
    vfx::BatchDev B;
    B.pool = b->pool;
    B.domains = b->dev_domains;
    const bool slabs = whole && b->xcd_run == 0;
    B.tiles = (slabs ? b->dev_slabs[bi] : b->dev_tiles[bi]) + first;
    B.ntiles = count;
    B.xcd_run = b->xcd_run == 0 ? 1 : b->xcd_run;
    for (int c = 0; c < 4; ++c) {
        B.ld[c] = b->ld[c];
        B.st[c] = b->st[c];
    }
    return B;
}

hipError_t launch_batch(const vfx_batch_s* b, const vfx::BatchDev& B, bool even, unsigned block, unsigned grid,
                        hipStream_t st)
{
    if (B.ntiles == 0) return hipSuccess;
    const unsigned g = grid == 0 || grid > B.ntiles ? B.ntiles : grid;
    switch (b->storage) {
    case VFX_FP64: return vfx::launch_batched<double>(B, even, b->variant, block, g, st);
    case VFX_FP32: return vfx::launch_batched<float>(B, even, b->variant, block, g, st);
    default: return vfx::launch_batched<_Float16>(B, even, b->variant, block, g, st);
    }
}

} // namespace

extern "C" {

vfx_status vfx_batch_create(void* pool, uint64_t pool_elements, int32_t storage, float fp16_scale,
                            const vfx_domain* domains, uint32_t n_domains, const double* quadric_limiter,
                            int32_t kernel_variant, vfx_batch* out)
{
    if (!out) return fail(VFX_EINVAL, "out is NULL");
    *out = nullptr;
    if (!pool || !domains || n_domains == 0) return fail(VFX_EINVAL, "pool/domains empty");
    if (storage < VFX_FP64 || storage > VFX_FP16_SCALED) return fail(VFX_EINVAL, "unknown storage");
    if (storage == VFX_FP16_SCALED && !(fp16_scale > 0.f)) return fail(VFX_EINVAL, "fp16_scale must be > 0");
    if (kernel_variant != 0 && kernel_variant != 1) return fail(VFX_EINVAL, "unknown kernel_variant");
    const uint64_t esz = storage == VFX_FP64 ? 8 : storage == VFX_FP32 ? 4 : 2;
    for (uint32_t i = 0; i < n_domains; ++i) {
        const vfx_domain& d = domains[i];
        if (d.dims[0] < 3 || d.dims[1] < 3 || d.dims[2] < 3) return fail(VFX_EINVAL, "domain %u: dims must be >= 3", i);
        const uint64_t N = (uint64_t)d.dims[0] * d.dims[1] * d.dims[2];
        // one buffer descriptor spans a domain's 27 planes
        if (27 * N * esz >= (1ull << 32)) return fail(VFX_EUNSUPPORTED, "domain %u too large (27 planes >= 4 GiB)", i);
        if (!(d.omega > 0.0 && d.omega < 2.0)) return fail(VFX_EINVAL, "domain %u: omega %g outside (0,2)", i, d.omega);
        if (d.f_offset > pool_elements || 27 * N > pool_elements - d.f_offset)
            return fail(VFX_EINVAL, "domain %u: planes end beyond the pool", i);
    }
    {
        std::vector<std::pair<uint64_t, uint64_t>> r(n_domains);
        for (uint32_t i = 0; i < n_domains; ++i)
            r[i] = {domains[i].f_offset,
                    domains[i].f_offset + 27ull * domains[i].dims[0] * domains[i].dims[1] * domains[i].dims[2]};
        std::sort(r.begin(), r.end());
        for (uint32_t i = 1; i < n_domains; ++i)
            if (r[i].first < r[i - 1].second) return fail(VFX_EINVAL, "domains overlap in the pool");
    }
    const double defl[3] = {0.01, 0.01, 0.01};
    const double* lim = quadric_limiter ? quadric_limiter : defl;
    auto* b = new vfx_batch_s();
    b->storage = storage;
    b->variant = kernel_variant;
    b->pool = pool;
    b->pool_elements = pool_elements;
    b->n = n_domains;
    b->domains.assign(domains, domains + n_domains);
    for (int c = 0; c < 4; ++c) {
        if (storage == VFX_FP16_SCALED) {
            b->ld[c] = (float)(kW[c] / (double)fp16_scale);
            b->st[c] = (float)((double)fp16_scale / kW[c]);
        } else {
            b->ld[c] = b->st[c] = 1.f;
        }
    }
    std::vector<char> host;
    if (storage == VFX_FP64) fill_domains<double>(*b, lim, host);
    else fill_domains<float>(*b, lim, host);
    hipError_t e = hipMalloc(&b->dev_domains, host.size());
    if (e == hipSuccess) e = hipMemcpy(b->dev_domains, host.data(), host.size(), hipMemcpyHostToDevice);
    const unsigned blocks[3] = {64, 128, 256};
    for (int bi = 0; bi < 3 && e == hipSuccess; ++bi) {
        std::vector<uint2> tiles;
        b->first_tile[bi].assign(n_domains + 1, 0);
        for (uint32_t i = 0; i < n_domains; ++i) {
            b->first_tile[bi][i] = (uint32_t)tiles.size();
            const vfx_domain& d = domains[i];
            const uint32_t nf = (d.dims[0] - 2) * (d.dims[1] - 2) * (d.dims[2] - 2);
            for (uint32_t s = 0; s < nf; s += blocks[bi]) tiles.push_back(make_uint2(i, s));
        }
        b->first_tile[bi][n_domains] = (uint32_t)tiles.size();
        b->ntiles[bi] = (uint32_t)tiles.size();
        e = hipMalloc((void**)&b->dev_tiles[bi], tiles.size() * sizeof(uint2));
        if (e == hipSuccess) e = hipMemcpy(b->dev_tiles[bi], tiles.data(), tiles.size() * sizeof(uint2), hipMemcpyHostToDevice);
        // XCD-slab order: workgroup w runs on XCD w mod 8, so position 8j + x
        // takes the j-th tile of XCD x's queue, and queue x holds the x-th of
        // eight contiguous slabs of every domain -- a domain's tiles meet their
        // z and y neighbours (other tiles of the same slab) in one L2.  The
        // remainder tiles of uneven splits rotate over the XCDs.
        std::vector<std::vector<uint2>> q(vfx::kXcds);
        uint32_t rot = 0;
        for (uint32_t i = 0; i < n_domains; ++i) {
            const uint32_t t0 = b->first_tile[bi][i], T = b->first_tile[bi][i + 1] - t0;
            const uint32_t base = T / vfx::kXcds, rem = T % vfx::kXcds;
            uint32_t pos = t0;
            for (uint32_t x = 0; x < vfx::kXcds; ++x) {
                const uint32_t size = base + ((x + vfx::kXcds - rot) % vfx::kXcds < rem ? 1u : 0u);
                for (uint32_t j = 0; j < size; ++j) q[x].push_back(tiles[pos + j]);
                pos += size;
            }
            rot = (rot + rem) % vfx::kXcds;
        }
        std::vector<uint2> slab;
        slab.reserve(tiles.size());
        size_t longest = 0;
        for (const auto& v : q) longest = std::max(longest, v.size());
        for (size_t j = 0; j < longest; ++j)
            for (uint32_t x = 0; x < vfx::kXcds; ++x)
                if (j < q[x].size()) slab.push_back(q[x][j]);
        if (e == hipSuccess) e = hipMalloc((void**)&b->dev_slabs[bi], slab.size() * sizeof(uint2));
        if (e == hipSuccess) e = hipMemcpy(b->dev_slabs[bi], slab.data(), slab.size() * sizeof(uint2), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        vfx_batch_destroy(b);
        return hip_status(e, "batch tables");
    }
    *out = b;
    return VFX_OK;
}

vfx_status vfx_batch_destroy(vfx_batch b)
{
    if (!b) return VFX_OK;
    if (b->dev_domains) (void)hipFree(b->dev_domains);
    for (auto* p : b->dev_slabs)
        if (p) (void)hipFree(p);
    for (auto* p : b->dev_tiles)
        if (p) (void)hipFree(p);
    delete b;
    return VFX_OK;
}

vfx_status vfx_batch_set_launch(vfx_batch b, uint32_t block, uint32_t grid)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    if (block_index(block) < 0) return fail(VFX_EINVAL, "block must be 64, 128 or 256");
    b->block = block;
    b->grid = grid;
    return VFX_OK;
}

vfx_status vfx_batch_set_xcd_run(vfx_batch b, uint32_t run)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    if (run > 4096) return fail(VFX_EINVAL, "xcd run %u outside 0..4096", run);
    b->xcd_run = run;
    return VFX_OK;
}

vfx_status vfx_batch_get_xcd_run(vfx_batch b, uint32_t* run)
{
    if (!b || !run) return fail(VFX_EINVAL, "NULL argument");
    *run = b->xcd_run;
    return VFX_OK;
}

vfx_status vfx_batch_get_launch(vfx_batch b, uint32_t* block, uint32_t* grid)
{
    if (!b || !block || !grid) return fail(VFX_EINVAL, "NULL argument");
    *block = b->block;
    *grid = b->grid;
    return VFX_OK;
}

vfx_status vfx_batch_init_rest(vfx_batch b, vfx_stream stream)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (b->storage) {
    case VFX_FP64: e = vfx::launch_batch_zero<double>(b->pool, b->pool_elements, st); break;
    case VFX_FP32: e = vfx::launch_batch_zero<float>(b->pool, b->pool_elements, st); break;
    default: e = vfx::launch_batch_zero<_Float16>(b->pool, b->pool_elements, st); break;
    }
    return hip_status(e, "batch init");
}

vfx_status vfx_batch_step(vfx_batch b, int32_t is_even, vfx_stream stream)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    const int bi = block_index(b->block);
    return hip_status(launch_batch(b, batch_view(b, bi, 0, b->ntiles[bi], true), is_even != 0, b->block, b->grid,
                                   (hipStream_t)stream),
                      "batch step");
}

vfx_status vfx_batch_step_domain(vfx_batch b, uint32_t domain, int32_t is_even, vfx_stream stream)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    if (domain >= b->n) return fail(VFX_EINVAL, "domain %u out of range", domain);
    const int bi = block_index(b->block);
    const uint32_t first = b->first_tile[bi][domain], count = b->first_tile[bi][domain + 1] - first;
    return hip_status(launch_batch(b, batch_view(b, bi, first, count), is_even != 0, b->block, 0, (hipStream_t)stream),
                      "batch domain step");
}

vfx_status vfx_batch_lattice(vfx_batch b, uint32_t domain, vfx_lattice* out)
{
    if (!b || !out) return fail(VFX_EINVAL, "NULL argument");
    if (domain >= b->n) return fail(VFX_EINVAL, "domain out of range");
    const vfx_domain& d = b->domains[domain];
    const uint64_t esz = b->storage == VFX_FP64 ? 8 : b->storage == VFX_FP32 ? 4 : 2;
    std::memset(out, 0, sizeof(*out));
    out->f = (char*)b->pool + d.f_offset * esz;
    out->storage = b->storage;
    out->fp16_scale = b->storage == VFX_FP16_SCALED ? (float)(kW[0] / (double)b->ld[0]) : 1.f;
    out->n_nodes = (uint64_t)d.dims[0] * d.dims[1] * d.dims[2];
    out->direct = 1;
    for (int c = 0; c < 3; ++c) out->dims[c] = d.dims[c];
    out->omega = d.omega;
    out->quadric_limiter[0] = out->quadric_limiter[1] = out->quadric_limiter[2] = 0.01;
    return VFX_OK;
}

// Times `steps` batched steps per candidate launch shape (blocks 64/128/256,
// one workgroup per tile or a persistent grid of 2/4/8 workgroups per CU-slot)
// with HIP events and keeps the fastest.  The state advances during tuning,
// starting from the parity is_even; every trial runs an even number of steps,
// so the parity is the same afterwards.
vfx_status vfx_batch_autotune(vfx_batch b, int32_t is_even, int32_t steps, vfx_stream stream, uint32_t* block_out,
                              uint32_t* grid_out, float* ms_out)
{
    if (!b) return fail(VFX_EINVAL, "batch is NULL");
    if (steps < 2) steps = 2;
    steps += steps & 1;
    const bool even0 = is_even != 0;
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(VFX_EHIP, "event create");
    }
    const unsigned keep_block = b->block, keep_grid = b->grid, keep_run = b->xcd_run;
    float best = 1e30f;
    unsigned bb = keep_block, bg = keep_grid, br = keep_run;
    vfx_status r = VFX_OK;
    const unsigned blocks[3] = {64, 128, 256};
    const unsigned runs[4] = {1, 4, 16, 0};
    for (unsigned block : blocks) {
        const unsigned grids[4] = {0, (unsigned)cus * 8, (unsigned)cus * 16, (unsigned)cus * 32};
        for (int cfg = 0; cfg < 16; ++cfg) {
            const unsigned grid = grids[cfg / 4], run = runs[cfg % 4];
            b->block = block;
            b->grid = grid;
            b->xcd_run = run;
            for (int s = 0; s < 2 && r == VFX_OK; ++s) r = vfx_batch_step(b, (s % 2 == 0) == even0, stream); // warm-up
            if (r == VFX_OK) r = hip_status(hipEventRecord(e0, st), "event record");
            for (int s = 0; s < steps && r == VFX_OK; ++s) r = vfx_batch_step(b, (s % 2 == 0) == even0, stream);
            if (r == VFX_OK) r = hip_status(hipEventRecord(e1, st), "event record");
            if (r == VFX_OK) r = hip_status(hipEventSynchronize(e1), "event synchronize");
            float ms = 0;
            if (r == VFX_OK) r = hip_status(hipEventElapsedTime(&ms, e0, e1), "event elapsed");
            if (r != VFX_OK) break;
            ms /= steps;
            if (ms < best) {
                best = ms;
                bb = block;
                bg = grid;
                br = run;
            }
        }
        if (r != VFX_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != VFX_OK) {
        b->block = keep_block;
        b->grid = keep_grid;
        b->xcd_run = keep_run;
        return r;
    }
    b->block = bb;
    b->grid = bg;
    b->xcd_run = br;
    if (block_out) *block_out = bb;
    if (grid_out) *grid_out = bg;
    if (ms_out) *ms_out = best;
    return VFX_OK;
}

// Captures n_steps batched steps (parity alternating from is_even_start) --
// or, with per_domain = 1, n_steps x n_domains single-domain launches -- into
// a HIP graph.  Capture runs on a private stream; the graph may be launched on
// any stream.
vfx_status vfx_batch_capture(vfx_batch b, int32_t n_steps, int32_t is_even_start, int32_t per_domain, vfx_graph* out)
{
    if (!b || !out) return fail(VFX_EINVAL, "NULL argument");
    if (n_steps < 1) return fail(VFX_EINVAL, "n_steps must be >= 1");
    *out = nullptr;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return fail(VFX_EHIP, "stream create");
    auto* g = new vfx_graph_s();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    bool even = is_even_start != 0;
    vfx_status s = VFX_OK;
    for (int i = 0; i < n_steps && e == hipSuccess && s == VFX_OK; ++i) {
        if (per_domain) {
            for (uint32_t d = 0; d < b->n && s == VFX_OK; ++d) s = vfx_batch_step_domain(b, d, even, cs);
        } else {
            s = vfx_batch_step(b, even, cs);
        }
        even = !even;
    }
    hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess && s == VFX_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (s != VFX_OK || e != hipSuccess) {
        vfx_graph_destroy(g);
        return s != VFX_OK ? s : hip_status(e, "graph capture");
    }
    *out = g;
    return VFX_OK;
}

vfx_status vfx_stream_create(vfx_stream* out)
{
    if (!out) return fail(VFX_EINVAL, "out is NULL");
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *out = e == hipSuccess ? (vfx_stream)s : nullptr;
    return hip_status(e, "stream create");
}

vfx_status vfx_stream_destroy(vfx_stream stream)
{
    if (!stream) return fail(VFX_EINVAL, "stream is NULL");
    return hip_status(hipStreamDestroy((hipStream_t)stream), "stream destroy");
}

vfx_status vfx_capture_begin(vfx_stream stream)
{
    if (!stream) return fail(VFX_EINVAL, "capture needs a created stream, not the NULL stream");
    return hip_status(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "begin capture");
}

vfx_status vfx_capture_end(vfx_stream stream, vfx_graph* out)
{
    if (!stream || !out) return fail(VFX_EINVAL, "NULL argument");
    *out = nullptr;
    auto* g = new vfx_graph_s();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        vfx_graph_destroy(g);
        return hip_status(e, "end capture");
    }
    *out = g;
    return VFX_OK;
}

vfx_status vfx_graph_launch(vfx_graph g, vfx_stream stream)
{
    if (!g || !g->exec) return fail(VFX_EINVAL, "graph is NULL");
    return hip_status(hipGraphLaunch(g->exec, (hipStream_t)stream), "graph launch");
}

vfx_status vfx_graph_destroy(vfx_graph g)
{
    if (!g) return VFX_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return VFX_OK;
}

} // extern "C"
