// api.hip -- the extern "C" surface declared in include/vfx.h.
//
// Validates every argument on the host before anything is launched (a kernel
// is never started with shapes it does not expect), translates vfx_lattice into
// the device view and dispatches on the storage type.  Errors are returned as
// vfx_status; the message is kept per host thread for vfx_last_error().
#include "kernels.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

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
    if (in->direct) {
        const uint64_t nx = in->dims[0], ny = in->dims[1], nz = in->dims[2];
        if (nx == 0 || ny == 0 || nz == 0) return fail(VFX_EINVAL, "direct addressing needs dims > 0");
        if (nx * ny * nz != in->n_nodes)
            return fail(VFX_EINVAL, "dims %llu*%llu*%llu != n_nodes %llu", (unsigned long long)nx,
                        (unsigned long long)ny, (unsigned long long)nz, (unsigned long long)in->n_nodes);
        if (ny * nz >= (1ull << 31) / ((nx + 255) / 256)) return fail(VFX_EUNSUPPORTED, "grid too large");
        L.nx = (uint32_t)nx;
        L.ny = (uint32_t)ny;
        L.nz = (uint32_t)nz;
        L.chunks = (uint32_t)((nx + 255) / 256);
        L.solid = in->solid_mask;
    } else {
        if (!in->neighbor_x || !in->neighbor_y || !in->neighbor_z)
            return fail(VFX_EINVAL, "sparse addressing needs neighbor_x/y/z");
        if (in->n_fluid && !in->fluid_nodes) return fail(VFX_EINVAL, "fluid_nodes is NULL");
        L.nbx = in->neighbor_x;
        L.nby = in->neighbor_y;
        L.nbz = in->neighbor_z;
        L.fluid = in->fluid_nodes;
        L.nfluid = in->n_fluid;
    }
    if (in->storage == VFX_FP16_SCALED) {
        if (!(in->fp16_scale > 0.f)) return fail(VFX_EINVAL, "fp16_scale must be > 0");
        for (int c = 0; c < 4; ++c) {
            L.ld[c] = (float)(kW[c] / (double)in->fp16_scale);
            L.st[c] = (float)((double)in->fp16_scale / kW[c]);
        }
    } else {
        for (int c = 0; c < 4; ++c) L.ld[c] = L.st[c] = 1.f;
    }
    L.rho = in->rho;
    L.vx = in->vx;
    L.vy = in->vy;
    L.vz = in->vz;
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
    if (tmpl != VFX_TEMPLATE_DEFAULT && tmpl != VFX_TEMPLATE_WRITE_MACRO)
        return fail(VFX_EUNSUPPORTED, "collision template %d not supported", tmpl);
    if (tmpl == VFX_TEMPLATE_WRITE_MACRO) {
        if (!lat->rho || !lat->vx || !lat->vy || !lat->vz)
            return fail(VFX_EINVAL, "WriteMacroVars needs rho/vx/vy/vz");
        L.write_macro = 1;
    }
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64:
        e = vfx::launch_step<double>(L, lat->direct, collision, is_even != 0, lat->kernel_variant, lat->omega, lat->quadric_limiter, lat->force, st);
        break;
    case VFX_FP32:
        e = vfx::launch_step<float>(L, lat->direct, collision, is_even != 0, lat->kernel_variant, lat->omega, lat->quadric_limiter, lat->force, st);
        break;
    default:
        e = vfx::launch_step<_Float16>(L, lat->direct, collision, is_even != 0, lat->kernel_variant, lat->omega, lat->quadric_limiter, lat->force, st);
        break;
    }
    return hip_status(e, "collide_stream launch");
}

vfx_status vfx_k17_step(const vfx_lattice* lat, int32_t is_even, int32_t tmpl, vfx_stream stream)
{
    return vfx_collide_stream(lat, VFX_K17, is_even, tmpl, stream);
}

vfx_status vfx_bgk_step(const vfx_lattice* lat, int32_t is_even, vfx_stream stream)
{
    return vfx_collide_stream(lat, VFX_BGK, is_even, VFX_TEMPLATE_DEFAULT, stream);
}

vfx_status vfx_bc_apply(const vfx_lattice* lat, const vfx_bc* bc, int32_t is_even, vfx_stream stream)
{
    vfx::Lat L;
    if (vfx_status s = make_lat(lat, L)) return s;
    if (!bc) return fail(VFX_EINVAL, "bc is NULL");
    if (bc->kind < VFX_BC_VELOCITY_BOUNCE_BACK || bc->kind > VFX_BC_PRESSURE_NEQ)
        return fail(VFX_EUNSUPPORTED, "bc kind %d not supported", bc->kind);
    if (bc->kind == VFX_BC_PRESSURE_NEQ) {
        if (bc->n && (!bc->k || !bc->k_neighbor || !bc->rho)) return fail(VFX_EINVAL, "pressure bc needs k, k_neighbor, rho");
        if (bc->normal < 1 || bc->normal > 26) return fail(VFX_EINVAL, "pressure bc normal %d invalid", bc->normal);
    } else if (bc->n && (!bc->k || !bc->q27)) {
        return fail(VFX_EINVAL, "bc.k / bc.q27 is NULL");
    }
    vfx::BCDev B{bc->n, bc->k, bc->q27, bc->vx, bc->vy, bc->vz, bc->cpu_float_velocity, lat->direct,
                 lat->omega, bc->rho, bc->k_neighbor, bc->normal};
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    switch (lat->storage) {
    case VFX_FP64: e = vfx::launch_bc<double>(L, B, bc->kind, is_even != 0, st); break;
    case VFX_FP32: e = vfx::launch_bc<float>(L, B, bc->kind, is_even != 0, st); break;
    default: e = vfx::launch_bc<_Float16>(L, B, bc->kind, is_even != 0, st); break;
    }
    return hip_status(e, "bc launch");
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
