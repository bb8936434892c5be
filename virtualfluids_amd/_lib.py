"""ctypes binding of libvfx.so (the C ABI declared in include/vfx.h).

The binding fails loudly when the HIP library is missing: there is no CPU
fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvfx.so")

VFX_OK, VFX_EINVAL, VFX_EHIP, VFX_EUNSUPPORTED = 0, 1, 2, 3
VFX_FP64, VFX_FP32, VFX_FP16_SCALED = 0, 1, 2
VFX_K17, VFX_BGK = 0, 1
VFX_TEMPLATE_DEFAULT, VFX_TEMPLATE_WRITE_MACRO, VFX_TEMPLATE_APPLY_BODY_FORCE, VFX_TEMPLATE_ALL_FEATURES = 0, 1, 2, 3
VFX_TURB_NONE, VFX_TURB_SMAGORINSKY, VFX_TURB_AMD, VFX_TURB_QR = 0, 1, 2, 3
VFX_BC_VELOCITY_BOUNCE_BACK, VFX_BC_NOSLIP_INTERPOLATED, VFX_BC_VELOCITY_INTERPOLATED = 0, 1, 2
VFX_BC_SLIP, VFX_BC_PRESSURE_NEQ = 3, 4
VFX_BC_VELOCITY_PRESSURE_INTERPOLATED, VFX_BC_OUTFLOW_NON_REFLECTING = 5, 6
VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION = 7
VFX_BC_SLIP_TURBULENT_VISCOSITY = 8
VFX_BC_STRESS_INTERPOLATED, VFX_BC_STRESS_BOUNCE_BACK, VFX_BC_STRESS_BOUNCE_BACK_PRESSURE = 9, 10, 11
VFX_BC_PRECURSOR_VELOCITY, VFX_BC_PRECURSOR_DISTRIBUTIONS = 12, 13

_u32p = C.POINTER(C.c_uint32)
_dp = C.POINTER(C.c_double)


class VfxLattice(C.Structure):
    _fields_ = [
        ("f", C.c_void_p), ("storage", C.c_int32), ("fp16_scale", C.c_float), ("n_nodes", C.c_uint64),
        ("direct", C.c_int32), ("dims", C.c_uint32 * 3), ("solid_mask", C.c_void_p),
        ("neighbor_x", C.c_void_p), ("neighbor_y", C.c_void_p), ("neighbor_z", C.c_void_p),
        ("fluid_nodes", C.c_void_p), ("n_fluid", C.c_uint32),
        ("omega", C.c_double), ("quadric_limiter", C.c_double * 3), ("force", C.c_double * 3),
        ("rho", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
        ("kernel_variant", C.c_int32), ("fp16_reference_velocity", C.c_double * 3),
        ("type_of_grid_node", C.c_void_p),
        ("turbulence_model", C.c_int32), ("sgs_constant", C.c_double), ("turbulent_viscosity", C.c_void_p),
        ("neighbor_inverse", C.c_void_p),
        ("body_force", C.c_void_p * 3), ("forces", C.c_double * 3), ("force_factor", C.c_double),
        ("tile_nodes", C.c_uint32), ("xcd_run", C.c_uint32), ("regular_groups", C.c_void_p),
    ]


class VfxBC(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n", C.c_uint32), ("k", C.c_void_p), ("q27", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p), ("cpu_float_velocity", C.c_int32),
        ("rho", C.c_void_p), ("k_neighbor", C.c_void_p), ("normal", C.c_int32),
        ("pressure_correction_factor", C.c_double),
        ("wm_normal", C.c_void_p * 3), ("wm_exchange_velocity", C.c_void_p * 3), ("wm_wall_velocity", C.c_void_p * 3),
        ("wm_sampling_offset", C.c_void_p), ("wm_z0", C.c_void_p), ("wm_u_star", C.c_void_p),
        ("wm_force", C.c_void_p * 3),
        ("n_precursor_nodes", C.c_uint32), ("plane_neighbor", C.c_void_p * 4), ("plane_weight", C.c_void_p * 4),
        ("precursor_last", C.c_void_p), ("precursor_current", C.c_void_p), ("time_ratio", C.c_double),
        ("velocity_ratio", C.c_double), ("precursor_velocity", C.c_double * 3),
    ]


class VfxFusedBC(C.Structure):
    _fields_ = [("bc", C.POINTER(VfxBC)), ("node_mask", C.c_void_p), ("word_base", C.c_void_p)]


class VfxLevelStep(C.Structure):
    _fields_ = [("lattice", C.POINTER(VfxLattice)), ("is_even_timestep", C.c_int32)]


class VfxDomain(C.Structure):
    _fields_ = [("f_offset", C.c_uint64), ("dims", C.c_uint32 * 3), ("omega", C.c_double),
                ("lid_velocity", C.c_double * 3)]


class VfxICells(C.Structure):
    _fields_ = [("fine_cells", C.c_void_p), ("coarse_cells", C.c_void_p), ("n_cells", C.c_uint32),
                ("offset_x", C.c_void_p), ("offset_y", C.c_void_p), ("offset_z", C.c_void_p)]


class VfxError(RuntimeError):
    pass


_lib = None

_SIGNATURES = {
    "vfx_collide_stream": [C.POINTER(VfxLattice), C.c_int32, C.c_int32, C.c_int32, C.c_void_p],
    "vfx_k17_step": [C.POINTER(VfxLattice), C.c_int32, C.c_int32, C.c_void_p],
    "vfx_bgk_step": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p],
    "vfx_sparse_regular_groups": [C.POINTER(VfxLattice), C.c_void_p, C.c_uint64, C.c_void_p],
    "vfx_k17_autotune": [C.POINTER(VfxLattice), C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(C.c_uint32),
                         C.POINTER(C.c_uint32), C.POINTER(C.c_float)],
    "vfx_k17_step_fused": [C.POINTER(VfxLattice), C.c_int32, C.POINTER(VfxFusedBC), C.c_int32,
                           C.POINTER(VfxFusedBC), C.c_void_p, C.c_int32, C.c_void_p],
    "vfx_bc_apply": [C.POINTER(VfxLattice), C.POINTER(VfxBC), C.c_int32, C.c_void_p],
    "vfx_init_equilibrium": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p],
    "vfx_read_populations": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p],
    "vfx_write_populations": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p],
    "vfx_export_distributions": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_int32, C.c_void_p],
    "vfx_turbulent_viscosity_amd": [C.POINTER(VfxLattice), C.c_void_p],
    "vfx_macroscopic_quantities": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p],
    "vfx_import_distributions": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_int32, C.c_void_p],
    "vfx_calculate_mean": [C.POINTER(VfxLattice), C.c_int32] + [C.c_void_p] * 6,
    "vfx_reset_mean": [C.POINTER(VfxLattice)] + [C.c_void_p] * 6,
    "vfx_macroscopic": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                        C.c_void_p],
    "vfx_reduce_mass_momentum": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p],
    "vfx_reduce_mass_momentum_noalloc": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p],
    "vfx_kinetic_energy": [C.POINTER(VfxLattice)] + [C.c_void_p] * 7,
    "vfx_enstrophy": [C.POINTER(VfxLattice)] + [C.c_void_p] * 7,
    "vfx_accumulate_stats": [C.POINTER(VfxLattice), C.c_int32, C.c_void_p, C.c_void_p],
    "vfx_scale_cf": [C.POINTER(VfxLattice), C.POINTER(VfxLattice), C.POINTER(VfxICells), C.c_int32, C.c_int32,
                     C.c_void_p],
    "vfx_scale_fc": [C.POINTER(VfxLattice), C.POINTER(VfxLattice), C.POINTER(VfxICells), C.c_int32, C.c_int32,
                     C.c_void_p],
    "vfx_k17_step_levels": [C.POINTER(VfxLevelStep), C.c_int32, C.c_void_p],
    "vfx_scale_fc_cf": [C.POINTER(VfxLattice), C.POINTER(VfxLattice), C.POINTER(VfxICells), C.POINTER(VfxICells),
                        C.c_int32, C.c_int32, C.c_void_p],
    "vfx_bc_set_create": [C.POINTER(C.POINTER(VfxLattice)), C.POINTER(C.POINTER(VfxBC)), C.c_int32,
                          C.POINTER(C.c_void_p)],
    "vfx_bc_set_apply": [C.c_void_p, C.c_uint32, C.c_void_p],
    "vfx_bc_set_destroy": [C.c_void_p],
    "vfx_batch_create": [C.c_void_p, C.c_uint64, C.c_int32, C.c_float, C.POINTER(VfxDomain), C.c_uint32, C.c_void_p, C.c_int32,
                         C.POINTER(C.c_void_p)],
    "vfx_batch_destroy": [C.c_void_p],
    "vfx_batch_init_rest": [C.c_void_p, C.c_void_p],
    "vfx_batch_set_launch": [C.c_void_p, C.c_uint32, C.c_uint32],
    "vfx_batch_get_launch": [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "vfx_batch_set_xcd_run": [C.c_void_p, C.c_uint32],
    "vfx_batch_get_xcd_run": [C.c_void_p, C.POINTER(C.c_uint32)],
    "vfx_batch_autotune": [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                           C.POINTER(C.c_float)],
    "vfx_batch_step": [C.c_void_p, C.c_int32, C.c_void_p],
    "vfx_batch_step_domain": [C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p],
    "vfx_batch_lattice": [C.c_void_p, C.c_uint32, C.POINTER(VfxLattice)],
    "vfx_batch_capture": [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)],
    "vfx_graph_launch": [C.c_void_p, C.c_void_p],
    "vfx_stream_create": [C.POINTER(C.c_void_p)],
    "vfx_stream_destroy": [C.c_void_p],
    "vfx_capture_begin": [C.c_void_p],
    "vfx_capture_end": [C.c_void_p, C.POINTER(C.c_void_p)],
    "vfx_graph_destroy": [C.c_void_p],
}


def lib():
    """Load libvfx.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VfxError(f"{LIB_PATH} is missing: build it with `python -m virtualfluids_amd.build` "
                           "(the MI355X stepper has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, argt in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = C.c_int32
        L.vfx_reduce_workspace_bytes.argtypes = [C.POINTER(VfxLattice)]
        L.vfx_reduce_workspace_bytes.restype = C.c_uint64
        L.vfx_reduce_noalloc_max_items.argtypes = []
        L.vfx_reduce_noalloc_max_items.restype = C.c_uint64
        L.vfx_sparse_regular_groups_bytes.argtypes = [C.POINTER(VfxLattice)]
        L.vfx_sparse_regular_groups_bytes.restype = C.c_uint64
        L.vfx_last_error.restype = C.c_char_p
        L.vfx_version.restype = C.c_int32
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != VFX_OK:
        msg = lib().vfx_last_error().decode(errors="replace")
        raise VfxError(f"{what or 'vfx call'} failed (status {status}): {msg}")
