"""Host-side lattice objects over the libvfx C ABI.

PyTorch only provides device memory and the stream here; every operation on
the populations is a libvfx kernel.  Mirrors the data VirtualFluids' GPU path
keeps per level (``LBMSimulationParameter``, src/gpu/core/Parameter/Parameter.h:66-255):
the distribution array, the neighbour arrays or the structured box, the
fluid-node list, omega and the quadric limiters.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as V

STORAGE = {"fp64": V.VFX_FP64, "fp32": V.VFX_FP32, "fp16": V.VFX_FP16_SCALED}
STORAGE_DTYPE = {"fp64": torch.float64, "fp32": torch.float32, "fp16": torch.float16}
BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}
COLLISION = {"k17": V.VFX_K17, "bgk": V.VFX_BGK}
# vf::lbm::TurbulenceModel of the K17 collision
TURBULENCE = {"none": V.VFX_TURB_NONE, "smagorinsky": V.VFX_TURB_SMAGORINSKY, "amd": V.VFX_TURB_AMD,
              "qr": V.VFX_TURB_QR}

# D3Q27 velocity set in VirtualFluids order (src/lbm/constants/D3Q27.h:42-72)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, 1, -1, -1, -1, -1])
INV = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 26, 25, 24, 23, 22, 21, 20, 19])
W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _stream(stream) -> int | None:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream


def pack_mask(solid: np.ndarray) -> np.ndarray:
    """Bit mask (uint32 words, bit k = node k) from a boolean node array."""
    flat = np.ascontiguousarray(solid, dtype=bool).ravel()
    nwords = (flat.size + 31) // 32
    padded = np.zeros(nwords * 32, dtype=bool)
    padded[: flat.size] = flat
    bits = padded.reshape(nwords, 32).astype(np.uint64) << np.arange(32, dtype=np.uint64)
    return bits.sum(axis=1).astype(np.uint32)


class Lattice:
    """One block of D3Q27 populations on the GPU.

    Structured (``dims=(nx, ny, nz)``, x fastest, periodic index wrap, walls as
    solid nodes) or VirtualFluids sparse (``neighbors=(nbx, nby, nbz)`` +
    ``fluid_nodes``)."""

    def __init__(self, dims=None, *, storage="fp64", omega=1.0, limiter=(0.01, 0.01, 0.01),
                 force=(0.0, 0.0, 0.0), solid=None, fp16_scale=None, neighbors=None, fluid_nodes=None,
                 n_nodes=None, device="cuda", variant="default", reference_velocity=(0.0, 0.0, 0.0),
                 node_types=None, turbulence="none", sgs_constant=0.0, neighbor_inverse=None,
                 regular_groups=True):
        if storage not in STORAGE:
            raise ValueError(f"storage must be one of {sorted(STORAGE)}")
        self.storage = storage
        if variant not in ("default", "plain"):
            raise ValueError("variant must be 'default' or 'plain'")
        self.variant = variant
        # launch shape of the structured step (0: built-in; see autotune())
        self.tile_nodes = 0
        self.xcd_run = 0
        self.device = torch.device(device)
        self.omega = float(omega)
        self.limiter = tuple(float(x) for x in limiter)
        self.force = tuple(float(x) for x in force)
        self.fp16_scale = float(fp16_scale if fp16_scale is not None else (2.0 ** 10 if storage == "fp16" else 1.0))
        # FP16 storage only: populations are stored as deviations from the
        # equilibrium at this velocity (pick the mean flow)
        self.reference_velocity = tuple(float(x) for x in reference_velocity)
        self.direct = neighbors is None
        self.solid_mask = None
        self.solid_host = None  # bool per node (structured lattice with a solid mask)
        self.nb = None
        self.fluid = None
        self.types = None
        if self.direct:
            if node_types is not None:
                raise ValueError("node_types (typeOfGridNode) belongs to the sparse layout; use solid")
            if dims is None:
                raise ValueError("structured lattice needs dims")
            self.dims = tuple(int(d) for d in dims)
            self.n_nodes = int(np.prod(self.dims))
            if solid is not None:
                solid = np.asarray(solid, dtype=bool).ravel()
                if solid.size != self.n_nodes:
                    raise ValueError("solid mask size does not match dims")
                self.solid_mask = torch.from_numpy(pack_mask(solid).view(np.int32)).to(self.device)
                self.solid_host = solid.copy()
                self.n_fluid = int((~solid).sum())
            else:
                self.n_fluid = self.n_nodes
        else:
            self.dims = (0, 0, 0)
            nbx, nby, nbz = (np.ascontiguousarray(a, dtype=np.uint32) for a in neighbors)
            self.n_nodes = int(n_nodes if n_nodes is not None else nbx.size)
            if not (nbx.size == nby.size == nbz.size == self.n_nodes):
                raise ValueError("neighbour arrays must have n_nodes entries")
            if max(nbx.max(), nby.max(), nbz.max()) >= self.n_nodes:
                raise ValueError("neighbour index out of range")
            self.nb = [torch.from_numpy(a.view(np.int32)).to(self.device) for a in (nbx, nby, nbz)]
            if node_types is not None:
                # VirtualFluids' typeOfGridNode: every node that is not
                # GEO_SOLID (15) or GEO_VOID (16) collides
                ty = np.ascontiguousarray(node_types, dtype=np.uint32).ravel()
                if ty.size != self.n_nodes:
                    raise ValueError("node_types needs n_nodes entries")
                self.types = torch.from_numpy(ty.view(np.int32)).to(self.device)
                self.n_fluid = int(((ty != 15) & (ty != 16)).sum())
            else:
                fl = np.ascontiguousarray(fluid_nodes, dtype=np.uint32)
                if fl.size and fl.max() >= self.n_nodes:
                    raise ValueError("fluid node index out of range")
                self.fluid = torch.from_numpy(fl.view(np.int32)).to(self.device)
                self.n_fluid = int(fl.size)
        self.f = torch.zeros(27 * self.n_nodes, dtype=STORAGE_DTYPE[storage], device=self.device)
        self.macro = None
        self._struct = None
        # eddy-viscosity model of the K17 collision (vf::lbm::TurbulenceModel):
        # the turbulent viscosity of every node starts at 0, as the reference's
        # zero-filled turbViscosity array
        if turbulence not in TURBULENCE:
            raise ValueError(f"turbulence must be one of {sorted(TURBULENCE)}")
        self.turbulence = turbulence
        self.sgs_constant = float(sgs_constant)
        # ApplyBodyForce: per-node body force (3, n_nodes) in the arithmetic
        # type, the global force unscaled and the level's force factor 1/2^level
        self.body_force = None
        self.forces = (0.0, 0.0, 0.0)
        self.force_factor = 1.0
        self.nut = None
        self.nbinv = None
        if turbulence != "none":
            self.nut = torch.zeros(self.n_nodes, dtype=self._macro_dtype(), device=self.device)
        if turbulence == "amd":
            self._alloc_macro()
            if not self.direct:
                if node_types is None or neighbor_inverse is None:
                    raise ValueError("the AMD model on the sparse layout needs node_types and neighbor_inverse")
                inv = np.ascontiguousarray(neighbor_inverse, dtype=np.uint32).ravel()
                if inv.size != self.n_nodes or (inv.size and inv.max() >= self.n_nodes):
                    raise ValueError("neighbor_inverse needs n_nodes entries in range")
                self.nbinv = torch.from_numpy(inv.view(np.int32)).to(self.device)
        # sparse layout: wavefronts of consecutive nodes with one y and one z
        # stride form their indices by arithmetic (vfx_sparse_regular_groups)
        self.rgroups = None
        if not self.direct and regular_groups:
            st = self.struct()
            nbytes = int(V.lib().vfx_sparse_regular_groups_bytes(C.byref(st)))
            if nbytes:
                t = torch.zeros(nbytes // 4, dtype=torch.int32, device=self.device)
                V.check(V.lib().vfx_sparse_regular_groups(C.byref(st), t.data_ptr(), nbytes, _stream(None)),
                        "vfx_sparse_regular_groups")
                self.rgroups = t

    @property
    def regular_fraction(self) -> float:
        """Share of the sparse step's wavefronts that form their indices by arithmetic."""
        if self.rgroups is None:
            return 0.0
        flags = self.rgroups.view(-1, 4)[:, 3].cpu().numpy()
        return float((flags & 1).mean()) if flags.size else 0.0

    def _macro_dtype(self):
        return torch.float64 if self.storage == "fp64" else torch.float32

    def _alloc_macro(self):
        if self.macro is None:
            self.macro = [torch.zeros(self.n_nodes, dtype=self._macro_dtype(), device=self.device) for _ in range(4)]

    # ------------------------------------------------------------ C view --
    def struct(self) -> V.VfxLattice:
        s = V.VfxLattice()
        s.f = self.f.data_ptr()
        s.storage = STORAGE[self.storage]
        s.fp16_scale = self.fp16_scale
        s.n_nodes = self.n_nodes
        s.direct = 1 if self.direct else 0
        s.dims[:] = list(self.dims)
        s.solid_mask = _ptr(self.solid_mask)
        if self.nb is not None:
            s.neighbor_x, s.neighbor_y, s.neighbor_z = (t.data_ptr() for t in self.nb)
            if self.types is not None:
                s.type_of_grid_node = self.types.data_ptr()
            else:
                s.fluid_nodes = _ptr(self.fluid) if self.n_fluid else None
                s.n_fluid = self.n_fluid
        s.omega = self.omega
        s.quadric_limiter[:] = list(self.limiter)
        s.force[:] = list(self.force)
        if self.macro is not None:
            s.rho, s.vx, s.vy, s.vz = (t.data_ptr() for t in self.macro)
        s.kernel_variant = 0 if self.variant == "default" else 1
        s.fp16_reference_velocity[:] = list(getattr(self, "reference_velocity", (0.0, 0.0, 0.0)))
        if getattr(self, "body_force", None) is not None:
            s.body_force[:] = [self.body_force[a].data_ptr() for a in range(3)]
            s.forces[:] = list(self.forces)
            s.force_factor = self.force_factor
        if getattr(self, "nut", None) is not None:
            s.turbulence_model = TURBULENCE[self.turbulence]
            s.sgs_constant = self.sgs_constant
            s.turbulent_viscosity = self.nut.data_ptr()
            s.neighbor_inverse = _ptr(self.nbinv)
        # lattices made without __init__ (domain views of an ensemble) have
        # the built-in launch shape
        s.tile_nodes = getattr(self, "tile_nodes", 0)
        s.xcd_run = getattr(self, "xcd_run", 0)
        s.regular_groups = _ptr(getattr(self, "rgroups", None))
        return s

    def autotune(self, even=True, steps=4, collision_template=0, stream=None):
        """Times the K17 step over every tile width (the sparse step: 256
        work items per workgroup) and XCD run (vfx_k17_autotune) and keeps the
        fastest in this lattice; the populations advance, the parity is
        unchanged.  Returns (tile_nodes, xcd_run, ms per step)."""
        s = self.struct()
        t, r, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        V.check(V.lib().vfx_k17_autotune(C.byref(s), int(bool(even)), int(collision_template), int(steps),
                                         _stream(stream), C.byref(t), C.byref(r), C.byref(ms)), "vfx_k17_autotune")
        self.tile_nodes, self.xcd_run = t.value, r.value
        return t.value, r.value, ms.value

    @property
    def bytes_per_population(self) -> int:
        return BYTES[self.storage]

    # ----------------------------------------------------------- kernels --
    def init_equilibrium(self, drho, ux, uy, uz, even=True, stream=None):
        f64 = [torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).ravel()).to(self.device)
               if not isinstance(a, torch.Tensor) else a.to(self.device, torch.float64).contiguous().view(-1)
               for a in (drho, ux, uy, uz)]
        for a in f64:
            if a.numel() != self.n_nodes:
                raise ValueError("initial fields need n_nodes entries")
        s = self.struct()
        V.check(V.lib().vfx_init_equilibrium(C.byref(s), int(bool(even)), *(a.data_ptr() for a in f64),
                                             _stream(stream)), "vfx_init_equilibrium")

    def set_body_force(self, fx, fy, fz, forces=(0.0, 0.0, 0.0), force_factor=1.0):
        """Per-node body force for the ApplyBodyForce collision (read once, then
        reset to 0 by it), with the global force and the level's force factor."""
        bf = torch.stack([torch.as_tensor(np.broadcast_to(np.asarray(a, np.float64), (self.n_nodes,)).copy())
                          for a in (fx, fy, fz)]).to(self.device, self._macro_dtype()).contiguous()
        self.body_force = bf
        self.forces = tuple(float(x) for x in forces)
        self.force_factor = float(force_factor)

    def collide_stream(self, collision="k17", even=True, write_macro=False, stream=None, apply_body_force=False):
        if write_macro:
            self._alloc_macro()
        if apply_body_force and self.body_force is None:
            raise ValueError("apply_body_force needs set_body_force first")
        tmpl = (V.VFX_TEMPLATE_ALL_FEATURES if write_macro else V.VFX_TEMPLATE_APPLY_BODY_FORCE) if apply_body_force \
            else (V.VFX_TEMPLATE_WRITE_MACRO if write_macro else V.VFX_TEMPLATE_DEFAULT)
        s = self.struct()
        V.check(V.lib().vfx_collide_stream(C.byref(s), COLLISION[collision], int(bool(even)), tmpl, _stream(stream)),
                "vfx_collide_stream")

    def macroscopic_quantities(self, even=True, stream=None):
        """calculateMacroscopicQuantitiesCompressible: (rho, vx, vy, vz, press)
        of every node in the lattice's real type, zero off the fluid nodes."""
        out = [torch.empty(self.n_nodes, dtype=self._macro_dtype(), device=self.device) for _ in range(5)]
        s = self.struct()
        V.check(V.lib().vfx_macroscopic_quantities(C.byref(s), int(bool(even)), *(t.data_ptr() for t in out),
                                                   _stream(stream)), "vfx_macroscopic_quantities")
        return out

    def turbulent_viscosity_amd(self, stream=None):
        """calcTurbulentViscosityAMD: the AMD eddy viscosity of every fluid
        node from the velocities the last collision wrote."""
        if self.nut is None or self.macro is None:
            raise ValueError("lattice has no turbulent viscosity / velocity fields")
        s = self.struct()
        V.check(V.lib().vfx_turbulent_viscosity_amd(C.byref(s), _stream(stream)), "vfx_turbulent_viscosity_amd")

    def apply_bc(self, bc: "BoundaryCondition", even=True, stream=None):
        if bc.max_node >= self.n_nodes:
            raise ValueError("boundary node index out of range")
        if bc.wm is not None:
            if bc.real != self._macro_dtype():
                raise ValueError("wall-model arrays must have the lattice's arithmetic type")
            self._alloc_macro()
        if bc.precursor is not None and bc.precursor.real != self._macro_dtype():
            raise ValueError("precursor weights and planes must have the lattice's arithmetic type")
        s = self.struct()
        b = bc.struct(self._macro_dtype())
        V.check(V.lib().vfx_bc_apply(C.byref(s), C.byref(b), int(bool(even)), _stream(stream)), "vfx_bc_apply")

    def k17_step_fused(self, even, post=(), pre=None, apply_pre=False, any_mask=None, stream=None):
        """K17 collide+stream with boundary groups applied inside the kernel
        (vfx_k17_step_fused): up to two post-collision FusedGroups and one
        pressure FusedGroup whose values the step takes from the previous
        step's populations (apply_pre False on the first step of a run);
        any_mask: the union of the groups' node masks (FusedPlan.any_mask)."""
        real = self._macro_dtype()
        for g in list(post) + ([pre] if pre is not None else []):
            if g.bc.max_node >= self.n_nodes or g.n_nodes != self.n_nodes:
                raise ValueError("fused group does not belong to this lattice")
        s = self.struct()
        ps = [g.struct(real) for g in post]
        arr = (V.VfxFusedBC * len(ps))(*ps) if ps else None
        pr = pre.struct(real) if pre is not None else None
        V.check(V.lib().vfx_k17_step_fused(C.byref(s), int(bool(even)), arr, len(ps),
                                           C.byref(pr) if pr is not None else None, _ptr(any_mask),
                                           int(bool(apply_pre)), _stream(stream)), "vfx_k17_step_fused")

    def read_populations(self, even=True, stream=None) -> torch.Tensor:
        out = torch.zeros(self.n_nodes, 27, dtype=torch.float64, device=self.device)
        s = self.struct()
        V.check(V.lib().vfx_read_populations(C.byref(s), int(bool(even)), out.data_ptr(), _stream(stream)),
                "vfx_read_populations")
        return out

    def write_populations(self, F, even=True, stream=None):
        F = torch.as_tensor(F, dtype=torch.float64).to(self.device).contiguous()
        if F.shape != (self.n_nodes, 27):
            raise ValueError("populations must be (n_nodes, 27)")
        s = self.struct()
        V.check(V.lib().vfx_write_populations(C.byref(s), int(bool(even)), F.data_ptr(), _stream(stream)),
                "vfx_write_populations")

    def export_distributions(self, even=True, dtype=torch.float32, stream=None) -> torch.Tensor:
        """The distribution array as VirtualFluids holds it (27 planes of
        n_nodes shifted populations at this parity) in float32 (VirtualFluids'
        default `real`) or float64, on the device."""
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("VirtualFluids' real is float32 or float64")
        out = torch.empty(27 * self.n_nodes, dtype=dtype, device=self.device)
        s = self.struct()
        V.check(V.lib().vfx_export_distributions(C.byref(s), int(bool(even)), out.data_ptr(),
                                                 STORAGE["fp64" if dtype == torch.float64 else "fp32"], _stream(stream)),
                "vfx_export_distributions")
        return out

    def import_distributions(self, f, even=True, stream=None):
        """Load a VirtualFluids-layout distribution array (float32 or float64,
        27 * n_nodes values) into this lattice at the given parity."""
        f = torch.as_tensor(f)
        if f.dtype not in (torch.float32, torch.float64):
            raise ValueError("VirtualFluids' real is float32 or float64")
        if f.numel() != 27 * self.n_nodes:
            raise ValueError(f"expected {27 * self.n_nodes} values, got {f.numel()}")
        f = f.to(self.device).contiguous().view(-1)
        s = self.struct()
        V.check(V.lib().vfx_import_distributions(C.byref(s), int(bool(even)), f.data_ptr(),
                                                 STORAGE["fp64" if f.dtype == torch.float64 else "fp32"], _stream(stream)),
                "vfx_import_distributions")

    def macroscopic(self, even=True, stream=None):
        outs = [torch.zeros(self.n_nodes, dtype=torch.float64, device=self.device) for _ in range(4)]
        s = self.struct()
        V.check(V.lib().vfx_macroscopic(C.byref(s), int(bool(even)), *(o.data_ptr() for o in outs),
                                        _stream(stream)), "vfx_macroscopic")
        return outs


    # ------------------------------------------------------- diagnostics --
    def _workspace(self):
        s = self.struct()
        nbytes = int(V.lib().vfx_reduce_workspace_bytes(C.byref(s)))
        ws = getattr(self, "_ws", None)
        if ws is None or ws.numel() * 8 < nbytes:
            ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=self.device)
            self._ws = ws
        return ws

    def mass_momentum(self, even=True, stream=None) -> torch.Tensor:
        """FP64 {sum drho, sum jx, sum jy, sum jz} of the pre-collision state (device tensor)."""
        out = torch.zeros(4, dtype=torch.float64, device=self.device)
        ws = self._workspace()
        s = self.struct()
        V.check(V.lib().vfx_reduce_mass_momentum(C.byref(s), int(bool(even)), ws.data_ptr(), out.data_ptr(),
                                                 _stream(stream)), "vfx_reduce_mass_momentum")
        return out

    def _energy(self, fn, name, even, fields, stream):
        if fields is None:
            fields = self.macroscopic(even, stream)
        out = torch.zeros(2, dtype=torch.float64, device=self.device)
        ws = self._workspace()
        s = self.struct()
        V.check(fn(C.byref(s), *(f.data_ptr() for f in fields), ws.data_ptr(), out.data_ptr(), _stream(stream)), name)
        return out

    def kinetic_energy(self, even=True, fields=None, stream=None) -> torch.Tensor:
        """{sum 1/2 rho |u|^2, count} over fluid nodes with a fluid 7-stencil (KineticEnergyAnalyzer)."""
        return self._energy(V.lib().vfx_kinetic_energy, "vfx_kinetic_energy", even, fields, stream)

    def enstrophy(self, even=True, fields=None, stream=None) -> torch.Tensor:
        """{sum 1/2 rho |curl u|^2, count} (EnstrophyAnalyzer)."""
        return self._energy(V.lib().vfx_enstrophy, "vfx_enstrophy", even, fields, stream)


class Statistics:
    """FP64 running sums of u and u_i u_j (CalcTurbulenceIntensity); mean and
    Reynolds stress on demand."""

    def __init__(self, lattice: Lattice):
        self.lattice = lattice
        self.sums = torch.zeros(9, lattice.n_nodes, dtype=torch.float64, device=lattice.device)
        self.samples = 0

    def accumulate(self, even, stream=None):
        s = self.lattice.struct()
        V.check(V.lib().vfx_accumulate_stats(C.byref(s), int(bool(even)), self.sums.data_ptr(), _stream(stream)),
                "vfx_accumulate_stats")
        self.samples += 1

    def mean(self) -> torch.Tensor:
        return self.sums[:3] / max(self.samples, 1)

    def reynolds_stress(self) -> torch.Tensor:
        """(6, n): <u'u'> for xx, yy, zz, xy, xz, yz."""
        m = self.mean()
        t = max(self.samples, 1)
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        return torch.stack([self.sums[3 + c] / t - m[a] * m[b] for c, (a, b) in enumerate(pairs)])


class MeanValues:
    """The time averages VirtualFluids writes to its _bin_mean files
    (Parameter::getCalcMean): running sums meanVelocityIn{X,Y,Z}direction,
    meanDensity, meanPressure in the lattice's real type, advanced by
    calculateMeanCompressible at every step of the averaging window
    (Simulation.cpp:448-470), divided by the steps since the previous output
    (calcMean, CalcMean.cpp:57-67) and reset (resetMean, :70-80)."""

    NAMES = ("vx", "vy", "vz", "rho", "press")

    def __init__(self, lattice: Lattice):
        self.lattice = lattice
        real = lattice._macro_dtype()
        self.sums = [torch.zeros(lattice.n_nodes, dtype=real, device=lattice.device) for _ in self.NAMES]

    def accumulate(self, even, stream=None):
        s = self.lattice.struct()
        V.check(V.lib().vfx_calculate_mean(C.byref(s), int(bool(even)), *(t.data_ptr() for t in self.sums),
                                           _stream(stream)), "vfx_calculate_mean")

    def mean(self, tdiff) -> dict:
        """sum / (real)tdiff of every field (host division in the real type)."""
        d = torch.tensor(float(tdiff), dtype=self.sums[0].dtype, device=self.sums[0].device)
        return {name: t / d for name, t in zip(self.NAMES, self.sums)}

    def reset(self, stream=None):
        s = self.lattice.struct()
        V.check(V.lib().vfx_reset_mean(C.byref(s), *(t.data_ptr() for t in self.sums), _stream(stream)),
                "vfx_reset_mean")


class BoundaryCondition:
    """A boundary-node list in the QforBoundaryConditions shape
    (src/gpu/core/Calculation/Calculation.h:105-117).

    kind: 0 velocity bounce-back, 1 no-slip interpolated, 2 velocity
    interpolated, 3 slip, 4 non-equilibrium pressure (pre-collision; needs
    ``rho``, ``neighbors`` (inward neighbour node) and ``normal`` (outward
    direction index)), 5 velocity interpolated with pressure, 6 non-reflecting
    outflow and 7 non-reflecting outflow with pressure correction
    (pre-collision; ``neighbors``, ``normal`` a face direction 1..6, kind 7
    also ``pressure_correction_factor``), 8 slip with the node's turbulent
    viscosity (the lattice needs a turbulence model), 9 / 10 / 11 wall models
    (StressCompressible, StressBounceBackCompressible,
    StressBounceBackPressureCompressible): ``neighbors`` is the sampling node,
    ``wall_normal`` (3, n), ``sampling_offset``, ``z0``, optional initial
    ``exchange_velocity`` / ``wall_velocity`` (3, n); ``real`` is the
    lattice's arithmetic type (float64 for FP64 storage, else float32).  The
    filtered velocities (state) and the monitors ``u_star`` and ``wall_force``
    live on the device and are updated by every call.

    The outflow kinds write the populations the collision after next reads
    (the kernel runs after a parity flip and fills the other parity's slots):
    applied only after every step, as VirtualFluids does, the second step
    reads whatever those elements held before.  Stepper applies them once with
    the starting parity before its first step to define them
    (Stepper.prime_outflow)."""

    PRE_COLLISION = {V.VFX_BC_PRESSURE_NEQ, V.VFX_BC_OUTFLOW_NON_REFLECTING,
                     V.VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION}
    OUTFLOW = {V.VFX_BC_OUTFLOW_NON_REFLECTING, V.VFX_BC_OUTFLOW_NON_REFLECTING_PRESSURE_CORRECTION}
    STRESS = {V.VFX_BC_STRESS_INTERPOLATED, V.VFX_BC_STRESS_BOUNCE_BACK, V.VFX_BC_STRESS_BOUNCE_BACK_PRESSURE}
    PRECURSOR = {V.VFX_BC_PRECURSOR_VELOCITY, V.VFX_BC_PRECURSOR_DISTRIBUTIONS}

    def __init__(self, kind, nodes, q27=None, vx=None, vy=None, vz=None, cpu_float_velocity=False, rho=None,
                 neighbors=None, normal=0, device="cuda", pressure_correction_factor=0.0, wall_normal=None,
                 sampling_offset=None, z0=None, exchange_velocity=None, wall_velocity=None, real=torch.float64,
                 precursor=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        n = nodes.size
        self.nodes = nodes
        self.kn_host = None if neighbors is None else np.ascontiguousarray(neighbors, dtype=np.uint32).ravel()
        self.kind = int(kind)
        self.precursor = precursor
        if self.kind in self.PRECURSOR:
            if precursor is None or precursor.n != n:
                raise ValueError("a precursor boundary needs PrecursorPlanes with one stencil per boundary node")
            want = 3 if self.kind == V.VFX_BC_PRECURSOR_VELOCITY else 9
            if precursor.components != want:
                raise ValueError(f"this precursor kind reads {want} plane components")
            if self.kind == V.VFX_BC_PRECURSOR_DISTRIBUTIONS and q27 is None:
                q27 = np.full((27, n), -1.0, np.float32)
        self.n = n
        dev = torch.device(device)
        self.max_node = int(nodes.max()) if n else -1
        self.k = torch.from_numpy(nodes.view(np.int32)).to(dev)
        self.wm = None
        if self.kind in self.STRESS:
            if neighbors is None or wall_normal is None or sampling_offset is None or z0 is None:
                raise ValueError("a wall model needs neighbors (sampling nodes), wall_normal, sampling_offset, z0")
            kn = np.ascontiguousarray(neighbors, dtype=np.uint32)
            if kn.size != n:
                raise ValueError("one sampling node per boundary node")
            self.max_node = max(self.max_node, int(kn.max()) if n else -1)
            self.kn = torch.from_numpy(kn.view(np.int32)).to(dev)

            def rows(a, default):
                a = np.zeros((3, n)) if a is None and default else np.asarray(a, np.float64).reshape(3, -1)
                return torch.from_numpy(np.array(np.broadcast_to(a, (3, n)), dtype=np.float64)).to(dev, real).contiguous()

            self.wm = {"normal": rows(wall_normal, False), "exchange": rows(exchange_velocity, True),
                       "wall": rows(wall_velocity, True),
                       "offset": torch.from_numpy(np.array(np.broadcast_to(sampling_offset, (n,)), np.int32)).to(dev),
                       "z0": torch.from_numpy(np.array(np.broadcast_to(z0, (n,)), np.float64)).to(dev, real),
                       "u_star": torch.zeros(n, dtype=real, device=dev),
                       "force": torch.zeros((3, n), dtype=real, device=dev)}
            self.real = real
        if self.kind == V.VFX_BC_PRESSURE_NEQ or self.kind in self.OUTFLOW:
            if self.kind == V.VFX_BC_PRESSURE_NEQ and (rho is None or neighbors is None or not 1 <= int(normal) <= 26):
                raise ValueError("pressure boundary needs rho, neighbors and normal in 1..26")
            if self.kind in self.OUTFLOW and (neighbors is None or not 1 <= int(normal) <= 6):
                raise ValueError("outflow boundary needs neighbors and a face normal in 1..6")
            q27 = np.zeros((27, n), np.float32) if q27 is None else q27
            kn = np.ascontiguousarray(neighbors, dtype=np.uint32)
            if kn.size != n:
                raise ValueError("one inward neighbour per boundary node")
            self.max_node = max(self.max_node, int(kn.max()) if n else -1)
            self.kn = torch.from_numpy(kn.view(np.int32)).to(dev)
        elif self.kind not in self.STRESS:
            self.kn = None
        q = np.ascontiguousarray(q27, dtype=np.float64)
        if q.shape != (27, n):
            raise ValueError("q27 must be (27, n)")

        def arr(v):
            if v is None:
                return None
            return np.array(np.broadcast_to(v, (n,)), dtype=np.float64)

        # q27, Vx/Vy/Vz and RhoBC as given (float64); the device copies take
        # the lattice's `real` (float64 for FP64 storage, float32 otherwise),
        # made once per type and handed to the C ABI as they are
        self._host = {"q": q, "vx": arr(vx), "vy": arr(vy), "vz": arr(vz), "rho": arr(rho)}
        self._dev = {}
        self.device = dev
        self.normal = int(normal)
        self.pressure_correction_factor = float(pressure_correction_factor)
        self.cpu_float_velocity = bool(cpu_float_velocity)

    @property
    def pre_collision(self) -> bool:
        return self.kind in self.PRE_COLLISION

    def subset(self, idx) -> "BoundaryCondition":
        """The entries idx (in that order) as a boundary of their own (interpolated
        and pressure kinds: no wall-model state, no precursor planes)."""
        if self.wm is not None or self.precursor is not None:
            raise ValueError("subset() covers boundaries without wall-model or precursor state")
        idx = np.asarray(idx, dtype=np.int64)
        h = self._host
        pick = (lambda a: None if a is None else a[idx])
        return BoundaryCondition(self.kind, self.nodes[idx], h["q"][:, idx], pick(h["vx"]), pick(h["vy"]),
                                 pick(h["vz"]), cpu_float_velocity=self.cpu_float_velocity, rho=pick(h["rho"]),
                                 neighbors=None if self.kn_host is None else self.kn_host[idx], normal=self.normal,
                                 device=self.device, pressure_correction_factor=self.pressure_correction_factor)

    def tables(self, real=torch.float64) -> dict:
        """Device q27 (27, n), vx, vy, vz, rho (n) in the arithmetic type `real`."""
        t = self._dev.get(real)
        if t is None:
            t = {k: None if v is None else torch.from_numpy(v).to(self.device, real).contiguous()
                 for k, v in self._host.items()}
            self._dev[real] = t
        return t

    def struct(self, real=torch.float64) -> V.VfxBC:
        """The vfx_bc of this boundary for a lattice whose arithmetic type is `real`."""
        t = self.tables(real)
        b = V.VfxBC()
        b.kind = self.kind
        b.n = self.n
        b.k = self.k.data_ptr() if self.n else None
        b.q27 = t["q"].data_ptr() if self.n else None
        b.vx, b.vy, b.vz = _ptr(t["vx"]), _ptr(t["vy"]), _ptr(t["vz"])
        b.cpu_float_velocity = int(self.cpu_float_velocity)
        b.rho = _ptr(t["rho"])
        b.k_neighbor = _ptr(self.kn) if self.n else None
        b.normal = self.normal
        b.pressure_correction_factor = self.pressure_correction_factor
        if self.wm is not None and self.n:
            w = self.wm
            b.wm_normal[:] = [w["normal"][j].data_ptr() for j in range(3)]
            b.wm_exchange_velocity[:] = [w["exchange"][j].data_ptr() for j in range(3)]
            b.wm_wall_velocity[:] = [w["wall"][j].data_ptr() for j in range(3)]
            b.wm_sampling_offset = w["offset"].data_ptr()
            b.wm_z0 = w["z0"].data_ptr()
            b.wm_u_star = w["u_star"].data_ptr()
            b.wm_force[:] = [w["force"][j].data_ptr() for j in range(3)]
        if self.precursor is not None and self.n:
            pc = self.precursor
            b.n_precursor_nodes = pc.n_precursor_nodes
            b.plane_neighbor[:] = [pc.neighbors[j].data_ptr() for j in range(4)]
            b.plane_weight[:] = [pc.weights[j].data_ptr() for j in range(4)]
            b.precursor_last = pc.last.data_ptr()
            b.precursor_current = pc.current.data_ptr()
            b.time_ratio = pc.time_ratio
            b.velocity_ratio = pc.velocity_ratio
        return b


class PrecursorPlanes:
    """Plane data of a precursor inflow (QforPrecursorBoundaryConditions,
    src/gpu/core/Calculation/Calculation.h:120-150) and its time cycling.

    neighbors / weights: (4, n) precursor node and inverse-distance weight of
    the 0PP, 0PM, 0MP, 0MM stencil of every boundary node; components: 3
    (velocities, VFX_BC_PRECURSOR_VELOCITY) or 9 (the positive-x populations,
    VFX_BC_PRECURSOR_DISTRIBUTIONS).  Planes are (components, n_precursor_nodes)
    arrays of the lattice's real type.

    With a reader -- reader(load_time) returning the plane array at a physical
    time -- the cycle follows VirtualFluids: GridGenerator loads t = 0 into
    last, time_steps_between_reads into current and twice that into next
    (GridGenerator.cpp:436-478, nPrecursorReads = 2); runPrecursorBCKernelPost
    (BoundaryConditionKernelManager.cpp:346-382) rotates last <- current <-
    next and reads the following plane once the level's time step reaches
    current's, and passes time_ratio = (t - lastTime) / time_steps_between_reads.
    Without a reader, set last, current and time_ratio directly."""

    def __init__(self, neighbors, weights, n_precursor_nodes, components, reader=None, time_steps_between_reads=1,
                 velocity_ratio=1.0, level=0, physical_time_ratio=1.0, real=torch.float64, device="cuda"):
        dev = torch.device(device)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32).reshape(4, -1)
        if nb.size and int(nb.max()) >= int(n_precursor_nodes):
            raise ValueError("plane neighbour beyond the precursor nodes")
        self.n = nb.shape[1]
        self.n_precursor_nodes = int(n_precursor_nodes)
        self.components = int(components)
        self.neighbors = torch.from_numpy(nb.view(np.int32)).to(dev)
        self.weights = torch.from_numpy(np.asarray(weights, np.float64).reshape(4, -1).copy()).to(dev, real)
        self.real, self.device = real, dev
        self.velocity_ratio = float(velocity_ratio)
        self.time_ratio = 0.0
        self.reader = reader
        self.steps_between_reads = int(time_steps_between_reads)
        self.level = int(level)
        self.physical_time_ratio = float(physical_time_ratio)
        zeros = np.zeros((self.components, self.n_precursor_nodes))
        self.last, self.current, self.next = (self.plane(zeros) for _ in range(3))
        self.n_reads = 0
        if reader is not None:
            dt = self.steps_between_reads * 2.0 ** -self.level * self.physical_time_ratio
            self.last, self.current, self.next = self.plane(reader(0.0)), self.plane(reader(dt)), self.plane(
                reader(2 * dt))
            self.n_reads = 2

    def plane(self, a):
        a = np.asarray(a, np.float64).reshape(self.components, self.n_precursor_nodes)
        return torch.from_numpy(a.copy()).to(self.device, self.real).contiguous()

    def _load(self, dst, a, stream=None):
        """Copy a plane read into the existing device array dst, on the stream
        the boundary kernels run on (torch's current stream when None), so the
        copy waits for the kernels that still read dst's previous contents."""
        src = torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float64).reshape(
            self.components, self.n_precursor_nodes))).to(self.real)
        if stream is None:
            dst.copy_(src)
        else:
            s = stream if isinstance(stream, torch.cuda.Stream) else torch.cuda.ExternalStream(stream)
            with torch.cuda.stream(s):
                dst.copy_(src)

    def advance(self, t_level, stream=None):
        """The cycling of runPrecursorBCKernelPost at level time step t_level.
        The three plane arrays are allocated once and rotate; a new read goes
        into the array that held `last`."""
        T = self.steps_between_reads
        last_time, current_time = (self.n_reads - 2) * T, (self.n_reads - 1) * T
        next_time = self.n_reads * T
        if t_level >= current_time:
            last_time, current_time, next_time = current_time, next_time, next_time + T
            self.last, self.current, self.next = self.current, self.next, self.last
            self._load(self.next, self.reader(next_time * 2.0 ** -self.level * self.physical_time_ratio), stream)
            self.n_reads += 1
        self.time_ratio = (t_level - last_time) / T
        return self.time_ratio


FUSED_POST_KINDS = (V.VFX_BC_NOSLIP_INTERPOLATED, V.VFX_BC_VELOCITY_INTERPOLATED, V.VFX_BC_SLIP,
                    V.VFX_BC_VELOCITY_PRESSURE_INTERPOLATED)


class FusedGroup:
    """A boundary whose entries the collide-stream kernel applies itself
    (vfx_fused_bc): entries sorted by node, one per node; node_mask has bit k
    set for every node k of the group, word_base[w] counts the entries of the
    nodes below 32 w."""

    def __init__(self, bc: BoundaryCondition, n_nodes: int):
        nodes = bc.nodes.astype(np.int64)
        if nodes.size and (np.any(np.diff(nodes) <= 0) or nodes[-1] >= n_nodes):
            raise ValueError("a fused group needs distinct nodes in increasing order inside the lattice")
        nw = (n_nodes + 31) // 32
        words = nodes >> 5
        mask = np.zeros(nw, np.uint32)
        np.bitwise_or.at(mask, words, np.left_shift(np.uint32(1), (nodes & 31).astype(np.uint32)))
        base = np.zeros(nw, np.uint32)
        base[1:] = np.cumsum(np.bincount(words, minlength=nw))[:-1]
        self.bc = bc
        self.n_nodes = n_nodes
        self.node_mask = torch.from_numpy(mask.view(np.int32)).to(bc.device)
        self.word_base = torch.from_numpy(base.view(np.int32)).to(bc.device)
        self._keep = None

    def struct(self, real=torch.float64) -> V.VfxFusedBC:
        self._keep = self.bc.struct(real)
        f = V.VfxFusedBC()
        f.bc = C.pointer(self._keep)
        f.node_mask = self.node_mask.data_ptr()
        f.word_base = self.word_base.data_ptr()
        return f


def _xyz(lattice, k):
    nx, ny, _ = lattice.dims
    k = np.asarray(k, dtype=np.int64)
    return k % nx, (k // nx) % ny, k // (nx * ny)


def _unique_nodes(nodes) -> np.ndarray:
    """True for the entries whose node appears once in the table."""
    _, inv, cnt = np.unique(nodes, return_inverse=True, return_counts=True)
    return cnt[inv] == 1


def fusable_post_entries(lattice, bc) -> np.ndarray:
    """Entries of an interpolated boundary the collide-stream kernel may apply
    itself: the node is fluid, appears once, and the far node of every wall
    link (0 <= q <= 1, in the storage real's rounding too) is solid, so that no
    other thread of the launch touches the elements the boundary writes or
    the node's own post-collision elements it reads back."""
    if lattice.solid_host is None or bc.n == 0:
        return np.zeros(bc.n, dtype=bool)
    nx, ny, nz = lattice.dims
    x, y, z = _xyz(lattice, bc.nodes)
    q64 = bc._host["q"]
    q32 = q64.astype(np.float32)
    ok = _unique_nodes(bc.nodes) & ~lattice.solid_host[bc.nodes]
    for d in range(1, 27):
        link = ((q64[d] >= 0) & (q64[d] <= 1)) | ((q32[d] >= 0) & (q32[d] <= 1))
        if not link.any():
            continue
        far = (x + CX[d]) % nx + nx * ((y + CY[d]) % ny + ny * ((z + CZ[d]) % nz))
        ok &= ~link | lattice.solid_host[far]
    return ok


def fusable_pressure_entries(lattice, bc) -> np.ndarray:
    """Entries of a non-equilibrium pressure boundary on a face normal to +-x
    the kernel may apply itself: fluid node appearing once, inward neighbour
    k -+ 1 in the same row and the same 64-node wavefront, fluid and no node of
    the boundary.  All False when any neighbour is a boundary node."""
    n = bc.n
    if bc.kind != V.VFX_BC_PRESSURE_NEQ or bc.normal not in (1, 2) or n == 0:
        return np.zeros(n, dtype=bool)
    k = bc.nodes.astype(np.int64)
    kn = bc.kn_host.astype(np.int64)
    if np.isin(kn, k).any():
        return np.zeros(n, dtype=bool)
    nx = lattice.dims[0]
    x = k % nx
    if bc.normal == 1:
        ok = (kn == k - 1) & (x >= 1) & (x % 64 != 0)
    else:
        ok = (kn == k + 1) & (x + 1 < nx) & (x % 64 != 63)
    ok &= _unique_nodes(k) & (kn < lattice.n_nodes) & (k < lattice.n_nodes)
    if lattice.solid_host is not None:
        ok &= ~lattice.solid_host[np.minimum(k, lattice.n_nodes - 1)]
        ok &= ~lattice.solid_host[np.minimum(kn, lattice.n_nodes - 1)]
    return ok


class BCSet:
    """Boundary groups of one or more lattices of one storage type applied by
    one launch (vfx_bc_set): kinds 0-5, each group at its own parity.  The
    groups must not write what another group of the set reads (the
    post-collision groups of a step never do; keep pressure groups apart from
    them)."""

    KINDS = {V.VFX_BC_VELOCITY_BOUNCE_BACK, V.VFX_BC_NOSLIP_INTERPOLATED, V.VFX_BC_VELOCITY_INTERPOLATED,
             V.VFX_BC_SLIP, V.VFX_BC_PRESSURE_NEQ, V.VFX_BC_VELOCITY_PRESSURE_INTERPOLATED}

    def __init__(self, pairs):
        self.pairs = list(pairs)
        self.handle = None
        if not self.pairs:
            return
        for lat, bc in self.pairs:
            if bc.kind not in self.KINDS:
                raise ValueError(f"kind {bc.kind} does not run in a boundary set")
            if bc.max_node >= lat.n_nodes:
                raise ValueError("boundary node index out of range")
        self._lats = [lat.struct() for lat, _ in self.pairs]
        self._bcs = [bc.struct(lat._macro_dtype()) for lat, bc in self.pairs]
        n = len(self.pairs)
        la = (C.POINTER(V.VfxLattice) * n)(*[C.pointer(x) for x in self._lats])
        ba = (C.POINTER(V.VfxBC) * n)(*[C.pointer(x) for x in self._bcs])
        h = C.c_void_p()
        V.check(V.lib().vfx_bc_set_create(la, ba, n, C.byref(h)), "vfx_bc_set_create")
        self.handle = h

    def apply(self, parities, stream=None):
        """parities: one bool per group (its is_even_timestep)."""
        if self.handle is None:
            return
        mask = 0
        for j, even in enumerate(parities):
            mask |= int(bool(even)) << j
        V.check(V.lib().vfx_bc_set_apply(self.handle, mask, _stream(stream)), "vfx_bc_set_apply")

    def __del__(self):
        if getattr(self, "handle", None) is not None:
            V.lib().vfx_bc_set_destroy(self.handle)
            self.handle = None


def step_levels(levels, stream=None):
    """The K17 collide-and-stream of one or two lattices in one launch
    (vfx_k17_step_levels): levels = [(lattice, is_even)], each the plain
    default-template step of that lattice at its parity."""
    n = len(levels)
    if not 1 <= n <= 2:
        raise ValueError("one or two levels per launch")
    structs = [lat.struct() for lat, _ in levels]
    arr = (V.VfxLevelStep * n)()
    for j, ((lat, even), s) in enumerate(zip(levels, structs)):
        arr[j].lattice = C.pointer(s)
        arr[j].is_even_timestep = int(bool(even))
    V.check(V.lib().vfx_k17_step_levels(arr, n, _stream(stream)), "vfx_k17_step_levels")


@dataclass
class FusedPlan:
    """Which boundary entries a fused step applies inside the collide-stream
    kernel (post: up to two groups, pre: one pressure group) and which run as
    boundary kernels of their own (residual_post / residual_pre, in the order
    of the unfused step)."""

    post: list
    pre: FusedGroup | None
    residual_post: list
    residual_pre: list
    any_mask: torch.Tensor | None = None  # union of the groups' node masks

    @staticmethod
    def build(lattice, bcs, post=True) -> "FusedPlan":
        """post=False: only the pressure group is fused (the interpolated
        kinds run as separate kernels)."""
        def split(b, ok):
            sel = np.flatnonzero(ok)
            sel = sel[np.argsort(b.nodes[sel], kind="stable")]
            rest = np.flatnonzero(~ok)
            return b.subset(sel), (b.subset(rest) if rest.size else None)

        post_all = [b for b in bcs if not b.pre_collision and b.kind not in BoundaryCondition.PRECURSOR]
        prec = [b for b in bcs if b.kind in BoundaryCondition.PRECURSOR]
        pre_all = [b for b in bcs if b.pre_collision]
        # the two interpolated boundaries with the most fusable entries
        cand = []
        for j, b in enumerate(post_all):
            if post and b.kind in FUSED_POST_KINDS and b.wm is None:
                ok = fusable_post_entries(lattice, b)
                if ok.any():
                    cand.append((int(ok.sum()), j, ok))
        chosen = {j: ok for _, j, ok in sorted(cand, key=lambda c: (-c[0], c[1]))[:2]}
        post, residual_post = [], []
        for j, b in enumerate(post_all):
            if j in chosen:
                fused, rest = split(b, chosen[j])
                post.append(FusedGroup(fused, lattice.n_nodes))
                if rest is not None:
                    residual_post.append(rest)
            else:
                residual_post.append(b)
        residual_post += prec
        pre, residual_pre = None, list(pre_all)
        # the pressure group is fused only as the sole pre-collision boundary
        if len(pre_all) == 1:
            b = pre_all[0]
            ok = fusable_pressure_entries(lattice, b)
            if ok.any():
                fused, rest = split(b, ok)
                pre = FusedGroup(fused, lattice.n_nodes)
                residual_pre = [rest] if rest is not None else []
        groups = post + ([pre] if pre is not None else [])
        any_mask = None
        if groups:
            u = groups[0].node_mask.clone()
            for g in groups[1:]:
                u |= g.node_mask
            any_mask = u
        return FusedPlan(post, pre, residual_post, residual_pre, any_mask)


@dataclass
class Stepper:
    """The per-level time step of UpdateGrid27::updateGrid without refinement
    (src/gpu/core/Calculation/UpdateGrid27.cpp:54-95): collide+stream, the
    post-collision boundary conditions, the parity flip, the AMD viscosity
    kernel when the lattice uses that model, then the pre-collision boundary
    conditions with the new parity."""

    lattice: Lattice
    collision: str = "k17"
    bcs: list = field(default_factory=list)
    even: bool = True
    apply_body_force: bool = False  # CollisionTemplate ApplyBodyForce (Lattice.set_body_force)
    t: int = 1  # time step of the next step; precursor planes with a reader cycle on it
    # Apply the interpolated and pressure boundaries inside the collide-stream
    # kernel where the geometry allows (FusedPlan); same populations as the
    # separate boundary kernels.
    fuse: bool = False
    # with fuse: also the interpolated kinds (False: the pressure group only).
    # Inside the kernel a boundary node costs its whole wavefront the boundary
    # code; on a face normal to x that is one node in 64, where a separate
    # kernel runs 64 boundary nodes per wavefront.
    fuse_post: bool = True
    # Before the first step, apply the non-reflecting outflow kinds once with
    # the starting parity: they write the populations the step after next
    # reads, so without it the second step would read whatever those elements
    # held (False leaves that to the caller, as VirtualFluids does).
    prime_outflow: bool = True
    _primed: bool = field(default=False, init=False, repr=False)
    _plan: FusedPlan | None = field(default=None, init=False, repr=False)
    _plan_key: tuple | None = field(default=None, init=False, repr=False)

    def fused_plan(self) -> FusedPlan:
        lat = self.lattice
        if not lat.direct or self.collision != "k17" or lat.variant != "default" or \
                getattr(lat, "turbulence", "none") != "none" or self.apply_body_force:
            raise ValueError("fused boundaries need a structured lattice stepped by the default K17 "
                             "without turbulence model or body-force arrays")
        if any(getattr(b, "wm", None) is not None for b in self.bcs):
            raise ValueError("fused boundaries do not run with wall models")
        key = (id(lat), tuple(id(b) for b in self.bcs), self.fuse_post)
        if self._plan_key != key:
            self._plan = FusedPlan.build(lat, self.bcs, post=self.fuse_post)
            self._plan_key = key
        return self._plan

    def _prime(self, stream):
        if self.prime_outflow and not self._primed:
            for bc in self.bcs:
                if bc.kind in BoundaryCondition.OUTFLOW:
                    self.lattice.apply_bc(bc, self.even, stream=stream)
        self._primed = True

    def _step_fused(self, n, stream):
        plan = self.fused_plan()
        lat = self.lattice
        self._prime(stream)
        for i in range(n):
            # the pressure group of the previous step of this run is applied
            # while the node's populations are gathered
            lat.k17_step_fused(self.even, plan.post, plan.pre, apply_pre=i > 0, any_mask=plan.any_mask,
                               stream=stream)
            for bc in plan.residual_post:
                if bc.precursor is not None and bc.precursor.reader is not None:
                    bc.precursor.advance(self.t, stream)
                lat.apply_bc(bc, self.even, stream=stream)
            self.t += 1
            self.even = not self.even
            for bc in plan.residual_pre:
                lat.apply_bc(bc, self.even, stream=stream)
        # leave the populations as the separate kernels would
        if n > 0 and plan.pre is not None:
            lat.apply_bc(plan.pre.bc, self.even, stream=stream)

    def step(self, n=1, stream=None):
        if self.fuse:
            return self._step_fused(n, stream)
        self._prime(stream)
        # precursor inflows run last among the post-collision conditions (UpdateGrid27.cpp:252)
        post = [b for b in self.bcs if not b.pre_collision and b.kind not in BoundaryCondition.PRECURSOR] + \
               [b for b in self.bcs if b.kind in BoundaryCondition.PRECURSOR]
        pre = [b for b in self.bcs if b.pre_collision]
        # wall models sample the velocities the collision writes
        write_macro = any(getattr(b, "wm", None) is not None for b in self.bcs)
        for _ in range(n):
            self.lattice.collide_stream(self.collision, self.even, write_macro=write_macro, stream=stream,
                                        apply_body_force=self.apply_body_force)
            for bc in post:
                if bc.precursor is not None and bc.precursor.reader is not None:
                    bc.precursor.advance(self.t, stream)
                self.lattice.apply_bc(bc, self.even, stream=stream)
            self.t += 1
            self.even = not self.even
            if self.collision == "k17" and getattr(self.lattice, "turbulence", "none") == "amd":
                self.lattice.turbulent_viscosity_amd(stream=stream)
            for bc in pre:
                self.lattice.apply_bc(bc, self.even, stream=stream)

    def populations(self) -> torch.Tensor:
        return self.lattice.read_populations(self.even)
