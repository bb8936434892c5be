"""Two-level (2:1) grid refinement: a nested fine box inside a coarse box.

Geometry (coarse units, coarse nodes at integer positions): the fine level
covers the coarse cells [lo, hi) on every axis with 2 (hi - lo) fine nodes per
axis at positions lo + 1/4 + m/2.  Every coarse cell of the nest holds a 2x2x2
block of fine nodes centred on the cell centre; every fine cell whose centre is
a coarse node (m odd) can be restricted to that coarse node.

Per coarse step the fine level takes two steps, the coarse level one, then the
interface is rebuilt (UpdateGrid27::updateGrid recursion,
src/gpu/core/Calculation/UpdateGrid27.cpp:54-95, and Refinement_noExchange,
src/gpu/core/GridScaling/RefinementStrategy.cpp:169-178):

* fine to coarse: coarse nodes two cells inside the nest (depth 2) are
  restricted from the fine cell around them (fine nodes m = 2 depth - 1, 2 depth,
  at least two layers inside the fine boundary, so still valid after the two
  fine steps);
* coarse to fine: the coarse cells of the outermost nest layer (depth 0-1
  corners, all active coarse nodes) are prolongated onto their 2x2x2 fine
  blocks, the two outermost fine layers, which the two fine steps contaminated
  from outside the nest.

Coarse nodes deeper than 2 are inactive (skipped by the coarse kernel; nothing
valid reads them).  The depth-2 coarse nodes collide with the rest but read
inactive neighbours, so their state is replaced by the restriction every step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as V
from .batch import Graph, capture
from .lattice import BCSet, FusedPlan, Lattice, _stream, step_levels

INACTIVE = 2  # value in the coarse node-state array


FACES = ("x-", "x+", "y-", "y+", "z-", "z+")


class NestedBox:
    """Index tables of one fine nest inside a coarse box (x fastest).

    wall_faces: nest faces that lie on a wall shared by both levels (for
    example the floor under a wall-mounted obstacle): no interface there, the
    nest reaches the wall layer.  coarse_solid / fine_solid (bool, box-shaped):
    solid nodes of each level; an interface cell with a solid source node is
    moved to the first all-fluid neighbour cell in D3Q27 direction order and
    interpolated back with the offset -c_d, as VirtualFluids' grid generator
    does (GridInterface.cpp:350-414, LevelGridBuilder.cpp:400-422)."""

    def __init__(self, coarse_dims, lo, hi, wall_faces=(), coarse_solid=None, fine_solid=None):
        self.coarse_dims = tuple(int(v) for v in coarse_dims)
        self.lo = np.array(lo, dtype=np.int64)
        self.hi = np.array(hi, dtype=np.int64)
        self.wall_faces = tuple(wall_faces)
        for f in self.wall_faces:
            if f not in FACES:
                raise ValueError(f"unknown face {f}")
        if np.any(self.hi - self.lo < 6):
            raise ValueError("the nest needs at least 6 coarse cells per axis")
        dims = np.array(self.coarse_dims)
        for a, name in enumerate("xyz"):
            if f"{name}-" not in self.wall_faces and self.lo[a] < 1:
                raise ValueError("an interface face must stay one coarse node clear of the box faces")
            if f"{name}+" not in self.wall_faces and self.hi[a] > dims[a] - 2:
                raise ValueError("an interface face must stay one coarse node clear of the box faces")
            if self.lo[a] < 0 or self.hi[a] > dims[a] - 1:
                raise ValueError("the nest must lie inside the coarse box")
        self.fine_dims = tuple(int(v) for v in 2 * (self.hi - self.lo))
        nx, ny, nz = self.coarse_dims
        fx, fy, fz = self.fine_dims
        csol = np.zeros((nz, ny, nx), bool) if coarse_solid is None else np.asarray(coarse_solid, bool).reshape(nz, ny, nx)
        fsol = np.zeros((fz, fy, fx), bool) if fine_solid is None else np.asarray(fine_solid, bool).reshape(fz, fy, fx)

        # coarse depth inside the nest, measured from the interface faces only (-1 outside)
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        pos = (x, y, z)
        inside = np.ones(x.shape, dtype=bool)
        big = np.iinfo(np.int64).max
        depth = np.full(x.shape, big, dtype=np.int64)
        for a, name in enumerate("xyz"):
            inside &= (pos[a] >= self.lo[a]) & (pos[a] <= self.hi[a])
            if f"{name}-" not in self.wall_faces:
                depth = np.minimum(depth, pos[a] - self.lo[a])
            if f"{name}+" not in self.wall_faces:
                depth = np.minimum(depth, self.hi[a] - pos[a])
        depth = np.where(inside, depth, -1)
        self.coarse_state = np.where(depth >= 3, INACTIVE, 0).astype(np.uint8).ravel()
        self.coarse_depth = depth.ravel()
        cvalid = (depth <= 1) & ~csol  # coarse nodes that hold valid data when the interface is rebuilt
        fvalid = ~fsol

        def cidx(cx, cy, cz):
            return cx + nx * (cy + ny * cz)

        def fidx(mx, my, mz):
            return mx + fx * (my + fy * mz)

        # fine to coarse: depth-2 coarse nodes (fluid ones)
        fz_, fy_, fx_ = np.nonzero((depth == 2) & ~csol)
        self.fc_coarse = cidx(fx_, fy_, fz_).astype(np.int64)
        fine_base = np.stack([2 * (fx_ - self.lo[0]) - 1, 2 * (fy_ - self.lo[1]) - 1, 2 * (fz_ - self.lo[2]) - 1])
        fine_base, self.fc_offset = _move_off_solid(fine_base, fvalid, (fx, fy, fz))
        self.fc_fine = fidx(*fine_base).astype(np.int64)
        # coarse to fine: coarse cells of the outermost layer of every interface face
        cz_, cy_, cx_ = np.meshgrid(np.arange(self.lo[2], self.hi[2]), np.arange(self.lo[1], self.hi[1]),
                                    np.arange(self.lo[0], self.hi[0]), indexing="ij")
        shell = np.zeros(cx_.shape, dtype=bool)
        for a, (c, name) in enumerate(zip((cx_, cy_, cz_), "xyz")):
            if f"{name}-" not in self.wall_faces:
                shell |= c == self.lo[a]
            if f"{name}+" not in self.wall_faces:
                shell |= c == self.hi[a] - 1
        cx_, cy_, cz_ = cx_[shell], cy_[shell], cz_[shell]
        self.cf_fine = fidx(2 * (cx_ - self.lo[0]), 2 * (cy_ - self.lo[1]), 2 * (cz_ - self.lo[2])).astype(np.int64)
        coarse_base, self.cf_offset = _move_off_solid(np.stack([cx_, cy_, cz_]), cvalid, (nx, ny, nz))
        self.cf_coarse = cidx(*coarse_base).astype(np.int64)

    @property
    def n_coarse_active(self) -> int:
        return int((self.coarse_state == 0).sum())

    def fine_positions(self):
        """Coarse-unit coordinates (x, y, z) of every fine node (x fastest)."""
        fx, fy, fz = self.fine_dims
        mz, my, mx = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
        return [(self.lo[a] + 0.25 + 0.5 * m).ravel() for a, m in enumerate((mx, my, mz))]


def _cell_valid(base, valid, dims):
    """All 8 corners of the cells with corner MMM `base` (3 x n) are valid."""
    nx, ny, nz = dims
    ok = np.ones(base.shape[1], dtype=bool)
    for c in range(8):
        x = base[0] + (c & 1)
        y = base[1] + ((c >> 1) & 1)
        z = base[2] + ((c >> 2) & 1)
        inb = (x >= 0) & (x < nx) & (y >= 0) & (y < ny) & (z >= 0) & (z < nz)
        ok &= inb
        ok[inb] &= valid[z[inb], y[inb], x[inb]]
    return ok


def _move_off_solid(base, valid, dims):
    """Cells with an invalid corner move to the first all-valid neighbour cell
    in D3Q27 direction order (GridInterface::findOffsetCF/FC); returns the
    moved bases and the offsets (-c_d; 0 for cells that stay) as 3 arrays or
    None when no cell moved."""
    from .lattice import CX, CY, CZ

    base = base.copy()
    ok = _cell_valid(base, valid, dims)
    if ok.all():
        return base, None
    off = np.zeros((3, base.shape[1]))
    todo = np.nonzero(~ok)[0]
    for d in range(1, 27):
        if todo.size == 0:
            break
        cand = base[:, todo] + np.array([CX[d], CY[d], CZ[d]])[:, None]
        good = _cell_valid(cand, valid, dims)
        sel = todo[good]
        base[:, sel] = cand[:, good]
        off[:, sel] = -np.array([CX[d], CY[d], CZ[d]], dtype=float)[:, None]
        todo = todo[~good]
    if todo.size:
        raise ValueError(f"{todo.size} interface cells have no all-fluid neighbour cell")
    return base, [off[0], off[1], off[2]]


class ICells:
    """Device copy of one interface-cell list (vfx_icells)."""

    def __init__(self, coarse, fine, device, offsets=None):
        self.n = int(len(coarse))
        for arr, name in ((coarse, "coarse"), (fine, "fine")):
            if self.n and (int(np.min(arr)) < 0 or int(np.max(arr)) >= 2 ** 32):
                raise ValueError(f"{name} cell index out of range")
        self.coarse = torch.from_numpy(np.ascontiguousarray(coarse, dtype=np.int64).astype(np.uint32).view(np.int32)
                                       ).to(device)
        self.fine = torch.from_numpy(np.ascontiguousarray(fine, dtype=np.int64).astype(np.uint32).view(np.int32)
                                     ).to(device)
        self.max_coarse = int(np.max(coarse)) if self.n else -1
        self.max_fine = int(np.max(fine)) if self.n else -1
        # ICellNeigh x/y/z as given (float64); device copies in the lattices'
        # `real` (float64 for FP64 storage, float32 otherwise), made once per type
        self.device = device
        self._offsets_host = None if offsets is None else [np.ascontiguousarray(o, dtype=np.float64) for o in offsets]
        self._offsets = {}

    def offsets(self, real=torch.float64):
        if self._offsets_host is None:
            return None
        t = self._offsets.get(real)
        if t is None:
            t = [torch.from_numpy(o).to(self.device, real).contiguous() for o in self._offsets_host]
            self._offsets[real] = t
        return t

    def struct(self, real=torch.float64) -> V.VfxICells:
        s = V.VfxICells()
        s.fine_cells = self.fine.data_ptr()
        s.coarse_cells = self.coarse.data_ptr()
        s.n_cells = self.n
        o = self.offsets(real)
        if o is not None:
            s.offset_x, s.offset_y, s.offset_z = (a.data_ptr() for a in o)
        return s


def _check_cells(coarse: Lattice, fine: Lattice, cells: ICells):
    if cells.max_coarse >= coarse.n_nodes or cells.max_fine >= fine.n_nodes:
        raise ValueError("interface cell index beyond its lattice")


def scale_cf(coarse: Lattice, fine: Lattice, cells: ICells, coarse_even: bool, fine_even: bool = True, stream=None):
    _check_cells(coarse, fine, cells)
    c, f, i = coarse.struct(), fine.struct(), cells.struct(coarse._macro_dtype())
    V.check(V.lib().vfx_scale_cf(C.byref(c), C.byref(f), C.byref(i), int(bool(coarse_even)), int(bool(fine_even)),
                                 _stream(stream)), "vfx_scale_cf")


def scale_fc(coarse: Lattice, fine: Lattice, cells: ICells, coarse_even: bool, fine_even: bool = True, stream=None):
    _check_cells(coarse, fine, cells)
    c, f, i = coarse.struct(), fine.struct(), cells.struct(coarse._macro_dtype())
    V.check(V.lib().vfx_scale_fc(C.byref(c), C.byref(f), C.byref(i), int(bool(coarse_even)), int(bool(fine_even)),
                                 _stream(stream)), "vfx_scale_fc")


def scale_fc_cf(coarse: Lattice, fine: Lattice, fc: ICells, cf: ICells, coarse_even: bool, fine_even: bool = True,
                stream=None):
    """scale_fc then scale_cf in one launch (vfx_scale_fc_cf); needs
    interface_cells_apart(...) to hold."""
    _check_cells(coarse, fine, fc)
    _check_cells(coarse, fine, cf)
    c, f = coarse.struct(), fine.struct()
    real = coarse._macro_dtype()
    a, b = fc.struct(real), cf.struct(real)
    V.check(V.lib().vfx_scale_fc_cf(C.byref(c), C.byref(f), C.byref(a), C.byref(b), int(bool(coarse_even)),
                                    int(bool(fine_even)), _stream(stream)), "vfx_scale_fc_cf")


def _cell_corners(base, dims):
    """The 8 corner nodes (x fastest, periodic wrap) of the cells at `base`
    of a structured box, as the interface kernels take them."""
    nx, ny, nz = dims
    base = np.asarray(base, dtype=np.int64)
    x, y, z = base % nx, (base // nx) % ny, base // (nx * ny)
    out = []
    for c in range(8):
        xx = (x + (c & 1)) % nx
        yy = (y + ((c >> 1) & 1)) % ny
        zz = (z + ((c >> 2) & 1)) % nz
        out.append(xx + nx * (yy + ny * zz))
    return np.concatenate(out) if base.size else np.zeros(0, np.int64)


def interface_cells_apart(nest) -> bool:
    """True when one launch may do fine-to-coarse and coarse-to-fine: no coarse
    node the fine-to-coarse cells write is a corner of a coarse-to-fine cell,
    and no fine node the coarse-to-fine cells write is a corner of a
    fine-to-coarse cell."""
    fc_targets = np.asarray(nest.fc_coarse, dtype=np.int64)
    cf_sources = _cell_corners(nest.cf_coarse, nest.coarse_dims)
    fc_sources = _cell_corners(nest.fc_fine, nest.fine_dims)
    cf_targets = _cell_corners(nest.cf_fine, nest.fine_dims)
    return not (np.intersect1d(fc_targets, cf_sources).size or np.intersect1d(fc_sources, cf_targets).size)


def fine_omega(omega_coarse: float) -> float:
    """Relaxation rate of the 2:1 finer level for the same physical viscosity
    (nu_fine = 2 nu_coarse in fine lattice units; Parameter.cpp:404)."""
    nu_c = (1.0 / omega_coarse - 0.5) / 3.0
    return 1.0 / (3.0 * 2.0 * nu_c + 0.5)


class RefinedStepper:
    """One coarse step = two fine steps, one coarse step, fine-to-coarse, then
    coarse-to-fine, on a NestedBox (both levels direct-addressed boxes;
    inactive coarse nodes masked).  Each level's step is collide+stream, its
    post-collision boundary conditions, the parity flip, the AMD viscosity
    kernel (AMD model) and its pre-collision boundary conditions
    (UpdateGrid27.cpp:54-95).  A turbulence model applies to both levels with
    one SGS constant; the interface then uses per-node rates
    (hasTurbulentViscosity)."""

    def __init__(self, nest: NestedBox, omega_coarse, storage="fp64", collision="k17", device="cuda",
                 fp16_scale=None, variant="default", coarse_solid=None, fine_solid=None, coarse_bcs=(),
                 fine_bcs=(), reference_velocity=(0.0, 0.0, 0.0), turbulence="none", sgs_constant=0.0,
                 merged=False, fuse_outlet=False):
        self.nest = nest
        self.collision = collision
        self.omega_c = float(omega_coarse)
        self.omega_f = fine_omega(self.omega_c)
        cmask = nest.coarse_state != 0
        if coarse_solid is not None:
            cmask = cmask | np.asarray(coarse_solid, bool).ravel()
        turb = dict(turbulence=turbulence, sgs_constant=sgs_constant)
        self.coarse = Lattice(nest.coarse_dims, storage=storage, omega=self.omega_c, solid=cmask, device=device,
                              fp16_scale=fp16_scale, variant=variant, reference_velocity=reference_velocity, **turb)
        self.fine = Lattice(nest.fine_dims, storage=storage, omega=self.omega_f, device=device,
                            solid=None if fine_solid is None else np.asarray(fine_solid, bool).ravel(),
                            fp16_scale=fp16_scale, variant=variant, reference_velocity=reference_velocity, **turb)
        self.fc = ICells(nest.fc_coarse, nest.fc_fine, device, offsets=nest.fc_offset)
        self.cf = ICells(nest.cf_coarse, nest.cf_fine, device, offsets=nest.cf_offset)
        self.coarse_bcs = list(coarse_bcs)
        self.fine_bcs = list(fine_bcs)
        self.even_c = True
        self.even_f = True
        # device copies of the boundary and interface tables are made here, not
        # on first use: a first use may be inside a graph capture
        real = self.coarse._macro_dtype()
        for bc in self.coarse_bcs + self.fine_bcs:
            bc.tables(real)
        self.fc.offsets(real)
        self.cf.offsets(real)
        self.merged = bool(merged)
        if self.merged:
            self._setup_merged()
        self.fuse_outlet = bool(fuse_outlet)
        self._plan_c = None
        if self.fuse_outlet:
            if self.merged:
                raise ValueError("fuse_outlet runs with the launch-per-kernel schedule")
            if collision != "k17" or variant != "default" or turbulence != "none":
                raise ValueError("fuse_outlet runs the default K17 without turbulence model")
            # the coarse level's pressure group inside its collide-stream kernel,
            # taken at the next step's gather (the interface kernels between two
            # coarse steps touch no node of the outlet or its neighbours)
            self._plan_c = FusedPlan.build(self.coarse, self.coarse_bcs, post=False)
            if self._plan_c.pre is None:
                raise ValueError("fuse_outlet: the coarse level has no pressure group the kernel can take")

    def _setup_merged(self):
        """Fewer launches per coarse step (same values): the coarse step and the
        first fine step in one launch (step_levels), each phase's boundary
        groups of both levels in one launch (BCSet), fine-to-coarse and
        coarse-to-fine in one launch (scale_fc_cf) where the nest keeps them
        apart.  Needs the default K17 without turbulence model."""
        if self.collision != "k17" or self.coarse.variant != "default" or self.coarse.turbulence != "none":
            raise ValueError("the merged schedule runs the default K17 without turbulence model")

        def split(bcs, pre):
            inset = [b for b in bcs if b.pre_collision == pre and b.kind in BCSet.KINDS]
            rest = [b for b in bcs if b.pre_collision == pre and b.kind not in BCSet.KINDS]
            return inset, rest

        cpost, self._c_post_rest = split(self.coarse_bcs, False)
        fpost, self._f_post_rest = split(self.fine_bcs, False)
        cpre, self._c_pre_rest = split(self.coarse_bcs, True)
        fpre, self._f_pre_rest = split(self.fine_bcs, True)
        self._n_cpost, self._n_cpre = len(cpost), len(cpre)
        self._set_post_both = BCSet([(self.coarse, b) for b in cpost] + [(self.fine, b) for b in fpost])
        self._set_post_fine = BCSet([(self.fine, b) for b in fpost])
        self._set_pre_both = BCSet([(self.coarse, b) for b in cpre] + [(self.fine, b) for b in fpre])
        self._set_pre_fine = BCSet([(self.fine, b) for b in fpre])
        self._n_fpost, self._n_fpre = len(fpost), len(fpre)
        self._one_interface = interface_cells_apart(self.nest)

    def _step_merged(self, stream):
        c, f = self.coarse, self.fine
        ec, ef = self.even_c, self.even_f
        # coarse step and first fine step
        step_levels([(c, ec), (f, ef)], stream)
        self._set_post_both.apply([ec] * self._n_cpost + [ef] * self._n_fpost, stream)
        for b in self._c_post_rest:
            c.apply_bc(b, ec, stream=stream)
        for b in self._f_post_rest:
            f.apply_bc(b, ef, stream=stream)
        ec, ef = not ec, not ef
        self._set_pre_both.apply([ec] * self._n_cpre + [ef] * self._n_fpre, stream)
        for b in self._c_pre_rest:
            c.apply_bc(b, ec, stream=stream)
        for b in self._f_pre_rest:
            f.apply_bc(b, ef, stream=stream)
        # second fine step
        step_levels([(f, ef)], stream)
        self._set_post_fine.apply([ef] * self._n_fpost, stream)
        for b in self._f_post_rest:
            f.apply_bc(b, ef, stream=stream)
        ef = not ef
        self._set_pre_fine.apply([ef] * self._n_fpre, stream)
        for b in self._f_pre_rest:
            f.apply_bc(b, ef, stream=stream)
        self.even_c, self.even_f = ec, ef
        if self._one_interface:
            scale_fc_cf(c, f, self.fc, self.cf, ec, ef, stream)
        else:
            scale_fc(c, f, self.fc, ec, ef, stream)
            scale_cf(c, f, self.cf, ec, ef, stream)

    @staticmethod
    def _level_step(lat, bcs, collision, even, stream):
        lat.collide_stream(collision, even, stream=stream)
        for bc in bcs:
            if not bc.pre_collision:
                lat.apply_bc(bc, even, stream=stream)
        if collision == "k17" and lat.turbulence == "amd":
            lat.turbulent_viscosity_amd(stream=stream)
        for bc in bcs:
            if bc.pre_collision:
                lat.apply_bc(bc, not even, stream=stream)
        return not even

    def _coarse_step_fused(self, even, first, stream):
        plan, lat = self._plan_c, self.coarse
        lat.k17_step_fused(even, [], plan.pre, apply_pre=not first, any_mask=plan.any_mask, stream=stream)
        for bc in plan.residual_post:
            lat.apply_bc(bc, even, stream=stream)
        for bc in plan.residual_pre:
            lat.apply_bc(bc, not even, stream=stream)
        return not even

    def step(self, n=1, stream=None):
        if self.merged:
            for _ in range(n):
                self._step_merged(stream)
            return
        for i in range(n):
            for _ in range(2):
                self.even_f = self._level_step(self.fine, self.fine_bcs, self.collision, self.even_f, stream)
            if self._plan_c is not None:
                self.even_c = self._coarse_step_fused(self.even_c, i == 0, stream)
            else:
                self.even_c = self._level_step(self.coarse, self.coarse_bcs, self.collision, self.even_c, stream)
            scale_fc(self.coarse, self.fine, self.fc, self.even_c, self.even_f, stream)
            scale_cf(self.coarse, self.fine, self.cf, self.even_c, self.even_f, stream)
        if n > 0 and self._plan_c is not None:
            # the outlet once more, so the stored state is the separate kernels'
            self.coarse.apply_bc(self._plan_c.pre.bc, self.even_c, stream=stream)

    def capture(self, n=1) -> Graph:
        """n coarse steps from the current parities as one HIP graph."""
        saved = (self.even_c, self.even_f)
        try:
            g = capture(lambda s: self.step(n, stream=s))
        finally:
            self.even_c, self.even_f = saved
        g.parity = saved
        return g

    def replay(self, graph: Graph, n, stream=None):
        """Launch a graph of n coarse steps captured at the current parities
        (an odd n therefore alternates between two graphs)."""
        if graph.parity is not None and graph.parity != (self.even_c, self.even_f):
            raise ValueError("graph was recorded from other time-step parities")
        graph.launch(stream)
        if n % 2:
            self.even_c = not self.even_c

    def autotune(self, steps=4):
        """Launch shape of each level's structured step (Lattice.autotune),
        with the populations restored afterwards.  The two-level launch of the
        merged schedule keeps 256-node tiles; one-level launches and the
        separate schedule use the tuned shapes.  Returns {level: (tile_nodes,
        xcd_run, ms)}."""
        out = {}
        for name, lat, even in (("coarse", self.coarse, self.even_c), ("fine", self.fine, self.even_f)):
            saved = lat.f.clone()
            out[name] = lat.autotune(even, steps=steps)
            lat.f.copy_(saved)
            del saved
        return out

    @property
    def launches_per_step(self) -> int:
        """Kernel launches of one coarse step in the current schedule."""
        if self._plan_c is not None:
            p = self._plan_c
            return 2 * (1 + len(self.fine_bcs)) + 1 + len(p.residual_post) + len(p.residual_pre) + 2
        if self.merged:
            sets = (self._set_post_both, self._set_pre_both, self._set_post_fine, self._set_pre_fine)
            rest = 2 * (len(self._f_post_rest) + len(self._f_pre_rest)) + len(self._c_post_rest) + len(self._c_pre_rest)
            return 2 + sum(s.handle is not None for s in sets) + rest + (1 if self._one_interface else 2)
        turb = int(self.collision == "k17" and self.coarse.turbulence == "amd")
        return 2 * (1 + turb + len(self.fine_bcs)) + (1 + turb + len(self.coarse_bcs)) + 2

    @property
    def lups_per_step(self) -> int:
        """Lattice updates per coarse step, fine counted twice (two fine steps)."""
        return self.coarse.n_fluid + 2 * self.fine.n_fluid
