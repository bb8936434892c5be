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
from .lattice import Lattice, _stream

INACTIVE = 2  # value in the coarse node-state array


class NestedBox:
    """Index tables of one fine nest inside a coarse box (x fastest)."""

    def __init__(self, coarse_dims, lo, hi):
        self.coarse_dims = tuple(int(v) for v in coarse_dims)
        self.lo = np.array(lo, dtype=np.int64)
        self.hi = np.array(hi, dtype=np.int64)
        if np.any(self.hi - self.lo < 6):
            raise ValueError("the nest needs at least 6 coarse cells per axis")
        if np.any(self.lo < 1) or np.any(self.hi > np.array(self.coarse_dims) - 2):
            raise ValueError("the nest must stay one coarse node clear of the box faces")
        self.fine_dims = tuple(int(v) for v in 2 * (self.hi - self.lo))
        nx, ny, nz = self.coarse_dims
        fx, fy, fz = self.fine_dims

        # coarse depth inside the nest (-1 outside)
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        pos = (x, y, z)
        inside = np.ones(x.shape, dtype=bool)
        depth = np.full(x.shape, np.iinfo(np.int64).max, dtype=np.int64)
        for a in range(3):
            inside &= (pos[a] >= self.lo[a]) & (pos[a] <= self.hi[a])
            depth = np.minimum(depth, np.minimum(pos[a] - self.lo[a], self.hi[a] - pos[a]))
        depth = np.where(inside, depth, -1)
        self.coarse_state = np.where(depth >= 3, INACTIVE, 0).astype(np.uint8).ravel()
        self.coarse_depth = depth.ravel()

        def cidx(cx, cy, cz):
            return cx + nx * (cy + ny * cz)

        def fidx(mx, my, mz):
            return mx + fx * (my + fy * mz)

        # fine to coarse: depth-2 coarse nodes
        fz_, fy_, fx_ = np.nonzero(depth == 2)
        self.fc_coarse = cidx(fx_, fy_, fz_).astype(np.int64)
        self.fc_fine = fidx(2 * (fx_ - self.lo[0]) - 1, 2 * (fy_ - self.lo[1]) - 1,
                            2 * (fz_ - self.lo[2]) - 1).astype(np.int64)
        # coarse to fine: coarse cells of the outermost nest layer
        cz_, cy_, cx_ = np.meshgrid(np.arange(self.lo[2], self.hi[2]), np.arange(self.lo[1], self.hi[1]),
                                    np.arange(self.lo[0], self.hi[0]), indexing="ij")
        shell = np.zeros(cx_.shape, dtype=bool)
        for a, c in enumerate((cx_, cy_, cz_)):
            shell |= (c == self.lo[a]) | (c == self.hi[a] - 1)
        cx_, cy_, cz_ = cx_[shell], cy_[shell], cz_[shell]
        self.cf_coarse = cidx(cx_, cy_, cz_).astype(np.int64)
        self.cf_fine = fidx(2 * (cx_ - self.lo[0]), 2 * (cy_ - self.lo[1]), 2 * (cz_ - self.lo[2])).astype(np.int64)

    @property
    def n_coarse_active(self) -> int:
        return int((self.coarse_state == 0).sum())

    def fine_positions(self):
        """Coarse-unit coordinates (x, y, z) of every fine node (x fastest)."""
        fx, fy, fz = self.fine_dims
        mz, my, mx = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
        return [(self.lo[a] + 0.25 + 0.5 * m).ravel() for a, m in enumerate((mx, my, mz))]


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
        self.offsets = None
        if offsets is not None:
            self.offsets = [torch.as_tensor(np.ascontiguousarray(o, dtype=np.float64)).to(device) for o in offsets]

    def struct(self) -> V.VfxICells:
        s = V.VfxICells()
        s.fine_cells = self.fine.data_ptr()
        s.coarse_cells = self.coarse.data_ptr()
        s.n_cells = self.n
        if self.offsets is not None:
            s.offset_x, s.offset_y, s.offset_z = (o.data_ptr() for o in self.offsets)
        return s


def _check_cells(coarse: Lattice, fine: Lattice, cells: ICells):
    if cells.max_coarse >= coarse.n_nodes or cells.max_fine >= fine.n_nodes:
        raise ValueError("interface cell index beyond its lattice")


def scale_cf(coarse: Lattice, fine: Lattice, cells: ICells, coarse_even: bool, fine_even: bool = True, stream=None):
    _check_cells(coarse, fine, cells)
    c, f, i = coarse.struct(), fine.struct(), cells.struct()
    V.check(V.lib().vfx_scale_cf(C.byref(c), C.byref(f), C.byref(i), int(bool(coarse_even)), int(bool(fine_even)),
                                 _stream(stream)), "vfx_scale_cf")


def scale_fc(coarse: Lattice, fine: Lattice, cells: ICells, coarse_even: bool, fine_even: bool = True, stream=None):
    _check_cells(coarse, fine, cells)
    c, f, i = coarse.struct(), fine.struct(), cells.struct()
    V.check(V.lib().vfx_scale_fc(C.byref(c), C.byref(f), C.byref(i), int(bool(coarse_even)), int(bool(fine_even)),
                                 _stream(stream)), "vfx_scale_fc")


def fine_omega(omega_coarse: float) -> float:
    """Relaxation rate of the 2:1 finer level for the same physical viscosity
    (nu_fine = 2 nu_coarse in fine lattice units; Parameter.cpp:404)."""
    nu_c = (1.0 / omega_coarse - 0.5) / 3.0
    return 1.0 / (3.0 * 2.0 * nu_c + 0.5)


class RefinedStepper:
    """One coarse step = two fine steps, one coarse step, fine-to-coarse, then
    coarse-to-fine, on a NestedBox (both levels direct-addressed periodic
    boxes; inactive coarse nodes masked)."""

    def __init__(self, nest: NestedBox, omega_coarse, storage="fp64", collision="k17", device="cuda",
                 fp16_scale=None, variant="default"):
        self.nest = nest
        self.collision = collision
        self.omega_c = float(omega_coarse)
        self.omega_f = fine_omega(self.omega_c)
        self.coarse = Lattice(nest.coarse_dims, storage=storage, omega=self.omega_c,
                              solid=nest.coarse_state != 0, device=device, fp16_scale=fp16_scale, variant=variant)
        self.fine = Lattice(nest.fine_dims, storage=storage, omega=self.omega_f, device=device,
                            fp16_scale=fp16_scale, variant=variant)
        self.fc = ICells(nest.fc_coarse, nest.fc_fine, device)
        self.cf = ICells(nest.cf_coarse, nest.cf_fine, device)
        self.even_c = True
        self.even_f = True

    def step(self, n=1, stream=None):
        for _ in range(n):
            for _ in range(2):
                self.fine.collide_stream(self.collision, self.even_f, stream=stream)
                self.even_f = not self.even_f
            self.coarse.collide_stream(self.collision, self.even_c, stream=stream)
            self.even_c = not self.even_c
            scale_fc(self.coarse, self.fine, self.fc, self.even_c, self.even_f, stream)
            scale_cf(self.coarse, self.fine, self.cf, self.even_c, self.even_f, stream)

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

    @property
    def lups_per_step(self) -> int:
        """Lattice updates per coarse step, fine counted twice (two fine steps)."""
        return self.nest.n_coarse_active + 2 * int(np.prod(self.nest.fine_dims))
