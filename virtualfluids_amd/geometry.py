"""Structured-box geometry for the stepper: solid masks and boundary-condition
tables built with whole-array operations (one pass per lattice direction), so
grids of 10^8 nodes are prepared in seconds.

* ``spheres_in_channel``: the flow-past-spheres channel — velocity inlet,
  non-equilibrium pressure outlet, no-slip side walls, spheres with per-link
  subgrid distances from the ray-sphere intersection of VirtualFluids'
  GridGenerator (src/gpu/GridGenerator/geometries/Sphere/Sphere.cpp:124-170:
  q is the parameter t of node + t c_d on the sphere surface).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import CX, CY, CZ

D_P00 = 1  # direction +x (outlet normal)


@dataclass
class BCTable:
    """One boundary-condition group in QforBoundaryConditions form: nodes,
    q27 (27 x n, q < 0 = no wall link), per-node velocity, density and inward
    neighbours (pressure boundary)."""
    kind: int
    nodes: np.ndarray
    q27: np.ndarray | None = None
    velocity: np.ndarray | None = None   # (3, n)
    rho: np.ndarray | None = None
    neighbors: np.ndarray | None = None
    normal: int = 0


@dataclass
class Channel:
    dims: tuple
    solid: np.ndarray              # bool, x fastest
    spheres: np.ndarray            # (n, 4): cx, cy, cz, r
    bcs: list = field(default_factory=list)
    omega: float = 0.0
    u_in: float = 0.05

    @property
    def n_fluid(self) -> int:
        return int((~self.solid).sum())


def place_spheres(dims, n=64, r_range=(6.0, 14.0), seed=1234, margin=2.0, x_clear=(48, 96), max_tries=100000):
    """Centres uniform in the channel (clear of the walls by `margin`, of the
    inlet and outlet by x_clear), radii uniform in r_range, a candidate
    rejected when it overlaps (gap < margin) a sphere already placed."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place the spheres without overlap")
        r = rng.uniform(*r_range)
        c = np.array([rng.uniform(x_clear[0] + r, nx - 1 - x_clear[1] - r),
                      rng.uniform(1 + margin + r, ny - 2 - margin - r),
                      rng.uniform(1 + margin + r, nz - 2 - margin - r)])
        if all(np.linalg.norm(c - s[:3]) >= r + s[3] + margin for s in out):
            out.append(np.array([c[0], c[1], c[2], r]))
    return np.array(out)


def _shift(a, d):
    """b[z, y, x] = a[z + cz, y + cy, x + cx] (periodic wrap; the box faces are
    solid, so wrapped values are never used for fluid nodes)."""
    return np.roll(a, shift=(-CZ[d], -CY[d], -CX[d]), axis=(0, 1, 2))


def spheres_in_channel(dims=(512, 256, 256), n_spheres=64, seed=1234, u_in=0.05, noise=0.01, noise_seed=1235,
                       re=200.0, r_range=(6.0, 14.0), x_clear=(48, 96)):
    """The channel: solid layers at x = 0 (inlet plane), x = nx-1 (outlet
    plane), y = 0 / ny-1, z = 0 / nz-1; velocity inlet (interpolated, q = 1/2)
    with u = (u_in, 0, 0) plus a uniform perturbation of +-noise*u_in per
    component and node; pressure outlet (drho = 0) on the last fluid layer;
    no-slip walls (q = 1/2) and spheres (q from the ray-sphere intersection).
    Viscosity from Re on the realised mean sphere diameter."""
    from ._lib import VFX_BC_NOSLIP_INTERPOLATED, VFX_BC_PRESSURE_NEQ, VFX_BC_VELOCITY_INTERPOLATED

    nx, ny, nz = dims
    sph = place_spheres(dims, n_spheres, r_range, seed, x_clear=x_clear)
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij", sparse=True)
    label = np.full((nz, ny, nx), -1, dtype=np.int16)
    for i, (cx, cy, cz, r) in enumerate(sph):
        x0, x1 = int(np.floor(cx - r)) - 1, int(np.ceil(cx + r)) + 2
        y0, y1 = int(np.floor(cy - r)) - 1, int(np.ceil(cy + r)) + 2
        z0, z1 = int(np.floor(cz - r)) - 1, int(np.ceil(cz + r)) + 2
        sub = ((x[:, :, x0:x1] - cx) ** 2 + (y[:, y0:y1, :] - cy) ** 2 + (z[z0:z1] - cz) ** 2) < r * r
        label[z0:z1, y0:y1, x0:x1][sub] = i
    wall = np.zeros((nz, ny, nx), dtype=bool)
    wall[:, :, 0] = wall[:, :, nx - 1] = True
    wall[:, 0, :] = wall[:, ny - 1, :] = True
    wall[0, :, :] = wall[nz - 1, :, :] = True
    solid = wall | (label >= 0)
    fluid = ~solid

    # per-node q27 of the inlet and no-slip groups (27 x N, only touched rows kept)
    q_in = {}
    q_ns = {}
    inlet_face = np.zeros((nz, ny, nx), dtype=bool)
    inlet_face[1:-1, 1:-1, 0] = True
    outlet_face = np.zeros((nz, ny, nx), dtype=bool)
    outlet_face[:, :, nx - 1] = True
    for d in range(1, 27):
        tgt_solid = _shift(solid, d) & fluid
        if not tgt_solid.any():
            continue
        t_inlet = _shift(inlet_face, d) & tgt_solid
        t_outlet = _shift(outlet_face, d) & tgt_solid
        t_label = _shift(label, d)
        t_sphere = (t_label >= 0) & tgt_solid
        t_wall = tgt_solid & ~t_inlet & ~t_outlet & ~t_sphere
        if t_inlet.any():
            q_in[d] = (np.flatnonzero(t_inlet), np.full(int(t_inlet.sum()), 0.5, np.float32))
        idx_w = np.flatnonzero(t_wall)
        idx_s = np.flatnonzero(t_sphere)
        qs = np.empty(0, np.float32)
        if idx_s.size:
            lab = t_label.ravel()[idx_s]
            px = (idx_s % nx).astype(np.float64)
            py = ((idx_s // nx) % ny).astype(np.float64)
            pz = (idx_s // (nx * ny)).astype(np.float64)
            c = sph[lab]
            rel = np.stack([px - c[:, 0], py - c[:, 1], pz - c[:, 2]])
            dv = np.array([CX[d], CY[d], CZ[d]], dtype=np.float64)[:, None]
            dd = float((dv * dv).sum())
            p = 2.0 * (rel * dv).sum(axis=0) / dd
            qq = ((rel * rel).sum(axis=0) - c[:, 3] ** 2) / dd
            disc = np.maximum(0.25 * p * p - qq, 0.0)
            r1 = -0.5 * p + np.sqrt(disc)
            r2 = -0.5 * p - np.sqrt(disc)
            r1 = np.where(r1 < 0, np.inf, r1)
            r2 = np.where(r2 < 0, np.inf, r2)
            t = np.minimum(r1, r2)
            # the node is outside and the link end inside, so 0 < t <= 1;
            # clip guards rounding at the ends
            qs = np.clip(t, 1e-6, 1.0).astype(np.float32)
        q_ns[d] = (np.concatenate([idx_w, idx_s]),
                   np.concatenate([np.full(idx_w.size, 0.5, np.float32), qs]))

    def table(per_dir):
        nodes = np.unique(np.concatenate([v[0] for v in per_dir.values()])) if per_dir else np.empty(0, np.int64)
        q27 = np.full((27, nodes.size), -1.0, np.float32)
        for d, (idx, q) in per_dir.items():
            q27[d, np.searchsorted(nodes, idx)] = q
        return nodes.astype(np.int64), q27

    bcs = []
    in_nodes, in_q = table(q_in)
    rng = np.random.default_rng(noise_seed)
    vel = np.zeros((3, in_nodes.size))
    vel[0] = u_in
    vel += noise * u_in * rng.uniform(-1.0, 1.0, vel.shape)
    bcs.append(BCTable(VFX_BC_VELOCITY_INTERPOLATED, in_nodes, in_q, velocity=vel))
    ns_nodes, ns_q = table(q_ns)
    bcs.append(BCTable(VFX_BC_NOSLIP_INTERPOLATED, ns_nodes, ns_q))
    zz, yy = np.meshgrid(np.arange(1, nz - 1), np.arange(1, ny - 1), indexing="ij")
    out_nodes = ((nx - 2) + nx * (yy + ny * zz)).ravel().astype(np.int64)
    out_nodes = out_nodes[fluid.ravel()[out_nodes]]
    bcs.append(BCTable(VFX_BC_PRESSURE_NEQ, out_nodes, rho=np.zeros(out_nodes.size), neighbors=out_nodes - 1,
                       normal=D_P00))
    d_mean = 2.0 * float(sph[:, 3].mean())
    nu = u_in * d_mean / re
    omega = 1.0 / (3.0 * nu + 0.5)
    return Channel(dims=tuple(dims), solid=solid.ravel(), spheres=sph, bcs=bcs, omega=omega, u_in=u_in)


def boundary_conditions(tables, device="cuda"):
    """Device BoundaryCondition objects for BCTable groups."""
    from .lattice import BoundaryCondition

    out = []
    for b in tables:
        vel = b.velocity if b.velocity is not None else (None, None, None)
        out.append(BoundaryCondition(b.kind, b.nodes, b.q27, vel[0], vel[1], vel[2], rho=b.rho,
                                     neighbors=b.neighbors, normal=b.normal, device=device))
    return out


def taylor_green(n, u0=0.04, re=1600.0):
    """Periodic Taylor-Green vortex on n^3 nodes (apps/gpu/TaylorGreenVortex/
    TaylorGreenVortex.cpp:97-100, 138-148): nodes at -pi + i 2pi/n,
    u = u0 (sin x cos y cos z, -cos x sin y cos z, 0),
    drho = 3 (u0^2/16)(cos 2x + cos 2y)(cos 2z + 2); viscosity from Re on
    L = n / 2pi.  Returns (drho, ux, uy, uz) node arrays (x fastest), omega, nu."""
    c = -np.pi + 2.0 * np.pi * np.arange(n) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    drho = 3.0 * (u0 * u0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    ux = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    uy = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    uz = np.zeros_like(ux)
    nu = (n / (2.0 * np.pi)) * u0 / re
    omega = 1.0 / (3.0 * nu + 0.5)
    return [a.ravel() for a in (drho, ux, uy, uz)], omega, nu


@dataclass
class RefinedCase:
    """A two-level setup: nest tables, per-level solids and boundary tables."""
    nest: object
    coarse_solid: np.ndarray
    fine_solid: np.ndarray
    coarse_bcs: list
    fine_bcs: list
    omega: float
    u_ref: float
    cube: tuple


def _links(solid, dims, target_mask, q=0.5):
    """Fluid nodes with links into `target_mask` nodes (periodic wrap):
    (nodes, q27) with q on those links and -1 elsewhere."""
    nx, ny, nz = dims
    fluid = ~solid
    per = {}
    for d in range(1, 27):
        hit = _shift(target_mask, d) & fluid
        if hit.any():
            per[d] = np.flatnonzero(hit)
    if not per:
        return np.empty(0, np.int64), np.full((27, 0), -1.0, np.float32)
    nodes = np.unique(np.concatenate(list(per.values())))
    q27 = np.full((27, nodes.size), -1.0, np.float32)
    for d, idx in per.items():
        q27[d, np.searchsorted(nodes, idx)] = q
    return nodes.astype(np.int64), q27


def wall_mounted_cube(coarse_dims=(256, 128, 128), nest_cells=(96, 64, 48), cube=32, cube_x=64, upstream=24,
                      u_ref=0.05, re=1.0e4, noise=0.02, noise_seed=2024):
    """Wall-mounted cube in a boundary layer, coarse level + one 2:1 nest.

    Coarse box (coarse units, x streamwise, z up): inlet plane x = 0 with the
    power-law profile u(z) = u_ref (z'/H)^0.2 (z' = height above the floor
    wall at z = 1/2, H = cube height) plus a uniform +-noise*u(z) perturbation
    of every component (seeded); pressure outlet on the last fluid layer before
    x = nx-1; no-slip floor (solid z = 0, halfway links); slip top (solid
    z = nz-1); periodic in y.  The cube occupies coarse cells
    [cube_x, cube_x+cube) x [centred in y) x [0, cube) above the floor; the
    nest covers coarse cells x in [cube_x-upstream, +nest_cells[0]), centred in
    y, z in [0, nest_cells[2]) and stands on the floor (no interface there).
    The cube is resolved on the fine level only (it lies in the inactive core
    of the coarse level).  Re on the cube height at u_ref."""
    from ._lib import (VFX_BC_NOSLIP_INTERPOLATED, VFX_BC_PRESSURE_NEQ, VFX_BC_SLIP, VFX_BC_VELOCITY_INTERPOLATED)
    from .refinement import NestedBox

    nx, ny, nz = coarse_dims
    H = float(cube)
    cy0 = (ny - cube) // 2
    lo = (cube_x - upstream, (ny - nest_cells[1]) // 2, 0)
    hi = (lo[0] + nest_cells[0], lo[1] + nest_cells[1], nest_cells[2])
    # coarse walls
    csol = np.zeros((nz, ny, nx), bool)
    inlet = np.zeros_like(csol)
    outlet = np.zeros_like(csol)
    floor = np.zeros_like(csol)
    top = np.zeros_like(csol)
    inlet[1:-1, :, 0] = True
    outlet[1:-1, :, nx - 1] = True
    floor[0, :, :] = True
    top[nz - 1, :, :] = True
    csol = inlet | outlet | floor | top
    nest = NestedBox(coarse_dims, lo, hi, wall_faces=("z-",), coarse_solid=None)
    fx, fy, fz = nest.fine_dims
    # fine solids: floor layer (z = 1/4 < 1/2) and the cube [x0 - 1/2, x0 + cube - 1/2] (coarse units)
    px, py, pz = [p.reshape(fz, fy, fx) for p in nest.fine_positions()]
    fcube = ((px > cube_x - 0.5) & (px < cube_x + cube - 0.5) & (py > cy0 - 0.5) & (py < cy0 + cube - 0.5) &
             (pz < cube + 0.5))
    ffloor = pz < 0.5
    fsol = fcube | ffloor
    nest = NestedBox(coarse_dims, lo, hi, wall_faces=("z-",), coarse_solid=csol, fine_solid=fsol)

    # coarse boundary tables
    cfl = ~csol
    zc = np.arange(nz, dtype=np.float64)
    prof = u_ref * np.clip((zc - 0.5) / H, 0.0, None) ** 0.2
    in_nodes, in_q = _links(csol, coarse_dims, inlet)
    rng = np.random.default_rng(noise_seed)
    zin = in_nodes // (nx * ny)
    vel = np.zeros((3, in_nodes.size))
    vel[0] = prof[zin]
    vel += noise * prof[zin] * rng.uniform(-1.0, 1.0, vel.shape)
    fl_nodes, fl_q = _links(csol, coarse_dims, floor)
    tp_nodes, tp_q = _links(csol, coarse_dims, top)
    zz, yy = np.meshgrid(np.arange(1, nz - 1), np.arange(ny), indexing="ij")
    out_nodes = ((nx - 2) + nx * (yy + ny * zz)).ravel().astype(np.int64)
    coarse_bcs = [
        BCTable(VFX_BC_VELOCITY_INTERPOLATED, in_nodes, in_q, velocity=vel),
        BCTable(VFX_BC_NOSLIP_INTERPOLATED, fl_nodes, fl_q),
        BCTable(VFX_BC_SLIP, tp_nodes, tp_q),
        BCTable(VFX_BC_PRESSURE_NEQ, out_nodes, rho=np.zeros(out_nodes.size), neighbors=out_nodes - 1, normal=D_P00),
    ]
    # links into the outlet plane are dropped (pressure boundary); links into
    # the inlet/outlet plane edges belong to the floor/top groups above
    # fine boundary tables: floor and cube, halfway
    ff_nodes, ff_q = _links(fsol, nest.fine_dims, ffloor)
    fc_nodes, fc_q = _links(fsol, nest.fine_dims, fcube)
    fine_bcs = [BCTable(VFX_BC_NOSLIP_INTERPOLATED, ff_nodes, ff_q),
                BCTable(VFX_BC_NOSLIP_INTERPOLATED, fc_nodes, fc_q)]
    nu = u_ref * H / re
    omega = 1.0 / (3.0 * nu + 0.5)
    return RefinedCase(nest=nest, coarse_solid=csol.ravel(), fine_solid=fsol.ravel(), coarse_bcs=coarse_bcs,
                       fine_bcs=fine_bcs, omega=omega, u_ref=u_ref, cube=(cube_x, cy0, 0, cube))
