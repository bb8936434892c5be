"""The oracle's refinement interpolation (coarse-to-fine / fine-to-coarse)
pinned without a GPU:

* golden vectors of the compiled reference header (g4: MomentsOnSourceNodeSet,
  interpolateCF, interpolateFC on 128 random cells, half of them with offsets);
* closed forms: a uniform state maps to the product-form (factorised)
  equilibrium of that state at every destination node; a linear velocity and
  density field is reproduced exactly at the destination positions and the
  destination's non-equilibrium strain is eps times the source strain;
* mirror symmetry in x, y and z;
* the nest tables: prolongation/restriction over a whole nest reproduce a
  linear field at the fine-node / coarse-node positions;
* two-level physics: a shear wave crossing the nest decays at the analytic
  viscous rate and the fine level follows the same wave.
"""
import copy
import os

import numpy as np
import pytest

import oracle
import virtualfluids_amd as vfx
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
W = cases.W
CX, CY, CZ = cases.CX, cases.CY, cases.CZ


def g4():
    return np.load(os.path.join(HERE, "golden", "g4_interp.npz"))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def test_golden_g4_coefficients_and_outputs():
    g = g4()
    worst = {"coeff_cf": 0.0, "coeff_fc": 0.0, "cf": 0.0, "fc": 0.0}
    for i in range(len(g["f8"])):
        wc, wf = g["omega"][i]
        k, out = oracle.interp_cell("cf", g["f8"][i], wc, wf, g["offsets"][i])
        worst["coeff_cf"] = max(worst["coeff_cf"], rel(k, g["coeff_cf"][i]))
        worst["cf"] = max(worst["cf"], rel(out, g["cf"][i]))
        k, out = oracle.interp_cell("fc", g["f8"][i], wf, wc, g["offsets"][i])
        worst["coeff_fc"] = max(worst["coeff_fc"], rel(k, g["coeff_fc"][i]))
        worst["fc"] = max(worst["fc"], rel(out, g["fc"][i]))
    print(worst)
    assert max(worst.values()) <= 1e-13


@pytest.mark.skipif(not oracle.ref_available(), reason="reference closure not built")
def test_live_reference_cells():
    rng = np.random.default_rng(99)
    for t in range(40):
        u = rng.uniform(-0.05, 0.05, (8, 3))
        f8 = np.stack([oracle.equilibrium(rng.uniform(-0.01, 0.01), *u[c]) for c in range(8)])
        f8 += rng.normal(0, 2e-3, (8, 27)) * W
        off = (0.0, 0.0, 0.0) if t % 3 else tuple(rng.choice([-0.5, 0.5], 3))
        for kind in ("cf", "fc"):
            k1, o1 = oracle.interp_cell(kind, f8, 1.93, 1.87, off)
            k2, o2 = oracle.interp_cell(kind, f8, 1.93, 1.87, off, impl="ref")
            assert rel(k1, k2) <= 1e-13 and rel(o1, o2) <= 1e-13


def product_equilibrium(drho, u):
    """Factorised D3Q27 equilibrium (shifted): rho * p(cx,ux) p(cy,uy) p(cz,uz) - w
    with the D1Q3 factors p(0,u) = 2/3 - u^2, p(+-1,u) = (1/3 + u^2 +- u) / 2."""
    def p(c, v):
        return 2.0 / 3.0 - v * v if c == 0 else 0.5 * (1.0 / 3.0 + v * v + c * v)
    rho = 1.0 + drho
    return np.array([rho * p(CX[d], u[0]) * p(CY[d], u[1]) * p(CZ[d], u[2]) - W[d] for d in range(27)])


def test_uniform_state_gives_product_equilibrium():
    drho, u = 0.004, np.array([0.03, -0.02, 0.045])
    f8 = np.stack([oracle.equilibrium(drho, *u)] * 8)
    expect = product_equilibrium(drho, u)
    _, cf = oracle.interp_cell("cf", f8, 1.9, 1.8)
    _, fc = oracle.interp_cell("fc", f8, 1.8, 1.9)
    for c in range(8):
        assert np.allclose(cf[c], expect, rtol=0, atol=1e-16)
    assert np.allclose(fc, expect, rtol=0, atol=1e-16)


def _linear_cell(rng, eq_noise=0.0, linear_density=True):
    u0 = rng.uniform(-0.04, 0.04, 3)
    G = rng.uniform(-0.01, 0.01, (3, 3))
    d0 = rng.uniform(-0.005, 0.005)
    g = rng.uniform(-0.002, 0.002, 3) if linear_density else np.zeros(3)
    f8 = []
    for c in range(8):
        s = 0.5 * np.array([1 if c & 1 else -1, 1 if c & 2 else -1, 1 if c & 4 else -1])
        f8.append(oracle.equilibrium(d0 + g @ s, *(u0 + G @ s)))
    f8 = np.array(f8)
    # the same traceless stress on every corner (no mass, no momentum); with a
    # uniform density the source strains k.. are then equal on all corners and
    # the field stays linear
    a, b, c = rng.normal(0, 1, 3)
    neq = W * (a * CX * CY + b * (CX ** 2 - CY ** 2) + c * CY * CZ)
    f8 += eq_noise * np.tile(neq, (8, 1))
    return u0, G, d0, g, f8


def _laplace(G):
    return -3.0 * (G[0, 0] ** 2 + G[1, 1] ** 2 + G[2, 2] ** 2) - 6.0 * (
        G[1, 0] * G[0, 1] + G[2, 0] * G[0, 2] + G[2, 1] * G[1, 2])


def _corner(c, h):
    return h * np.array([1 if c & 1 else -1, 1 if c & 2 else -1, 1 if c & 4 else -1])


def test_linear_field_reproduced():
    """Linear velocity and density at the corners (equilibrium): the fine nodes
    and the shifted coarse point get exactly the linear field (density with the
    reference's curvature term 3 x^2 laplace_rho for an unshifted cell)."""
    rng = np.random.default_rng(5)
    for _ in range(10):
        u0, G, d0, g, f8 = _linear_cell(rng)
        _, cf = oracle.interp_cell("cf", f8, 1.9, 1.8)
        for c in range(8):
            x = _corner(c, 0.25)
            mac = oracle.macroscopic(cf[c])
            assert abs(mac[0] - (d0 + g @ x + 3 * x[0] ** 2 * _laplace(G))) < 1e-16
            assert np.allclose(mac[1:], u0 + G @ x, rtol=0, atol=1e-16)
        off = np.array([0.3, -0.2, 0.1])
        _, fc = oracle.interp_cell("fc", f8, 1.8, 1.9, off)
        mac = oracle.macroscopic(fc)
        assert np.allclose(mac[1:], u0 + G @ off, rtol=0, atol=1e-16)
        assert abs(mac[0] - (d0 + g @ off)) < 1e-16  # no curvature term for a shifted cell
        _, fc0 = oracle.interp_cell("fc", f8, 1.8, 1.9)
        assert abs(oracle.macroscopic(fc0)[0] - (d0 - _laplace(G) / 4)) < 1e-16


def test_destination_strain_is_eps_times_source_strain():
    """The destination's non-equilibrium strain, read back in its own lattice
    units, is eps x the source velocity gradient: eps = 1/2 coarse->fine, 2
    fine->coarse; a constant stress on the source corners does not enter."""
    rng = np.random.default_rng(6)
    omega_src, omega_dst = 1.9, 1.8
    for _ in range(10):
        u0, G, d0, g, f8 = _linear_cell(rng, eq_noise=1e-3, linear_density=False)
        _, cf = oracle.interp_cell("cf", f8, omega_src, omega_dst)
        for c in range(8):
            assert np.allclose(oracle.macroscopic(cf[c])[1:], u0 + G @ _corner(c, 0.25), rtol=0, atol=1e-16)
            s = oracle.source_moments(cf[c], omega_dst)
            want = 0.5 * np.array([G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0],
                                   G[0, 0] - G[1, 1], G[0, 0] - G[2, 2]])
            assert np.allclose(s[4:], want, rtol=0, atol=1e-14)
        _, fc = oracle.interp_cell("fc", f8, omega_dst, omega_src)
        s = oracle.source_moments(fc, omega_src)
        want = 2.0 * np.array([G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0],
                               G[0, 0] - G[1, 1], G[0, 0] - G[2, 2]])
        assert np.allclose(s[4:], want, rtol=0, atol=1e-14)


def _mirror_dirs(axis):
    c = [CX, CY, CZ]
    out = []
    for d in range(27):
        v = [c[0][d], c[1][d], c[2][d]]
        v[axis] = -v[axis]
        out.append(next(e for e in range(27) if (CX[e], CY[e], CZ[e]) == tuple(v)))
    return np.array(out)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_mirror_symmetry(axis):
    """Reflecting the source cell (corners, populations) reflects the outputs."""
    rng = np.random.default_rng(11 + axis)
    md = _mirror_dirs(axis)
    bit = 1 << axis
    _, _, _, _, f8 = _linear_cell(rng)
    f8 += rng.normal(0, 1e-3, (8, 27)) * W
    f8m = np.stack([f8[c ^ bit][md] for c in range(8)])
    _, cf = oracle.interp_cell("cf", f8, 1.9, 1.85)
    _, cfm = oracle.interp_cell("cf", f8m, 1.9, 1.85)
    for c in range(8):
        assert np.allclose(cfm[c], cf[c ^ bit][md], rtol=0, atol=1e-16)
    _, fc = oracle.interp_cell("fc", f8, 1.85, 1.9)
    _, fcm = oracle.interp_cell("fc", f8m, 1.85, 1.9)
    assert np.allclose(fcm, fc[md], rtol=0, atol=1e-16)


def _linear_fields(nest, u0, G):
    nx, ny, nz = nest.coarse_dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pc = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(float)
    pf = np.stack(nest.fine_positions(), axis=1)
    uc = u0 + pc @ G.T
    uf = u0 + pf @ G.T
    Fc = np.stack([oracle.equilibrium(0.0, *u) for u in uc])
    Ff = np.stack([oracle.equilibrium(0.0, *u) for u in uf])
    return Fc, Ff, uc, uf


def test_nest_tables_reproduce_linear_field():
    nest = vfx.NestedBox((16, 14, 14), (3, 3, 4), (11, 10, 11))
    G = np.array([[0.001, 0.002, -0.001], [0.0005, -0.001, 0.002], [0.001, 0.0, -0.0015]])
    u0 = np.array([0.01, -0.02, 0.015])
    Fc, Ff, uc, uf = _linear_fields(nest, u0, G)
    # prolongation onto a zeroed fine level
    Ff0 = oracle.scale_cf(Fc, np.zeros_like(Ff), nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine,
                          1.9, vfx.fine_omega(1.9))
    written = np.abs(Ff0).sum(axis=1) > 0
    fx, fy, fz = nest.fine_dims
    mz, my, mx = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
    ring = ((mx < 2) | (mx >= fx - 2) | (my < 2) | (my >= fy - 2) | (mz < 2) | (mz >= fz - 2)).ravel()
    assert np.array_equal(written, ring)  # exactly the two outer fine layers
    for k in np.nonzero(ring)[0][::7]:
        assert np.allclose(oracle.macroscopic(Ff0[k])[1:], uf[k], rtol=0, atol=1e-15)
    # restriction onto a zeroed coarse level
    Fc0 = oracle.scale_fc(np.zeros_like(Fc), Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine,
                          1.9, vfx.fine_omega(1.9))
    written = np.abs(Fc0).sum(axis=1) > 0
    assert np.array_equal(written, nest.coarse_depth == 2)
    for k in np.nonzero(written)[0]:
        assert np.allclose(oracle.macroscopic(Fc0[k])[1:], uc[k], rtol=0, atol=1e-15)


def refined_oracle_run(nest, Fc, Ff, steps, omega_c):
    omega_f = vfx.fine_omega(omega_c)
    for _ in range(steps):
        Ff = oracle.run(Ff, nest.fine_dims, 2, omega_f)
        Fc = oracle.run(Fc, nest.coarse_dims, 1, omega_c, solid=nest.coarse_state)
        Fc = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, omega_c, omega_f)
        Ff = oracle.scale_cf(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, omega_c, omega_f)
    return Fc, Ff


def shear_wave_nest():
    nest = vfx.NestedBox((16, 32, 10), (3, 8, 2), (11, 24, 8))
    return nest


def shear_wave_init(nest, U=0.02):
    ny = nest.coarse_dims[1]
    k = 2 * np.pi / ny
    nx, _, nz = nest.coarse_dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    Fc = np.stack([oracle.equilibrium(0.0, U * np.sin(k * yy), 0.0, 0.0) for yy in y.ravel()])
    _, yf, _ = nest.fine_positions()
    Ff = np.stack([oracle.equilibrium(0.0, U * np.sin(k * yy), 0.0, 0.0) for yy in yf])
    return Fc, Ff, k


def test_shear_wave_through_nest_decays_at_viscous_rate():
    nest = shear_wave_nest()
    omega_c, U, steps = 1.9, 0.02, 120
    Fc, Ff, k = shear_wave_init(nest, U)
    Fc, Ff = refined_oracle_run(nest, Fc, Ff, steps, omega_c)
    nu = (1.0 / omega_c - 0.5) / 3.0
    amp = U * np.exp(-nu * k * k * steps)
    act = nest.coarse_state == 0
    nx, ny, nz = nest.coarse_dims
    y = np.tile(np.repeat(np.arange(ny), nx), nz)
    ux = np.array([oracle.macroscopic(f)[1] for f in Fc])
    fit_c = (ux[act] @ np.sin(k * y[act])) / (np.sin(k * y[act]) @ np.sin(k * y[act]))
    _, yf, _ = nest.fine_positions()
    uxf = np.array([oracle.macroscopic(f)[1] for f in Ff])
    fit_f = (uxf @ np.sin(k * yf)) / (np.sin(k * yf) @ np.sin(k * yf))
    print(f"analytic {amp:.6e} coarse {fit_c:.6e} fine {fit_f:.6e}")
    assert abs(fit_c / amp - 1) < 0.01
    assert abs(fit_f / amp - 1) < 0.01
    # the fine level carries the same wave: residual small against the amplitude
    assert np.abs(uxf - fit_f * np.sin(k * yf)).max() < 0.02 * amp


# ---------------------------------------------------------------- wall nest
SMALL_CUBE = dict(coarse_dims=(40, 24, 20), nest_cells=(20, 14, 12), cube=6, cube_x=14, upstream=6,
                  u_ref=0.05, re=200.0)


def small_cube_case():
    from virtualfluids_amd import geometry
    return geometry.wall_mounted_cube(**SMALL_CUBE)


def cube_oracle_groups(tables):
    out = []
    for b in tables:
        if b.kind == 4:
            out.append(oracle.BCGroup(b.nodes, kind=4, rho=b.rho, nbr=b.neighbors, normal=b.normal))
        elif b.velocity is not None:
            out.append(oracle.BCGroup(b.nodes, b.q27, *b.velocity, kind=b.kind))
        else:
            out.append(oracle.BCGroup(b.nodes, b.q27, kind=b.kind))
    return out


def cube_masks(case):
    """Oracle node states: 1 solid with boundary links, 2 not computed (inactive
    coarse core; the fine floor layer, which the fine box's periodic wrap
    touches from the top layers that the coarse-to-fine step rebuilds)."""
    nest = case.nest
    cm = np.where(nest.coarse_state != 0, 2, 0).astype(np.uint8)
    cm[case.coarse_solid] = np.maximum(cm[case.coarse_solid], 1)
    fx, fy, fz = nest.fine_dims
    fm = case.fine_solid.astype(np.uint8)
    fm.reshape(fz, fy, fx)[0] = 2
    return cm, fm


def refined_oracle_run_walls(case, Fc, Ff, steps, omega_c):
    nest = case.nest
    omega_f = vfx.fine_omega(omega_c)
    cm, fm = cube_masks(case)
    cb, fb = cube_oracle_groups(case.coarse_bcs), cube_oracle_groups(case.fine_bcs)
    for _ in range(steps):
        Ff = oracle.run(Ff, nest.fine_dims, 1, omega_f, solid=fm, bcs=fb)
        Ff = oracle.run(Ff, nest.fine_dims, 1, omega_f, solid=fm, bcs=fb)
        Fc = oracle.run(Fc, nest.coarse_dims, 1, omega_c, solid=cm, bcs=cb)
        Fc = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, omega_c, omega_f,
                             nest.fc_offset)
        Ff = oracle.scale_cf(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, omega_c, omega_f,
                             nest.cf_offset)
    return Fc, Ff


def test_wall_nest_tables():
    """The nest standing on the floor: no interface on the floor face, the
    coarse-to-fine cells that touch the floor move up one cell with offset
    (0, 0, -1), every interpolation source cell is all fluid and valid."""
    case = small_cube_case()
    nest = case.nest
    nx, ny, nz = nest.coarse_dims
    fx, fy, fz = nest.fine_dims
    csol = case.coarse_solid.reshape(nz, ny, nx)
    fsol = case.fine_solid.reshape(fz, fy, fx)
    depth = nest.coarse_depth.reshape(nz, ny, nx)
    off = np.stack(nest.cf_offset)
    moved = np.abs(off).sum(axis=0) > 0
    assert moved.any() and (off[:, moved] == np.array([[0.0], [0.0], [-1.0]])).all()
    for base in nest.cf_coarse:
        x, y, z = base % nx, (base // nx) % ny, base // (nx * ny)
        for c in range(8):
            cx, cy, cz = x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1)
            assert not csol[cz, cy, cx] and depth[cz, cy, cx] <= 1
    for base in nest.fc_fine:
        x, y, z = base % fx, (base // fx) % fy, base // (fx * fy)
        for c in range(8):
            assert not fsol[z + ((c >> 2) & 1), y + ((c >> 1) & 1), x + (c & 1)]
    # coarse-to-fine cells sit on the five interface faces only: each one is on
    # an x or y face or the top face, and the floor row is covered through the
    # side faces
    mx, my, mz = nest.cf_fine % fx, (nest.cf_fine // fx) % fy, nest.cf_fine // (fx * fy)
    cx, cy, cz = nest.lo[0] + mx // 2, nest.lo[1] + my // 2, nest.lo[2] + mz // 2
    on_face = ((cx == nest.lo[0]) | (cx == nest.hi[0] - 1) | (cy == nest.lo[1]) | (cy == nest.hi[1] - 1) |
               (cz == nest.hi[2] - 1))
    assert on_face.all() and (cz == 0).any()


def test_wall_nest_oracle_runs():
    """The oracle advances the small wall-mounted cube case on both levels
    (every wall link closed, no interface link left open) and the flow stays
    bounded and attached to the inflow."""
    case = small_cube_case()
    nest = case.nest
    Nc, Nf = int(np.prod(nest.coarse_dims)), int(np.prod(nest.fine_dims))
    u0 = case.u_ref
    Fc = np.tile(oracle.equilibrium(0.0, u0, 0.0, 0.0), (Nc, 1))
    Ff = np.tile(oracle.equilibrium(0.0, u0, 0.0, 0.0), (Nf, 1))
    Fc, Ff = refined_oracle_run_walls(case, Fc, Ff, 40, case.omega)
    mc = np.array([oracle.macroscopic(f) for f in Fc])
    act = (nest.coarse_state == 0) & ~case.coarse_solid
    assert np.isfinite(mc).all() and np.abs(mc[act, 1:]).max() < 0.2
    mf = np.array([oracle.macroscopic(f) for f in Ff])
    assert np.isfinite(mf).all() and np.abs(mf[~case.fine_solid, 1:]).max() < 0.2


def test_interface_cells_apart_on_the_nests():
    """Fine-to-coarse writes coarse nodes two cells inside the nest, coarse-to-
    fine reads the outermost layer: the cell sets never meet, so one launch may
    do both (planting an overlap is detected)."""
    from virtualfluids_amd.refinement import _cell_corners, interface_cells_apart

    case = small_cube_case()
    for nest in (case.nest, shear_wave_nest()):
        assert interface_cells_apart(nest)
    nest = case.nest
    bad = copy.copy(nest)
    bad.fc_coarse = np.concatenate([nest.fc_coarse, _cell_corners(nest.cf_coarse[:1], nest.coarse_dims)[:1]])
    assert not interface_cells_apart(bad)
