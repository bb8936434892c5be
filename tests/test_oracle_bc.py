"""CPU pins of the oracle's boundary conditions beyond the reference's CPU
strategies (VelocityBounceBack is pinned bit-exactly by the cavity goldens):
the GPU-only interpolated no-slip / velocity / slip and non-equilibrium
pressure kernels have no CPU twin in the reference, so they are pinned by
flows with closed-form solutions whose wall positions follow from q."""
import numpy as np
import pytest

import oracle
from tests import cases

OMEGA = 1.8  # K17 is stable here (it is not for omega below about 1.7)


def channel_profile(ny, qb, qt, kind, force=0.0, top_u=0.0, omega=OMEGA):
    nx = nz = 2
    solid, groups = cases.plate_channel(nx, ny, nz, qb, qt)
    bcs = []
    for name, (nodes, q) in groups.items():
        u = top_u if name == "top" else 0.0
        bcs.append(oracle.BCGroup(nodes, q, ux=u, kind=kind))
    nu = (1 / omega - 0.5) / 3
    H = ny - 2
    steps = int(12 * H * H / nu / np.pi ** 2) + 1000
    F = oracle.run(np.zeros((nx * ny * nz, 27)), (nx, ny, nz), steps, omega, solid=solid, bcs=bcs,
                   force=(force, 0.0, 0.0))
    ux = np.array([oracle.macroscopic(f)[1] for f in F]).reshape(nz, ny, nx)[1, :, 1]
    return np.arange(1, ny - 1), ux[1:-1], nu


@pytest.mark.parametrize("qb,qt", [(0.3, 0.7), (0.5, 0.5), (0.8, 0.2)])
def test_noslip_interpolated_poiseuille_walls_sit_at_q(qb, qt):
    """Body-force Poiseuille flow between plates placed at fractions q of the
    wall links: the fitted parabola vanishes at y = 1 - q_bottom and
    y = ny - 2 + q_top, and its curvature gives the lattice viscosity."""
    ny = 24
    Fh = 1e-5 / (ny - 2) ** 2
    y, ux, nu = channel_profile(ny, qb, qt, kind=1, force=Fh)
    c = np.polyfit(y, ux, 2)
    r = np.sort(np.roots(c).real)
    assert abs(r[0] - (1 - qb)) < 0.05 and abs(r[1] - (ny - 2 + qt)) < 0.05, r
    nu_eff = 2 * Fh / (2 * -c[0])
    assert abs(nu_eff / nu - 1) < 0.01


@pytest.mark.parametrize("qb,qt", [(0.3, 0.7), (0.5, 0.5)])
def test_velocity_interpolated_couette_is_linear(qb, qt):
    """Couette flow: the top plate moves with U (velocity-interpolated
    boundary), the bottom one is a no-slip wall; the steady profile is linear
    and vanishes at the bottom wall position and reaches U at the top one."""
    ny, U = 20, 0.01
    y, ux, _ = channel_profile(ny, qb, qt, kind=2, top_u=U)
    c = np.polyfit(y, ux, 1)
    y0, y1 = 1 - qb, ny - 2 + qt
    assert abs(np.polyval(c, y0)) < 0.01 * U
    assert abs(np.polyval(c, y1) - U) < 0.01 * U
    assert np.abs(np.polyval(c, y) - ux).max() < 2e-3 * U


def test_slip_keeps_uniform_flow_uniform():
    """A uniform stream along slip walls is a steady state (no shear at the
    wall); along no-slip walls it is not."""
    nx, ny, nz, U = 2, 10, 2, 0.02
    solid, groups = cases.plate_channel(nx, ny, nz, 0.5, 0.5)
    N = nx * ny * nz
    F0 = oracle.init_field(np.zeros(N), np.full(N, U), np.zeros(N), np.zeros(N))
    res = {}
    for kind in (3, 1):
        bcs = [oracle.BCGroup(nodes, q, kind=kind) for nodes, q in groups.values()]
        F = oracle.run(F0, (nx, ny, nz), 300, OMEGA, solid=solid, bcs=bcs)
        ux = np.array([oracle.macroscopic(f)[1] for f in F]).reshape(nz, ny, nx)[1, 1:-1, 1]
        res[kind] = np.abs(ux - U).max() / U
    assert res[3] < 1e-3
    assert res[1] > 0.1


@pytest.mark.parametrize("normal", [1, 2, 3, 6])
def test_pressure_boundary_replaces_the_density_of_an_equilibrium_neighbour(normal):
    """The inward neighbour holds the equilibrium of a density drho1 that
    differs from the prescribed rho_bc, with a velocity along all three axes:
    the non-equilibrium pressure boundary writes exactly feq(rho_bc, u) on the
    directions with c . normal <= 0 (the neighbour's non-equilibrium part is
    zero) and nothing else.  Dropping the density replacement, or replacing
    the wrong directions, fails by about w * |drho1 - rho_bc|."""
    drho1, rho_bc, u = 5e-3, -1e-3, (0.01, 0.004, -0.003)
    fn = oracle.equilibrium(drho1, *u)
    out, wr = oracle.pressure_bc(fn, rho_bc, normal)
    c = np.stack([cases.CX, cases.CY, cases.CZ])
    expect = (c.T @ c[:, normal]) <= 0
    assert np.array_equal(wr, expect)
    np.testing.assert_allclose(out[wr], oracle.equilibrium(rho_bc, *u)[wr], rtol=0, atol=2e-17)
    assert np.abs(oracle.equilibrium(rho_bc, *u) - fn)[wr].min() > 1e-5


def test_pressure_boundary_in_a_channel_step():
    """The same in a one-step channel run: the outlet node's replaced
    populations come from its neighbour's state after the collision and
    streaming (K17 moves an equilibrium by O(u^3)), the others stay as the
    streaming left them."""
    nx, ny, nz = 6, 5, 5
    solid, groups, onodes, onbrs = cases.inflow_channel(nx, ny, nz)
    N = nx * ny * nz
    drho1, rho_bc, u = 5e-3, -1e-3, (0.01, 0.004, -0.003)
    F0 = oracle.init_field(np.full(N, drho1), np.full(N, u[0]), np.full(N, u[1]), np.full(N, u[2]))
    bcs = [oracle.BCGroup(onodes, kind=4, rho=rho_bc, nbr=onbrs, normal=1),
           oracle.BCGroup(*groups["wall"], kind=1), oracle.BCGroup(*groups["inlet"], ux=u[0], kind=2)]
    F_nobc = oracle.run(F0, (nx, ny, nz), 1, OMEGA, solid=solid, bcs=bcs[1:] + [
        oracle.BCGroup(onodes, kind=4, rho=drho1, nbr=onbrs, normal=1)])
    F = oracle.run(F0, (nx, ny, nz), 1, OMEGA, solid=solid, bcs=bcs)
    k = onodes[len(onodes) // 2]
    kn = onbrs[len(onodes) // 2]
    out, wr = oracle.pressure_bc(F[kn], rho_bc, 1)
    np.testing.assert_array_equal(F[k][wr], out[wr])
    np.testing.assert_array_equal(F[k][~wr], F_nobc[k][~wr])
    assert np.abs(F[k][wr] - F_nobc[k][wr]).min() > 1e-5


@pytest.mark.parametrize("q", [0.3, 0.5, 0.8])
def test_slip_walls_let_a_body_force_accelerate_a_plug_flow(q):
    """Slip walls on both sides of a periodic channel exert no tangential
    friction: a body force (acceleration 2 F per step for the K17 force F)
    accelerates the fluid as a plug, u_x = 2 F t everywhere.  No-slip walls at
    the same positions leave a strongly sheared profile instead."""
    nx, ny, nz, steps, Fh = 2, 16, 2, 400, 1e-6
    solid, groups = cases.plate_channel(nx, ny, nz, q, q)
    N = nx * ny * nz
    prof = {}
    for kind in (3, 1):
        bcs = [oracle.BCGroup(*groups["bottom"], kind=kind), oracle.BCGroup(*groups["top"], kind=kind)]
        F = oracle.run(np.zeros((N, 27)), (nx, ny, nz), steps, OMEGA, solid=solid, bcs=bcs, force=(Fh, 0.0, 0.0))
        prof[kind] = np.array([oracle.macroscopic(f)[1] for f in F]).reshape(nz, ny, nx)[1, 1:-1, 1]
    ux = prof[3]
    assert abs(ux.mean() / (2 * Fh * steps) - 1) < 0.02
    assert (ux.max() - ux.min()) / ux.mean() < 0.03
    assert (prof[1].max() - prof[1].min()) / prof[1].mean() > 0.5


@pytest.mark.parametrize("q", [0.3, 0.5])
def test_slip_walls_stop_the_normal_flow_and_keep_the_tangential(q):
    """A uniform stream u = (0.02, 0.01, 0) into slip walls at y: the normal
    component is stopped (mean u_y from 0.01 to below 1e-4 after 2000 steps)
    while the tangential momentum stays (mean u_x within 1e-3 of 0.02).  A
    slip kernel that kept the fluid's normal velocity as the wall velocity
    lets the normal flow through the wall and goes unstable."""
    nx, ny, nz = 2, 16, 2
    solid, groups = cases.plate_channel(nx, ny, nz, q, q)
    N = nx * ny * nz
    F0 = oracle.init_field(np.zeros(N), np.full(N, 0.02), np.full(N, 0.01), np.zeros(N))
    F0[solid] = 0.0
    bcs = [oracle.BCGroup(*groups["bottom"], kind=3), oracle.BCGroup(*groups["top"], kind=3)]
    F = oracle.run(F0, (nx, ny, nz), 2000, OMEGA, solid=solid, bcs=bcs)
    m = np.array([oracle.macroscopic(f) for f in F]).reshape(nz, ny, nx, 4)[:, 1:-1, :]
    assert np.isfinite(m).all()
    assert abs(m[..., 2].mean()) < 1e-4
    assert abs(m[..., 1].mean() / 0.02 - 1) < 1e-3


def test_inflow_outflow_channel_reaches_prescribed_outlet_density():
    """Velocity inlet + pressure outlet: the steady density next to the outlet
    is the prescribed one and the mass flux is the same through every cross
    section."""
    nx, ny, nz, U = 24, 8, 8, 0.02
    solid, groups, onodes, onbrs = cases.inflow_channel(nx, ny, nz)
    N = nx * ny * nz
    bcs = [oracle.BCGroup(onodes, kind=4, rho=0.0, nbr=onbrs, normal=1),
           oracle.BCGroup(*groups["wall"], kind=1), oracle.BCGroup(*groups["inlet"], ux=U, kind=2)]
    F = oracle.run(np.zeros((N, 27)), (nx, ny, nz), 3000, 1.9, solid=solid, bcs=bcs)
    m = np.array([oracle.macroscopic(f) for f in F]).reshape(nz, ny, nx, 4)
    inner = m[1:-1, 1:-1, :, :]
    assert np.abs(inner[:, :, nx - 2, 0]).max() < 2e-3
    flux = ((1 + inner[..., 0]) * inner[..., 1]).sum(axis=(0, 1))[1:-1]
    assert np.abs(flux / flux.mean() - 1).max() < 2e-2


# ------------------------------------ velocity with pressure, outflow ------
W = cases.W
C27 = np.stack([cases.CX, cases.CY, cases.CZ], axis=1)
FACES = [1, 2, 3, 4, 5, 6]  # dP00, dM00, d0P0, d0M0, d00P, d00M


def _random_links(rng):
    q = np.full(27, -1.0, np.float32)
    links = rng.choice(np.arange(1, 27), 9, replace=False)
    q[links] = rng.uniform(0.0, 1.0, links.size)
    return q


def test_velocity_with_pressure_closed_forms():
    """VelocityWithPressureInterpolatedCompressible (KernelUtilities.h:207-213):
    at rest with density deviation drho0 the plain velocity boundary writes
    the equilibrium w_d drho0 and the pressure variant the bulk state 0; a
    moving equilibrium at drho = 0 with the wall velocity is written back
    unchanged by both; in general they differ by exactly -w_d drho."""
    rng = np.random.default_rng(40)
    for _ in range(50):
        q = _random_links(rng)
        om = rng.uniform(1.2, 1.95)
        drho0 = rng.uniform(-0.02, 0.02)
        rest = W * drho0
        a, wa = oracle.interp_bc(2, rest, q, om)
        b, wb = oracle.interp_bc(5, rest, q, om)
        assert np.array_equal(wa, wb) and wa.sum() == 9
        np.testing.assert_allclose(a[wa], rest[wa], rtol=0, atol=1e-17)
        np.testing.assert_allclose(b[wb], 0.0, rtol=0, atol=1e-17)
        u = rng.uniform(-0.05, 0.05, 3)
        feq = oracle.equilibrium(0.0, *u)
        for kind in (2, 5):
            v, w = oracle.interp_bc(kind, feq, q, om, u)
            np.testing.assert_allclose(v[w], feq[w], rtol=0, atol=2e-17)
        p = feq + rng.normal(0, 1e-3, 27) * W + W * drho0
        dr = oracle.macroscopic(p)[0]
        a, wa = oracle.interp_bc(2, p, q, om, u)
        b, wb = oracle.interp_bc(5, p, q, om, u)
        np.testing.assert_allclose(b[wb] - a[wa], -W[wb] * dr, rtol=1e-12, atol=1e-19)


@pytest.mark.parametrize("kind", [6, 7])
def test_outflow_writes_the_entering_directions(kind):
    rng = np.random.default_rng(41)
    f = rng.normal(0, 1e-3, 27) * W
    fn = rng.normal(0, 1e-3, 27) * W
    for normal in FACES:
        _, w = oracle.outflow_bc(kind, f, fn, normal)
        assert np.array_equal(w, C27 @ C27[normal] < 0), normal
        assert w.sum() == 9
    for normal in (0, 7, 19):  # not a face direction: nothing (the kernels' default branch)
        assert not oracle.outflow_bc(kind, f, fn, normal)[1].any()


@pytest.mark.parametrize("kind", [6, 7])
def test_outflow_keeps_a_uniform_state(kind):
    """Node and neighbour in the same equilibrium: every entering population
    stays what it is (the boundary passes a uniform stream through)."""
    rng = np.random.default_rng(42)
    for _ in range(30):
        f = oracle.equilibrium(rng.uniform(-0.01, 0.01), *rng.uniform(-0.05, 0.05, 3))
        for normal in FACES:
            v, w = oracle.outflow_bc(kind, f, f, normal)
            np.testing.assert_allclose(v[w], f[w], rtol=0, atol=1e-17)


def test_outflow_characteristic_and_pressure_correction():
    """Entering population = f + cs (fn - f) with cs = 1/sqrt(3) (upwind
    advection at the sound speed); the corrected kernel subtracts w_d factor
    drho, so a node at rest with density drho0 and factor 1 lets the bulk
    state 0 enter; its neighbour values of directions with a negative
    component are the node's own (the kernel's neighbour indices)."""
    rng = np.random.default_rng(43)
    cs = 1.0 / np.sqrt(3.0)
    for normal in FACES:
        f = rng.normal(0, 1e-3, 27) * W
        fn = rng.normal(0, 1e-3, 27) * W
        v, w = oracle.outflow_bc(6, f, fn, normal)
        # cs is 1 / sqrtf(3): single-precision root, exact only to ~3e-8
        np.testing.assert_allclose(v[w], (f + cs * (fn - f))[w], rtol=0, atol=1e-7 * np.abs(fn - f).max())
        drho0 = rng.uniform(-0.02, 0.02)
        v, w = oracle.outflow_bc(7, W * drho0, W * drho0, normal, correction=1.0)
        np.testing.assert_allclose(v[w], 0.0, rtol=0, atol=1e-17)
        neg = (C27 < 0).any(axis=1)
        a, w = oracle.outflow_bc(7, f, fn, normal, correction=0.3)
        fn2 = fn.copy()
        fn2[neg] += 1e-3
        b, _ = oracle.outflow_bc(7, f, fn2, normal, correction=0.3)
        assert np.array_equal(a, b)
        fn3 = fn.copy()
        fn3[~neg] += 1e-3
        c, _ = oracle.outflow_bc(7, f, fn3, normal, correction=0.3)
        assert np.array_equal(c[w & neg], a[w & neg])
        if normal in (1, 3, 5):  # outflow on a + face: every entering direction has a negative component
            assert not (w & ~neg).any()
        else:
            assert (w & ~neg).sum() == 4 and not np.any(c[w & ~neg] == a[w & ~neg])


def _pulse_run(right_kind, steps, nx=100, x0=70, amp=1e-3):
    """Density pulse in a tube along x (2 x 2 periodic cross section): no-slip
    bounce-back wall on the left, `right_kind` on the right (6 outflow, 0
    bounce-back wall)."""
    ny = nz = 2
    solid = np.zeros((nz, ny, nx), bool)
    solid[:, :, 0] = solid[:, :, nx - 1] = True
    x = np.arange(nx)
    drho = np.tile(amp * np.exp(-0.5 * ((x - x0) / 3.0) ** 2), ny * nz)
    z = np.zeros(nx * ny * nz)
    F = oracle.init_field(drho, z, z, z)

    def face(xf, toward):
        nodes = np.array([xf + nx * (y + ny * zz) for zz in range(nz) for y in range(ny)], np.int64)
        q = np.full((27, nodes.size), -1.0, np.float32)
        for d in range(1, 27):
            if cases.CX[d] == toward:
                q[d, :] = 0.5
        return nodes, q

    ln, lq = face(1, -1)
    bcs = [oracle.BCGroup(ln, lq)]
    rn, rq = face(nx - 2, 1)
    if right_kind == 0:
        bcs.append(oracle.BCGroup(rn, rq))
    else:
        bcs.append(oracle.BCGroup(rn, kind=right_kind, nbr=rn - 1, normal=1))
    F = oracle.run(F, (nx, ny, nz), steps, 1.9, solid=solid.ravel(), bcs=bcs)
    return np.array([oracle.macroscopic(f)[0] for f in F]).reshape(nz, ny, nx)[0, 0]


def test_outflow_absorbs_a_pressure_pulse():
    """The right-going half of the pulse meets the boundary after ~50 steps; at
    step 100 its reflection would sit near x = 70.  A bounce-back wall reflects
    it whole; the non-reflecting outflow lets it leave."""
    wall = _pulse_run(0, 100)
    out = _pulse_run(6, 100)
    r_wall = np.abs(wall[45:95]).max()
    r_out = np.abs(out[45:95]).max()
    print(f"reflected amplitude: wall {r_wall:.3e}, outflow {r_out:.3e}")
    assert r_wall > 3e-4
    assert r_out < 0.15 * r_wall


# ------------------------------------------------------------- wall models --
def _wall_node(rng):
    """A node above a floor (links with cz < 0 are wall links, q = 0.5) in a
    random post-collision state."""
    p = oracle.equilibrium(rng.uniform(-0.01, 0.01), *rng.uniform(-0.05, 0.05, 3)) + rng.normal(0, 1e-3, 27) * W
    q = np.full(27, -1.0, np.float32)
    q[C27[:, 2] < 0] = 0.5
    return p, q


def test_wall_model_without_sampled_velocity_is_no_slip():
    """Sampled velocity 0 and filtered state 0: the clip bounds are 0, so the
    wall velocity is 0 -- StressBounceBack is the plain bounce-back, the
    pressure variant the bounce-back minus w drho, StressCompressible the
    interpolated no-slip minus w drho (the NE link with 2/27); the force
    monitor is the momentum exchange 2 sum c_d p_d over the wall links."""
    rng = np.random.default_rng(50)
    for _ in range(20):
        p, q = _wall_node(rng)
        om = rng.uniform(1.5, 1.95)
        dr = sum(p)  # plain sum as the kernels take it (difference to the pyramid sum is rounding)
        links = np.nonzero(q >= 0)[0]
        v10, w10, *_, f10 = oracle.stress_bc(10, p, q, om, (0, 0, 0), (0, 0, 1), 2, 0.1)
        v11, w11, *_ = oracle.stress_bc(11, p, q, om, (0, 0, 0), (0, 0, 1), 2, 0.1)
        v9, w9, *_ = oracle.stress_bc(9, p, q, om, (0, 0, 0), (0, 0, 1), 2, 0.1)
        inv = cases.INV
        assert sorted(np.nonzero(w10)[0]) == sorted(inv[links])
        for d in links:
            assert v10[inv[d]] == p[d]
            assert abs(v11[inv[d]] - (p[d] - W[d] * dr)) < 1e-17
        ns, _ = oracle.interp_bc(1, p, q, om)
        for d in links:
            wp = 2 / 27 if d == 7 else W[d]
            assert abs(v9[inv[d]] - (ns[inv[d]] - wp * dr)) < 1e-16
        np.testing.assert_allclose(f10, 2 * (C27[links].T * p[links]).sum(axis=1), rtol=0, atol=1e-15)


def test_wall_model_log_law_and_clip():
    """Filtered exchange velocity (U0, V0, 0) over a floor: u* = kappa |U| /
    ln((offset + q_B) / z0); the wall velocity is -3 F, clipped to twice the
    filtered tangential velocity; the filters move by eps = 0.001f."""
    rng = np.random.default_rng(51)
    for trial in range(20):
        p, q = _wall_node(rng)
        U = np.array([rng.uniform(0.01, 0.05), rng.uniform(-0.02, 0.02), 0.0])
        z0 = [1e-3, 0.5, 2.4][trial % 3]
        v, w, el, wm, us, force = oracle.stress_bc(10, p, q, 1.9, U, (0, 0, 1), 2, z0, exchange_velocity=U,
                                                   wall_velocity=(0.01, 0, 0))
        eps = float(np.float32(0.001))
        # the sampled z velocity is read from the y array (Stress.cu:57-59)
        np.testing.assert_allclose(el, eps * U[[0, 1, 1]] + (1 - eps) * U, rtol=1e-15)
        assert abs(us - 0.4 * np.hypot(U[0], U[1]) / np.log(2.5 / z0)) <= 1e-14 * us
        # wall velocity from the BW link (-x, -z): out = p - 6 w (-Ux - Uz), Uz = 0
        bw, te = 12, 11
        Ux = (v[te] - p[bw]) / (6 * W[bw])
        assert abs(Ux) <= 2 * el[0] * (1 + 1e-12)
        if trial % 3 == 2:  # z0 close to z: large stress, the clip holds
            assert abs(abs(Ux) - 2 * el[0]) < 1e-12
