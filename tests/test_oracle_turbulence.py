"""CPU pins of the oracle's eddy-viscosity models of the K17 collision
(vf::lbm::TurbulenceModel: Smagorinsky, AMD, QR).

  * golden vectors of the compiled reference closure
    (runK17CompressibleNavierStokes<Smagorinsky / AMD / QR>,
    tests/golden/make_golden.py g6) -- bit exact, and the live reference where
    it is built here;
  * closed forms: a node at rest carrying the Chapman-Enskog non-equilibrium
    part of a known traceless strain rate S has the velocity gradients S, so
    Smagorinsky gives Cs^2 sqrt(2 S:S) and QR gives C max(-det S, 0) / (S:S / 2);
    the AMD kernel on a linear velocity field (central and one-sided
    differences are exact) gives max(0, -C (G G^T):S) / (G:G);
  * special cases: SGS constant 0 is the plain K17; a given turbulent viscosity
    (AMD, no update in the collision) is the plain K17 at
    omega / (1 + 3 omega nu_t); the viscosity of a state is invariant under
    axis permutations.
"""
import os

import numpy as np
import pytest

import oracle
from tests import cases

GOLD = os.path.join(os.path.dirname(__file__), "golden")
W = cases.W
MODELS = ["smagorinsky", "amd", "qr"]
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference closure not built here")


@pytest.mark.parametrize("model", MODELS)
def test_turbulence_golden(model):
    g = np.load(os.path.join(GOLD, "g6_k17_turbulence.npz"))
    f_in, om, nut_in, sgs = (g[f"{model}_{k}"] for k in ("f_in", "omega", "nut_in", "sgs"))
    for i in range(len(om)):
        f, mv, nut = oracle.k17_collide_turb(f_in[i], om[i], model, sgs[i], nut_in[i])
        assert np.array_equal(f, g[f"{model}_f_out"][i]), i
        assert np.array_equal(mv, g[f"{model}_macro"][i]), i
        assert nut == g[f"{model}_nut_out"][i], i


@needs_ref
@pytest.mark.parametrize("model", MODELS)
def test_turbulence_matches_reference_random(model):
    rng = np.random.default_rng(17)
    for i in range(150):
        f = rng.uniform(-1, 1, 27) * [1e-3, 3e-2, 2e-1][i % 3] * W
        om = rng.uniform(0.5, 1.999)
        nut = rng.uniform(0.0, 0.1)
        sgs = rng.uniform(0.0, 1.0)
        force = rng.uniform(-1e-5, 1e-5, 3) if i % 5 == 0 else np.zeros(3)
        a = oracle.k17_collide_turb(f, om, model, sgs, nut, force=force)
        b = oracle.k17_collide_turb(f, om, model, sgs, nut, force=force, impl="ref")
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2], i


def test_zero_smagorinsky_constant_is_plain_k17():
    rng = np.random.default_rng(3)
    for i in range(100):
        f = rng.uniform(-1, 1, 27) * 3e-2 * W
        om = rng.uniform(0.6, 1.99)
        a, mva, nut = oracle.k17_collide_turb(f, om, "smagorinsky", 0.0, 0.0)
        b, mvb = oracle.k17_collide(f, om)
        assert nut == 0.0
        assert np.array_equal(a, b) and np.array_equal(mva, mvb)


def test_given_viscosity_relaxes_at_effective_omega():
    """AMD's viscosity comes from outside the collision: the node is the plain
    K17 at the rate of the total viscosity nu + nu_t, 1 / (3 (nu + nu_t) + 1/2)."""
    rng = np.random.default_rng(4)
    for i in range(100):
        f = rng.uniform(-1, 1, 27) * 3e-2 * W
        om = rng.uniform(0.6, 1.99)
        nut = rng.uniform(0.0, 0.05)
        a, mva, nut_out = oracle.k17_collide_turb(f, om, "amd", 0.3, nut)
        b, mvb = oracle.k17_collide(f, om / (1.0 + 3.0 * om * nut))
        assert nut_out == nut
        assert np.array_equal(a, b) and np.array_equal(mva, mvb)
        nu, nu_t = (1.0 / om - 0.5) / 3.0, nut
        assert abs(om / (1.0 + 3.0 * om * nut) - 1.0 / (3.0 * (nu + nu_t) + 0.5)) < 1e-14


def _strained_node(S, omega):
    """Populations (f - w) of a node at rest (rho = 1) whose non-equilibrium part
    is the Chapman-Enskog one of the velocity gradient S:
    f_i = -w_i / (c_s^2 omega) c_i.S.c_i (S traceless)."""
    c = np.stack([cases.CX, cases.CY, cases.CZ], axis=1).astype(float)
    return -W / (omega / 3.0) * np.einsum("ia,ab,ib->i", c, S, c)


def _random_strain(rng, amp):
    A = rng.uniform(-1, 1, (3, 3))
    S = 0.5 * (A + A.T)
    S -= np.trace(S) / 3.0 * np.eye(3)
    return amp * S


@pytest.mark.parametrize("model", ["smagorinsky", "qr"])
def test_model_viscosity_closed_form(model):
    rng = np.random.default_rng(11)
    for i in range(200):
        S = _random_strain(rng, [1e-4, 1e-3, 1e-2][i % 3])
        om = rng.uniform(0.7, 1.95)
        C = rng.uniform(0.05, 0.5)
        f = _strained_node(S, om)
        _, mv, nut = oracle.k17_collide_turb(f, om, model, C, 0.0)
        assert np.abs(mv).max() < 1e-17  # at rest, no mass added
        SS = np.sum(S * S)
        want = C * C * np.sqrt(2.0 * SS) if model == "smagorinsky" else C * max(-np.linalg.det(S), 0.0) / (0.5 * SS)
        assert abs(nut - want) <= 1e-10 * max(abs(want), C * np.sqrt(SS) * 1e-3), (i, nut, want)


def _perm(f, axes):
    """Populations after relabelling the axes (axes = image of x, y, z)."""
    c = np.stack([cases.CX, cases.CY, cases.CZ], axis=1)
    idx = {tuple(v): d for d, v in enumerate(c)}
    g = np.zeros_like(f)
    for d in range(27):
        g[idx[tuple(c[d][list(axes)])]] = f[d]
    return g


@pytest.mark.parametrize("model", ["smagorinsky", "qr"])
def test_model_viscosity_is_rotation_invariant(model):
    rng = np.random.default_rng(12)
    for i in range(60):
        f = rng.uniform(-1, 1, 27) * 3e-2 * W
        om = rng.uniform(0.7, 1.95)
        ref = oracle.k17_collide_turb(f, om, model, 0.2, 0.01)
        for axes in [(1, 0, 2), (2, 1, 0), (0, 2, 1)]:
            got = oracle.k17_collide_turb(_perm(f, axes), om, model, 0.2, 0.01)
            assert abs(got[2] - ref[2]) <= 1e-12 * abs(ref[2]) + 1e-300
            np.testing.assert_allclose(got[0], _perm(ref[0], axes), rtol=0, atol=1e-15)


def _amd_closed_form(G, C):
    S = 0.5 * (G + G.T)
    num = np.einsum("ik,jk,ij->", G, G, S)
    return max(0.0, -C * num) / np.sum(G * G)


def test_amd_linear_field_closed_form():
    """A box walled by a solid layer holding u = G x: every fluid node, including
    those next to the walls (one-sided differences), sees the gradient G."""
    rng = np.random.default_rng(5)
    dims = (9, 7, 6)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()]).astype(float)
    solid = np.ones((nz, ny, nx), bool)
    solid[1:-1, 1:-1, 1:-1] = False
    fluid = ~solid.ravel()
    for trial in range(20):
        G = rng.uniform(-1e-2, 1e-2, (3, 3))
        u0 = rng.uniform(-0.05, 0.05, 3)
        u = u0[:, None] + G @ pos
        C = rng.uniform(0.1, 0.4)
        nut = oracle.turbulent_viscosity_amd(dims, u[0], u[1], u[2], C, solid=solid)
        want = _amd_closed_form(G, C)
        np.testing.assert_allclose(nut[fluid], want, rtol=1e-9, atol=1e-18)
        assert np.all(nut[~fluid] == 0.0)


def test_amd_special_cases():
    dims = (6, 5, 4)
    n = 6 * 5 * 4
    rng = np.random.default_rng(6)
    # uniform flow: no gradient, zero viscosity (the 0/0 branch)
    u = np.full(n, 0.03)
    assert np.all(oracle.turbulent_viscosity_amd(dims, u, -u, 0 * u, 0.3) == 0.0)
    # a node whose two x neighbours are solid has no x derivative
    solid = np.zeros((4, 5, 6), bool)
    solid[2, 2, 1] = solid[2, 2, 3] = True
    k = 2 + 6 * (2 + 5 * 2)
    vx = rng.uniform(-0.01, 0.01, n)
    vy = rng.uniform(-0.01, 0.01, n)
    vz = rng.uniform(-0.01, 0.01, n)
    a = oracle.turbulent_viscosity_amd(dims, vx, vy, vz, 0.3, solid=solid)
    vx2, vy2, vz2 = vx.copy(), vy.copy(), vz.copy()
    for v in (vx2, vy2, vz2):  # values of the solid neighbours are never read
        v[k - 1] += 1.0
        v[k + 1] -= 1.0
    b = oracle.turbulent_viscosity_amd(dims, vx2, vy2, vz2, 0.3, solid=solid)
    assert a[k] == b[k]


def test_run_zero_constant_is_plain_run():
    dims = (6, 5, 4)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=2, u_amp=0.05))
    plain = oracle.run(F0, dims, 6, 1.8)
    F, nut = oracle.run_turb(F0, dims, 6, 1.8, "smagorinsky", 0.0)
    assert np.array_equal(F, plain) and np.all(nut == 0.0)


def test_run_amd_viscosity_follows_last_velocities():
    """After a step the AMD viscosity is the kernel applied to the velocities
    the collision just wrote (computed after the collision, used by the next)."""
    dims = (7, 6, 5)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=8, u_amp=0.05))
    F, nut, macro = oracle.run_turb(F0, dims, 5, 1.9, "amd", 0.3, want_macro=True)
    want = oracle.turbulent_viscosity_amd(dims, macro[:, 1], macro[:, 2], macro[:, 3], 0.3)
    assert np.array_equal(nut, want)
    assert nut.max() > 0.0
    # the viscosity changes the flow: the run differs from the plain K17
    assert not np.array_equal(F, oracle.run(F0, dims, 5, 1.9))


# ------------------------------------------------ refinement interpolation --
def _nest_states():
    import virtualfluids_amd as vfx

    nest = vfx.NestedBox((16, 14, 14), (3, 3, 4), (11, 10, 11))
    rng = np.random.default_rng(21)
    nc, nf = int(np.prod(nest.coarse_dims)), int(np.prod(nest.fine_dims))
    Fc = oracle.init_field(*cases.random_flow(nest.coarse_dims, seed=1)) + rng.normal(0, 1e-3, (nc, 27)) * W
    Ff = oracle.init_field(*cases.random_flow(nest.fine_dims, seed=2)) + rng.normal(0, 1e-3, (nf, 27)) * W
    return nest, Fc, Ff, rng


def test_scaling_zero_and_constant_viscosity():
    """nu_t = 0 on both levels is the plain interpolation; a constant nu_t is the
    plain interpolation at the effective rates of both levels."""
    import virtualfluids_amd as vfx

    nest, Fc, Ff, _ = _nest_states()
    oc = 1.93
    of = vfx.fine_omega(oc)
    dc, df = nest.coarse_dims, nest.fine_dims
    zc, zf = np.zeros(len(Fc)), np.zeros(len(Ff))
    for nut_c, nut_f in [(0.0, 0.0), (0.013, 0.004)]:
        eff_c, eff_f = oc / (1.0 + 3.0 * oc * nut_c), of / (1.0 + 3.0 * of * nut_f)
        nut = (zc + nut_c, zf + nut_f)
        a = oracle.scale_cf(Fc, Ff, dc, df, nest.cf_coarse, nest.cf_fine, oc, of, nut=nut)
        b = oracle.scale_cf(Fc, Ff, dc, df, nest.cf_coarse, nest.cf_fine, eff_c, eff_f)
        assert np.array_equal(a, b)
        a = oracle.scale_fc(Fc, Ff, dc, df, nest.fc_coarse, nest.fc_fine, oc, of, nut=nut)
        b = oracle.scale_fc(Fc, Ff, dc, df, nest.fc_coarse, nest.fc_fine, eff_c, eff_f)
        assert np.array_equal(a, b)


def _corners(base, dims):
    nx, ny, nz = dims
    bx, by, bz = base % nx, (base // nx) % ny, base // (nx * ny)
    return [((bx + (c & 1)) % nx) + nx * (((by + ((c >> 1) & 1)) % ny) + ny * ((bz + ((c >> 2) & 1)) % nz))
            for c in range(8)]


def test_scaling_uses_each_nodes_viscosity():
    """Random nu_t per node: every source corner's moments use that corner's
    rate and every destination node its own (scaleCoarseToFineCompressible.cu:94,
    369, scaleFineToCoarseCompressible.cu:64) -- rebuilt cell by cell from the
    single-cell interpolation functions the golden vectors pin."""
    import virtualfluids_amd as vfx

    nest, Fc, Ff, rng = _nest_states()
    oc = 1.9
    of = vfx.fine_omega(oc)
    dc, df = nest.coarse_dims, nest.fine_dims
    nut_c = rng.uniform(0.0, 0.02, len(Fc))
    nut_f = rng.uniform(0.0, 0.02, len(Ff))
    rate = lambda om, nu: om / (1.0 + 3.0 * om * nu)  # noqa: E731
    got = oracle.scale_cf(Fc, Ff, dc, df, nest.cf_coarse, nest.cf_fine, oc, of, nut=(nut_c, nut_f))
    for i in range(0, len(nest.cf_coarse), 7):
        cc, fc = _corners(nest.cf_coarse[i], dc), _corners(nest.cf_fine[i], df)
        k = oracle.interp_coefficients_omegas(Fc[cc], [rate(oc, nut_c[j]) for j in cc])
        for c in range(8):
            x, y, z = (0.25 if c & 1 else -0.25), (0.25 if c & 2 else -0.25), (0.25 if c & 4 else -0.25)
            want = oracle.interpolate_cf(k, rate(of, nut_f[fc[c]]), x, y, z)
            assert np.array_equal(got[fc[c]], want), (i, c)
    got = oracle.scale_fc(Fc, Ff, dc, df, nest.fc_coarse, nest.fc_fine, oc, of, nut=(nut_c, nut_f))
    for i in range(0, len(nest.fc_coarse), 5):
        fc = _corners(nest.fc_fine[i], df)
        k = oracle.interp_coefficients_omegas(Ff[fc], [rate(of, nut_f[j]) for j in fc])
        want = oracle.interpolate_fc(k, rate(oc, nut_c[nest.fc_coarse[i]]))
        assert np.array_equal(got[nest.fc_coarse[i]], want), i


# ------------------------------------------- slip with turbulent viscosity --
def _slip_channel(kind):
    dims = (6, 12, 5)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.3, q_top=0.6)
    bcs = [oracle.BCGroup(nodes, q, kind=kind) for nodes, q in groups.values()]
    F0 = oracle.init_field(*cases.random_flow(dims, seed=14, u_amp=0.03))
    return dims, solid, bcs, F0


def test_slip_turbulent_viscosity_rate():
    """SlipTurbulentViscosityCompressible: the slip formula at the node's rate
    omega / (1 + 3 omega max(0, nu_t)).  With a fixed nu_t (AMD: the collision
    does not change it) one step equals the plain slip run at the effective
    rate; with nu_t = 0 (Smagorinsky constant 0) it is the plain slip run."""
    dims, solid, bcs8, F0 = _slip_channel(8)
    _, _, bcs3, _ = _slip_channel(3)
    om, c = 1.9, 0.012
    F, _ = oracle.run_turb(F0, dims, 1, om, "amd", 0.3, nut=np.full(len(F0), c), solid=solid, bcs=bcs8)
    want = oracle.run(F0, dims, 1, om / (1.0 + 3.0 * om * c), solid=solid, bcs=bcs3)
    assert np.array_equal(F[~solid], want[~solid])
    F, nut = oracle.run_turb(F0, dims, 20, om, "smagorinsky", 0.0, solid=solid, bcs=bcs8)
    assert np.array_equal(F[~solid], oracle.run(F0, dims, 20, om, solid=solid, bcs=bcs3)[~solid])
