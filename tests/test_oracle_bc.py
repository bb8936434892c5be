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


def test_pressure_boundary_keeps_matching_equilibrium():
    """If the inward neighbour is at the equilibrium of the prescribed density
    the non-equilibrium pressure boundary reproduces that equilibrium, and it
    replaces exactly the directions with c . normal <= 0."""
    nx, ny, nz = 6, 5, 5
    solid, groups, onodes, onbrs = cases.inflow_channel(nx, ny, nz)
    N = nx * ny * nz
    rho_bc = 2e-3
    F0 = oracle.init_field(np.full(N, rho_bc), np.full(N, 0.01), np.zeros(N), np.zeros(N))
    bcs = [oracle.BCGroup(onodes, kind=4, rho=rho_bc, nbr=onbrs, normal=1),
           oracle.BCGroup(*groups["wall"], kind=1), oracle.BCGroup(*groups["inlet"], ux=0.01, kind=2)]
    F = oracle.run(F0, (nx, ny, nz), 1, OMEGA, solid=solid, bcs=bcs)
    # a single step from a uniform equilibrium: interior stays at equilibrium,
    # the outlet node's replaced populations equal the neighbour's equilibrium
    k = onodes[len(onodes) // 2]
    feq = oracle.equilibrium(rho_bc, 0.01, 0.0, 0.0)
    replaced = [d for d in range(27) if cases.CX[d] <= 0]
    np.testing.assert_allclose(F[k][replaced], feq[replaced], rtol=0, atol=1e-15)


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
