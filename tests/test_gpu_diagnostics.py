"""GPU diagnostics: FP64 mass/momentum reduction, kinetic energy and enstrophy
(KineticEnergyAnalyzer / EnstrophyAnalyzer), and running velocity statistics
(CalcTurbulenceIntensity), against the oracle and closed forms."""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12  # FP64 sums in a different order: relative agreement


def close(a, b, tol=TOL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale <= tol


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_mass_momentum_matches_oracle(cuda, storage):
    dims = (40, 24, 18)
    fields = cases.random_flow(dims, seed=31, u_amp=0.04)
    F0 = oracle.init_field(*fields)
    lat = vfx.Lattice(dims, storage=storage, omega=1.7)
    lat.write_populations(F0)
    got = lat.mass_momentum(True).cpu().numpy()
    Fd = lat.read_populations(True).cpu().numpy()  # what the lattice actually stores
    ref = oracle.mass_momentum(Fd)
    assert close(got, ref, 1e-11)


def test_mass_is_conserved_over_steps(cuda):
    """Periodic box, K17 FP64: global mass and momentum stay constant to
    rounding over 200 steps (the FP64 conservation monitor)."""
    dims = (32, 32, 32)
    lat = vfx.Lattice(dims, storage="fp64", omega=1.9)
    lat.init_equilibrium(*cases.random_flow(dims, seed=2, u_amp=0.05))
    st = vfx.Stepper(lat, "k17")
    m0 = lat.mass_momentum(st.even).cpu().numpy()
    st.step(200)
    m1 = lat.mass_momentum(st.even).cpu().numpy()
    assert abs(m1[0] - m0[0]) < 1e-11
    assert np.abs(m1[1:] - m0[1:]).max() < 1e-11


def test_mass_error_of_fp16_storage_is_reported(cuda):
    """FP16 storage: the mass drift is what the quantisation makes it; the
    monitor measures it (bounded, non-zero in general)."""
    dims = (32, 32, 32)
    fields = cases.random_flow(dims, seed=2, u_amp=0.05)
    out = {}
    for storage in ("fp64", "fp16"):
        lat = vfx.Lattice(dims, storage=storage, omega=1.9)
        lat.init_equilibrium(*fields)
        st = vfx.Stepper(lat, "k17")
        st.step(100)
        out[storage] = lat.mass_momentum(st.even).cpu().numpy()
    n = np.prod(dims)
    assert abs(out["fp16"][0] - out["fp64"][0]) / n < 1e-5


@pytest.mark.parametrize("which", ["ke", "enstrophy"])
def test_energy_matches_oracle(cuda, which):
    dims = (30, 26, 20)
    fields = cases.random_flow(dims, seed=5, u_amp=0.03)
    F0 = oracle.init_field(*fields)
    # a small solid block exercises the stencil masks and the order reduction
    solid = np.zeros(dims[::-1], dtype=bool)
    solid[8:11, 10:13, 12:16] = True
    lat = vfx.Lattice(dims, storage="fp64", omega=1.7, solid=solid)
    lat.write_populations(F0)
    f = lat.kinetic_energy if which == "ke" else lat.enstrophy
    got = f(True).cpu().numpy()
    ref = oracle.energy(F0, dims, enstrophy=(which == "enstrophy"), solid=solid)
    assert got[1] == ref[1]
    assert close(got[0], ref[0])


def test_taylor_green_kinetic_energy_closed_form(cuda):
    """E0 = <1/2 rho u^2> = u0^2 / 8 (+ O(u0^4) from the density) for the TGV."""
    n, u0 = 48, 0.04
    lat = vfx.Lattice((n, n, n), storage="fp64", omega=1.9)
    lat.init_equilibrium(*cases.taylor_green(n, u0))
    s, c = lat.kinetic_energy(True).cpu().numpy()
    assert c == n ** 3
    assert abs(s / c / (u0 ** 2 / 8) - 1) < 1e-3


def test_statistics_mean_and_reynolds_stress(cuda):
    """Running sums equal the average of the per-step velocity fields."""
    dims = (16, 12, 10)
    lat = vfx.Lattice(dims, storage="fp32", omega=1.8)
    lat.init_equilibrium(*cases.random_flow(dims, seed=8, u_amp=0.05))
    st = vfx.Stepper(lat, "k17")
    stats = vfx.Statistics(lat)
    us = []
    for _ in range(12):
        st.step(1)
        stats.accumulate(st.even)
        _, ux, uy, uz = lat.macroscopic(st.even)
        us.append(torch.stack([ux, uy, uz]).cpu().numpy())
    U = np.stack(us)  # (T, 3, N)
    mean = U.mean(0)
    assert close(stats.mean().cpu().numpy(), mean, 1e-12)
    rs = stats.reynolds_stress().cpu().numpy()
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for c, (a, b) in enumerate(pairs):
        expect = (U[:, a] * U[:, b]).mean(0) - mean[a] * mean[b]
        np.testing.assert_allclose(rs[c], expect, rtol=0, atol=1e-13)
