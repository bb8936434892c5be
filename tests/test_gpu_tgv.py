"""Taylor-Green vortex on a 64^3 grid, K17 and BGK in FP64 and FP32 on the GPU.

* Viscous decay: u0 = 0.005 and Re = 10 keep the flow linear, so the kinetic
  energy decays as exp(-6 nu k^2 t) (k = 2 pi / n); the rate is fitted over
  steps 100..400 (after the initial non-equilibrium layer, over several
  acoustic periods) and must be within 1% of the analytic one.
* FP32 vs FP64: the Re = 100, u0 = 0.04 vortex, kinetic energy and enstrophy
  within 1e-4 relative over 300 steps.
The FP64 kernels themselves are bit-identical to the oracle (test_gpu_parity.py)."""
import math

import numpy as np
import pytest
import torch

import virtualfluids_amd as vfx
from virtualfluids_amd import geometry

pytestmark = pytest.mark.gpu


def _run(n, coll, storage, steps, every, u0=0.04, re=100.0):
    fields, omega, nu = geometry.taylor_green(n, u0=u0, re=re)
    lat = vfx.Lattice((n, n, n), storage=storage, omega=omega)
    lat.init_equilibrium(*fields)
    st = vfx.Stepper(lat, coll)
    ke, en = [], []
    for i in range(steps // every + 1):
        if i:
            st.step(every)
        ke.append(lat.kinetic_energy(st.even))
        en.append(lat.enstrophy(st.even))
    torch.cuda.synchronize()
    k = torch.stack(ke).cpu().numpy()
    e = torch.stack(en).cpu().numpy()
    return k[:, 0] / k[:, 1], e[:, 0] / e[:, 1], nu


@pytest.mark.parametrize("coll", ["k17", "bgk"])
def test_tgv_viscous_decay(cuda, coll):
    n, steps, every = 64, 400, 20
    print(f"[tgv viscous {coll} {n}^3 x {steps}]", flush=True)
    e64, _, nu = _run(n, coll, "fp64", steps, every, u0=0.005, re=10.0)
    t = np.arange(len(e64)) * every
    k = 2 * math.pi / n
    fit = t >= 100
    rate = -np.polyfit(t[fit], np.log(e64[fit]), 1)[0]
    err = rate / (6 * nu * k * k) - 1
    print(f"decay rate {rate:.4e} analytic {6 * nu * k * k:.4e} rel {err:+.3e}")
    # initial energy per node of the Taylor-Green field: u0^2 / 8 (rho = 1)
    assert abs(e64[0] / (0.005 ** 2 / 8) - 1) < 2e-3
    assert abs(err) < 0.01


@pytest.mark.parametrize("coll", ["k17", "bgk"])
def test_tgv_fp32_against_fp64(cuda, coll):
    n, steps, every = 64, 300, 25
    print(f"[tgv fp32 {coll} {n}^3 x {steps}]", flush=True)
    e64, z64, _ = _run(n, coll, "fp64", steps, every)
    e32, z32, _ = _run(n, coll, "fp32", steps, every)
    ke, en = np.abs(e32 / e64 - 1).max(), np.abs(z32 / z64 - 1).max()
    print(f"fp32 vs fp64: KE {ke:.2e} enstrophy {en:.2e}")
    assert ke < 1e-4 and en < 1e-4
