"""Taylor-Green vortex on a 64^3 grid (Re = 100, so the early decay is
viscous within the 300 steps): K17 and BGK in FP64 and FP32 on the GPU.  Early-time kinetic-energy decay against the
analytic rate 6 nu k^2 (k = 2 pi / n) and FP32 against FP64 energy and
enstrophy; the FP64 kernels themselves are bit-identical to the oracle
(test_gpu_parity.py)."""
import math

import numpy as np
import pytest
import torch

import virtualfluids_amd as vfx
from virtualfluids_amd import geometry

pytestmark = pytest.mark.gpu


def _run(n, coll, storage, steps, every):
    fields, omega, nu = geometry.taylor_green(n, re=100.0)
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
def test_tgv_decay_and_fp32(cuda, coll):
    n, steps, every = 64, 300, 25
    print(f"[tgv {coll} {n}^3 x {steps}]", flush=True)
    e64, z64, nu = _run(n, coll, "fp64", steps, every)
    e32, z32, _ = _run(n, coll, "fp32", steps, every)
    t = np.arange(len(e64)) * every
    k = 2 * math.pi / n
    rate = -np.polyfit(t, np.log(e64), 1)[0]
    err = rate / (6 * nu * k * k) - 1
    print(f"decay rate {rate:.4e} analytic {6 * nu * k * k:.4e} rel {err:+.3e}")
    print(f"fp32 vs fp64: KE {np.abs(e32 / e64 - 1).max():.2e} enstrophy {np.abs(z32 / z64 - 1).max():.2e}")
    # initial energy per node of the Taylor-Green field: u0^2 / 8 (rho = 1)
    assert abs(e64[0] / (0.04 ** 2 / 8) - 1) < 2e-3
    assert abs(err) < 0.03
    assert np.abs(e32 / e64 - 1).max() < 1e-5
    assert np.abs(z32 / z64 - 1).max() < 1e-4
