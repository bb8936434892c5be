"""Checkpoint and resume through VirtualFluids restart files (RestartObject.cpp).

* The exported array is VirtualFluids' own distribution array: for FP32/FP64
  storage the device array itself, for FP16 storage element (plane p, index j)
  is the decoded population of direction p (even parity) or inv(p) (odd) of
  the node that owns it -- checked against the node-major read-out.
* Stop after an odd number of steps, write a binary restart file, load it into
  a fresh lattice and continue: the stored populations equal those of the
  uninterrupted run bit for bit, in every storage form (FP16 with a reference
  velocity, so the odd-parity direction mapping of the offsets is exercised).
"""
import numpy as np
import pytest
import torch

import virtualfluids_amd as vfx
from tests import cases
from virtualfluids_amd import restart

pytestmark = pytest.mark.gpu

DIMS = (24, 18, 14)
CX, CY, CZ = cases.CX, cases.CY, cases.CZ
INV = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 26, 25, 24, 23, 22, 21, 20, 19]


def _lattice(storage):
    ref = (0.03, 0.01, 0.0) if storage == "fp16" else (0.0, 0.0, 0.0)
    lat = vfx.Lattice(DIMS, storage=storage, omega=1.9, device="cuda", reference_velocity=ref)
    lat.init_equilibrium(*cases.random_flow(DIMS, seed=3, u_amp=0.04))
    return lat


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_export_is_virtualfluids_array(cuda, storage):
    lat = _lattice(storage)
    st = vfx.Stepper(lat, "k17")
    st.step(3)  # odd parity
    even = st.even
    dt = torch.float64 if storage == "fp64" else torch.float32
    ex = lat.export_distributions(even, dtype=dt)
    if storage != "fp16":
        assert torch.equal(ex, lat.f)
    pop = lat.read_populations(even).to(dt).cpu().numpy()  # node-major, direction d of node k
    exn = ex.cpu().numpy()
    nx, ny, nz = DIMS
    N = nx * ny * nz
    k = np.arange(N)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
    for d in range(27):
        p = d if even else INV[d]
        # the element direction d of node k lives at: one step along each axis where c_d < 0
        j = ((x + (CX[d] < 0)) % nx) + nx * (((y + (CY[d] < 0)) % ny) + ny * ((z + (CZ[d] < 0)) % nz))
        assert np.array_equal(exn[p * N + j], pop[:, d]), d


@pytest.mark.parametrize("storage,real", [("fp64", np.float64), ("fp32", np.float32), ("fp16", np.float32)])
def test_resume_from_restart_file_is_bitwise(cuda, tmp_path, storage, real):
    a = _lattice(storage)
    sa = vfx.Stepper(a, "k17")
    sa.step(7)
    name = tmp_path / f"cp_{storage}"
    restart.save(name, [(a, sa.even)], "binary", real)
    saved_even = sa.even
    sa.step(5)
    b = vfx.Lattice(DIMS, storage=storage, omega=1.9, device="cuda", reference_velocity=a.reference_velocity)
    restart.load(name, [(b, saved_even)], "binary", real)
    sb = vfx.Stepper(b, "k17", even=saved_even)
    sb.step(5)
    torch.cuda.synchronize()
    print(f"[restart {storage}] resumed run identical: {torch.equal(a.f, b.f)}")
    assert sa.even == sb.even
    assert torch.equal(a.f, b.f)


def test_ascii_restart_rounds_to_six_digits(cuda, tmp_path):
    a = _lattice("fp64")
    vfx.Stepper(a, "k17").step(2)
    restart.save(tmp_path / "cp", [(a, True)], "ascii", np.float64)
    b = vfx.Lattice(DIMS, storage="fp64", omega=1.9, device="cuda")
    restart.load(tmp_path / "cp", [(b, True)], "ascii", np.float64)
    fa, fb = a.f.cpu().numpy(), b.f.cpu().numpy()
    rel = np.abs(fb - fa) / np.maximum(np.abs(fa), 1e-300)
    print(f"[ascii restart] max relative change {rel.max():.2e}")
    assert rel.max() <= 5e-6
