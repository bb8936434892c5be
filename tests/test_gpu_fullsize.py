"""The headline workload at its full size (512 x 256 x 256, 33.5 M nodes) and the
full 512-cavity ensemble, checked through properties that need no oracle run
(the CPU oracle cannot step these sizes in seconds):

* translation invariance: the flow shifted by one node along x, y or z gives
  exactly the shifted macroscopic fields after 20 steps, in every storage form
  (a node's arithmetic does not depend on where it sits, so any mistake in the
  row/plane wrap-around, the block-to-XCD mapping or the esoteric-twist offsets
  at full size breaks the equality);
* FP64: global mass and momentum are conserved to rounding;
* FP32 and FP16 storage against FP64 at full size: velocity L2 error after 100
  steps within 1e-4 (FP32) and 1e-3 (FP16, the accuracy bar of the FP16 form);
* the 512-cavity ensemble: one batched launch per step equals one launch per
  domain bit for bit.
"""
import numpy as np
import pytest
import torch

import bench
import virtualfluids_amd as vfx

pytestmark = pytest.mark.gpu

DIMS = bench.DIMS
OMEGA = bench.OMEGA


def _flow(shift=(0, 0, 0)):
    """bench's smooth shear flow plus seeded noise (breaks the flow's own
    symmetries), rolled by `shift` nodes."""
    nx, ny, nz = DIMS
    fields = bench.initial_flow(torch, DIMS, "cuda")
    g = torch.Generator(device="cuda").manual_seed(11)
    out = []
    for f in fields:
        f = f + 1e-3 * (torch.rand(f.numel(), generator=g, device="cuda", dtype=torch.float64) - 0.5)
        f = torch.roll(f.view(nz, ny, nx), shifts=(shift[2], shift[1], shift[0]), dims=(0, 1, 2))
        out.append(f.contiguous().view(-1))
    return out


def _run(storage, steps, shift=(0, 0, 0)):
    lat = vfx.Lattice(DIMS, storage=storage, omega=OMEGA, device="cuda")
    lat.init_equilibrium(*_flow(shift), even=True)
    st = vfx.Stepper(lat, "k17")
    st.step(steps)
    torch.cuda.synchronize()
    return lat, st


def _roll(t, shift):
    nx, ny, nz = DIMS
    return torch.roll(t.view(nz, ny, nx), shifts=(shift[2], shift[1], shift[0]), dims=(0, 1, 2)).reshape(-1)


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_translation_invariance_full_size(cuda, storage):
    print(f"[full-size translation {storage}]", flush=True)
    lat, st = _run(storage, 20)
    ref = [m.clone() for m in lat.macroscopic(st.even)]
    del lat, st
    torch.cuda.empty_cache()
    for shift in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (257, 3, 129)]:
        lat, st = _run(storage, 20, shift)
        got = lat.macroscopic(st.even)
        same = all(torch.equal(g, _roll(r, shift)) for g, r in zip(got, ref))
        print(f"  shift {shift}: identical {same}", flush=True)
        assert same, shift
        del lat, st, got
        torch.cuda.empty_cache()


def test_fp64_conservation_full_size(cuda):
    lat = vfx.Lattice(DIMS, storage="fp64", omega=OMEGA, device="cuda")
    lat.init_equilibrium(*_flow(), even=True)
    st = vfx.Stepper(lat, "k17")
    m0 = lat.mass_momentum(st.even).cpu().numpy()
    st.step(20)
    m1 = lat.mass_momentum(st.even).cpu().numpy()
    print(f"[full-size conservation] mass {m0[0]:.6e} -> {m1[0]:.6e}, momentum change "
          f"{np.abs(m1[1:] - m0[1:]).max():.3e}", flush=True)
    # rounding of 33.5 M per-node sums over 20 steps stays below 1e-8; one lost
    # population per node would move the sums by ~1e3
    assert abs(m1[0] - m0[0]) < 1e-8
    assert np.abs(m1[1:] - m0[1:]).max() < 1e-8


def test_reduced_precision_full_size(cuda):
    steps = 100
    res = {}
    for storage in ("fp64", "fp32", "fp16"):
        lat, st = _run(storage, steps)
        res[storage] = [m.clone() for m in lat.macroscopic(st.even)]
        del lat, st
        torch.cuda.empty_cache()
    ref = res["fp64"]
    for storage, bar in (("fp32", 1e-4), ("fp16", 1e-3)):
        l2 = bench.velocity_l2(res[storage], ref)
        drift = bench.density_drift(res[storage], ref)
        print(f"[full-size {storage}] velocity L2 {l2:.3e}, density drift {drift:.3e}", flush=True)
        assert l2 <= bar


def test_ensemble_batched_equals_per_domain_full_size(cuda):
    dims, omegas, lids = vfx.cavity_ensemble(512, seed=7)
    boxes = [(a + 2, b + 2, c + 2) for a, b, c in dims]
    a = vfx.Ensemble(boxes, omegas, lids, storage="fp16", device="cuda")
    a.init_rest()
    a.step(4)
    pool_a = a.pool.clone()
    del a
    torch.cuda.empty_cache()
    b = vfx.Ensemble(boxes, omegas, lids, storage="fp16", device="cuda")
    b.init_rest()
    b.step_per_domain(4)
    torch.cuda.synchronize()
    same = torch.equal(pool_a, b.pool)
    print(f"[full ensemble] {b.fluid_nodes} interior nodes, batched == per-domain: {same}", flush=True)
    assert same
