"""GPU parity of the K17 eddy-viscosity models (vf::lbm::TurbulenceModel:
Smagorinsky, AMD, QR) against the oracle.

* FP64, periodic box and walled cavity with bounce-back links: populations and
  the turbulent viscosity of every node after N steps within 1e-12 relative
  (expected exactly equal: reference operation order, no FMA contraction);
* the AMD model on VirtualFluids' sparse layout (typeOfGridNode, neighbor
  arrays, neighborInverse, permuted node order) equals the direct layout bit
  for bit, and so do the Smagorinsky and QR models;
* SGS constant 0 reproduces the plain K17 bit for bit (FP64);
* FP32 and scaled-FP16 storage stay close to the FP64 oracle (velocity L2
  relative error below 3e-5 / 1e-3, turbulent viscosity L2 relative error
  below 1e-3 / 3e-2).

Every flow starts from a seeded random field: the QR model divides by the
second strain invariant without a guard (as the reference does), so a node
with exactly zero strain -- a fluid at rest -- gets nu_t = 0/0.
"""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu

REL = 1e-12
MODELS = ["smagorinsky", "amd", "qr"]
SGS = {"smagorinsky": 0.17, "amd": 0.33, "qr": 0.3}


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("model", MODELS)
def test_periodic_fp64_matches_oracle(cuda, model):
    dims = (16, 12, 10)
    omega, steps = 1.93, 12
    print(f"[turbulence periodic {model}]", flush=True)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=21, u_amp=0.08))
    lat = vfx.Lattice(dims, storage="fp64", omega=omega, turbulence=model, sgs_constant=SGS[model])
    lat.write_populations(F0, even=True)
    st = vfx.Stepper(lat, "k17")
    st.step(steps)
    torch.cuda.synchronize()
    got = st.populations().cpu().numpy()
    nut = lat.nut.cpu().numpy()
    ref, rnut = oracle.run_turb(F0, dims, steps, omega, model, SGS[model])
    e_f, e_n = rel_err(got, ref), rel_err(nut, rnut)
    print(f"  populations {e_f:.2e} (exact {np.array_equal(got, ref)}), nu_t {e_n:.2e} "
          f"(exact {np.array_equal(nut, rnut)}), max nu_t {rnut.max():.3e}", flush=True)
    assert rnut.max() > 0.0
    assert e_f <= REL and e_n <= REL


def _cavity(n, omega, model, storage="fp64"):
    solid, groups = cases.cavity(n, lid=(0.1, 0.02, 0.0))
    m = n + 2
    lat = vfx.Lattice((m, m, m), storage=storage, omega=omega, solid=solid, turbulence=model,
                      sgs_constant=SGS[model])
    bcs = [vfx.BoundaryCondition(0, nodes, q, u[0], u[1], u[2]) for nodes, q, u in groups]
    obcs = [oracle.BCGroup(nodes, q, u[0], u[1], u[2]) for nodes, q, u in groups]
    return lat, bcs, obcs, solid, groups


@pytest.mark.parametrize("model", MODELS)
def test_cavity_fp64_matches_oracle(cuda, model):
    """Walls: the AMD differences turn one-sided next to them."""
    n, omega, steps = 10, 1.96, 40
    print(f"[turbulence cavity {model}]", flush=True)
    lat, bcs, obcs, solid, _ = _cavity(n, omega, model)
    m = n + 2
    fields = cases.random_flow((m, m, m), seed=31, u_amp=0.04)
    lat.init_equilibrium(*fields)
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    torch.cuda.synchronize()
    got = st.populations().cpu().numpy()
    nut = lat.nut.cpu().numpy()
    ref, rnut = oracle.run_turb(oracle.init_field(*fields), (m, m, m), steps, omega, model, SGS[model], solid=solid,
                                bcs=obcs)
    fl = ~solid
    e_f, e_n = rel_err(got[fl], ref[fl]), rel_err(nut[fl], rnut[fl])
    print(f"  populations {e_f:.2e}, nu_t {e_n:.2e}, max nu_t {rnut.max():.3e}", flush=True)
    assert rnut.max() > 0.0
    assert e_f <= REL and e_n <= REL


@pytest.mark.parametrize("model", MODELS)
def test_sparse_layout_equals_direct(cuda, model):
    """The cavity on VirtualFluids' sparse layout: solid nodes typed GEO_SOLID,
    fluid nodes GEO_FLUID, node order permuted, neighborInverse for the AMD
    differences."""
    n, omega, steps = 9, 1.95, 30
    print(f"[turbulence sparse {model}]", flush=True)
    a, bcs, _, solid, groups = _cavity(n, omega, model)
    m = n + 2
    dims = (m, m, m)
    nbx, nby, nbz, _, sidx = cases.sparse_layout(dims, seed=9)
    types = np.full(sidx.size + 1, 19, np.uint32)
    types[0] = 15
    types[sidx[solid]] = 15
    b = vfx.Lattice(storage="fp64", omega=omega, neighbors=(nbx, nby, nbz), node_types=types, turbulence=model,
                    sgs_constant=SGS[model], neighbor_inverse=cases.sparse_inverse(dims, sidx))
    bcs_b = [vfx.BoundaryCondition(0, sidx[nodes], q, u[0], u[1], u[2]) for nodes, q, u in groups]
    fields = cases.random_flow(dims, seed=32, u_amp=0.04)
    a.init_equilibrium(*fields)
    sparse_fields = []
    for f in fields:
        g = np.zeros(sidx.size + 1)
        g[sidx] = f
        sparse_fields.append(g)
    b.init_equilibrium(*sparse_fields)
    sa, sb = vfx.Stepper(a, "k17", bcs), vfx.Stepper(b, "k17", bcs_b)
    sa.step(steps)
    sb.step(steps)
    torch.cuda.synchronize()
    fl = ~solid
    pa = sa.populations().cpu().numpy()[fl]
    pb = sb.populations().cpu().numpy()[sidx][fl]
    na = a.nut.cpu().numpy()[fl]
    nb = b.nut.cpu().numpy()[sidx][fl]
    print(f"  identical populations {np.array_equal(pa, pb)}, nu_t {np.array_equal(na, nb)}", flush=True)
    assert na.max() > 0.0
    assert np.array_equal(pa, pb) and np.array_equal(na, nb)


def test_zero_constant_is_plain_k17(cuda):
    dims = (14, 12, 10)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=5, u_amp=0.06))
    a = vfx.Lattice(dims, storage="fp64", omega=1.9, turbulence="smagorinsky", sgs_constant=0.0)
    b = vfx.Lattice(dims, storage="fp64", omega=1.9)
    for lat in (a, b):
        lat.write_populations(F0, even=True)
        vfx.Stepper(lat, "k17").step(7)
    torch.cuda.synchronize()
    assert torch.equal(a.f, b.f)
    assert torch.count_nonzero(a.nut).item() == 0


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_reduced_precision_close_to_fp64(cuda, storage, model):
    n = 32
    dims = (n, n, n)
    omega, steps = 1.9, 100
    fields = cases.random_flow(dims, seed=33, u_amp=0.05)
    lat = vfx.Lattice(dims, storage=storage, omega=omega, turbulence=model, sgs_constant=SGS[model])
    lat.init_equilibrium(*fields)
    st = vfx.Stepper(lat, "k17")
    st.step(steps)
    r, ux, uy, uz = (t.cpu().numpy() for t in lat.macroscopic(st.even))
    nut = lat.nut.cpu().numpy().astype(np.float64)
    F0 = oracle.init_field(*fields)
    ref, rnut = oracle.run_turb(F0, dims, steps, omega, model, SGS[model])
    ro = np.array([oracle.macroscopic(f) for f in ref])
    du = np.sqrt(((ux - ro[:, 1]) ** 2 + (uy - ro[:, 2]) ** 2 + (uz - ro[:, 3]) ** 2).sum())
    l2 = du / np.sqrt((ro[:, 1] ** 2 + ro[:, 2] ** 2 + ro[:, 3] ** 2).sum())
    ln = np.sqrt(((nut - rnut) ** 2).sum() / (rnut ** 2).sum())
    print(f"[turbulence {storage} {model}] velocity L2 {l2:.3e}, nu_t L2 {ln:.3e}, max nu_t {rnut.max():.3e}",
          flush=True)
    assert l2 < (3e-5 if storage == "fp32" else 1e-3)
    assert ln < (1e-3 if storage == "fp32" else 3e-2)


# ------------------------------------------------ refinement interpolation --
NEST = ((16, 14, 14), (3, 3, 4), (11, 10, 11))


def _random_state(dims, seed):
    return oracle.init_field(*cases.random_flow(dims, seed=seed)) + \
        np.random.default_rng(seed).normal(0, 1e-3, (int(np.prod(dims)), 27)) * cases.W


def test_scale_with_turbulent_viscosity(cuda):
    """hasTurbulentViscosity: random nu_t on both levels, one FC and one CF
    call, every node vs the oracle."""
    print("[turbulent scaling]", flush=True)
    nest = vfx.NestedBox(*NEST)
    oc = 1.93
    of = vfx.fine_omega(oc)
    Fc = _random_state(nest.coarse_dims, 1)
    Ff = _random_state(nest.fine_dims, 2)
    rng = np.random.default_rng(3)
    nut_c = rng.uniform(0.0, 0.02, len(Fc))
    nut_f = rng.uniform(0.0, 0.02, len(Ff))
    C = vfx.Lattice(nest.coarse_dims, storage="fp64", omega=oc, turbulence="smagorinsky", sgs_constant=0.1)
    F = vfx.Lattice(nest.fine_dims, storage="fp64", omega=of, turbulence="smagorinsky", sgs_constant=0.1)
    C.nut.copy_(torch.from_numpy(nut_c))
    F.nut.copy_(torch.from_numpy(nut_f))
    C.write_populations(Fc, even=False)
    F.write_populations(Ff, even=True)
    vfx.scale_fc(C, F, vfx.ICells(nest.fc_coarse, nest.fc_fine, "cuda"), False, True)
    vfx.scale_cf(C, F, vfx.ICells(nest.cf_coarse, nest.cf_fine, "cuda"), False, True)
    torch.cuda.synchronize()
    got_c = C.read_populations(False).cpu().numpy()
    got_f = F.read_populations(True).cpu().numpy()
    nut = (nut_c, nut_f)
    ref_c = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, oc, of, nut=nut)
    ref_f = oracle.scale_cf(ref_c, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, oc, of,
                            nut=nut)
    plain_f = oracle.scale_cf(ref_c, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, oc, of)
    ec, ef = rel_err(got_c, ref_c), rel_err(got_f, ref_f)
    print(f"  coarse {ec:.2e} fine {ef:.2e}", flush=True)
    assert not np.array_equal(ref_f, plain_f)  # the viscosity matters
    assert ec <= REL and ef <= REL


@pytest.mark.parametrize("model", ["smagorinsky", "amd"])
def test_refined_steps_match_oracle(cuda, model):
    """RefinedStepper with a turbulence model on both levels vs the oracle's
    two-level run (two fine steps, one coarse step, FC, CF per coarse step)."""
    steps = 8
    print(f"[turbulent refined {model}]", flush=True)
    nest = vfx.NestedBox(*NEST)
    oc = 1.9
    of = vfx.fine_omega(oc)
    Fc = _random_state(nest.coarse_dims, 4)
    Ff = _random_state(nest.fine_dims, 5)
    st = vfx.RefinedStepper(nest, oc, storage="fp64", turbulence=model, sgs_constant=SGS[model])
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    st.step(steps)
    torch.cuda.synchronize()
    got_c = st.coarse.read_populations(st.even_c).cpu().numpy()
    got_f = st.fine.read_populations(st.even_f).cpu().numpy()
    nc, nf = np.zeros(len(Fc)), np.zeros(len(Ff))
    for _ in range(steps):
        Ff, nf = oracle.run_turb(Ff, nest.fine_dims, 2, of, model, SGS[model], nut=nf)
        Fc, nc = oracle.run_turb(Fc, nest.coarse_dims, 1, oc, model, SGS[model], nut=nc, solid=nest.coarse_state)
        Fc = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, oc, of,
                             nut=(nc, nf))
        Ff = oracle.scale_cf(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, oc, of,
                             nut=(nc, nf))
    act = nest.coarse_state == 0
    ec, ef = rel_err(got_c[act], Fc[act]), rel_err(got_f, Ff)
    en = rel_err(st.fine.nut.cpu().numpy(), nf)
    print(f"  coarse {ec:.2e} fine {ef:.2e} fine nu_t {en:.2e}", flush=True)
    assert nf.max() > 0.0
    assert ec <= REL and ef <= REL and en <= REL


# ------------------------------------------- slip with turbulent viscosity --
@pytest.mark.parametrize("model", ["smagorinsky", "qr"])
def test_slip_turbulent_viscosity_matches_oracle(cuda, model):
    """SlipTurbulentViscosityCompressible walls (q 0.3 / 0.6) with the model's
    viscosity, FP64 vs the oracle after 40 steps."""
    print(f"[slip turbulent viscosity {model}]", flush=True)
    dims = (6, 12, 5)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.3, q_top=0.6)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=15, u_amp=0.03))
    lat = vfx.Lattice(dims, storage="fp64", omega=1.9, solid=solid, turbulence=model, sgs_constant=SGS[model])
    lat.write_populations(F0)
    bcs = [vfx.BoundaryCondition(8, nodes, q) for nodes, q in groups.values()]
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(40)
    torch.cuda.synchronize()
    got = st.populations().cpu().numpy()
    obcs = [oracle.BCGroup(nodes, q, kind=8) for nodes, q in groups.values()]
    ref, rnut = oracle.run_turb(F0, dims, 40, 1.9, model, SGS[model], solid=solid, bcs=obcs)
    fl = ~solid
    e_f, e_n = rel_err(got[fl], ref[fl]), rel_err(lat.nut.cpu().numpy()[fl], rnut[fl])
    print(f"  populations {e_f:.2e}, nu_t {e_n:.2e}", flush=True)
    assert rnut[fl].max() > 0.0
    assert e_f <= REL and e_n <= REL


def test_slip_turbulent_viscosity_needs_a_model(cuda):
    dims = (6, 12, 5)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.3, q_top=0.6)
    lat = vfx.Lattice(dims, storage="fp64", omega=1.9, solid=solid)
    nodes, q = next(iter(groups.values()))
    with pytest.raises(vfx.VfxError, match="turbulence model"):
        lat.apply_bc(vfx.BoundaryCondition(8, nodes, q))
