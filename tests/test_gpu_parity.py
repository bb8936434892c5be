"""GPU parity: the gfx950 stepper (libvfx through its C ABI) against the oracle.

FP64 bar: every population after N steps equal to the oracle within 1e-12
relative to the largest population magnitude (the kernels keep the
reference's operation order and the FP64 unit is built without FMA
contraction, so the observed difference is expected to be exactly 0).
"""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu

REL = 1e-12


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def structured(dims, storage="fp64", omega=1.6, solid=None, **kw):
    return vfx.Lattice(dims, storage=storage, omega=omega, solid=solid, **kw)


@pytest.mark.parametrize("collision", ["k17", "bgk"])
@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 10, 7)])
def test_periodic_fp64_matches_oracle(cuda, collision, dims):
    print(f"[periodic {collision} {dims}]", flush=True)
    omega = 1.7
    fields = cases.random_flow(dims, seed=11)
    F0 = oracle.init_field(*fields)
    lat = structured(dims, omega=omega)
    lat.write_populations(F0, even=True)
    st = vfx.Stepper(lat, collision)
    st.step(25)
    torch.cuda.synchronize()
    got = st.populations().cpu().numpy()
    ref = oracle.run(F0, dims, 25, omega, collision=collision)
    err = rel_err(got, ref)
    print(f"max rel err {err:.3e}  exact={np.array_equal(got, ref)}")
    assert err <= REL


def test_init_equilibrium_matches_oracle(cuda):
    dims = (20, 12, 9)
    fields = cases.random_flow(dims, seed=3)
    for even in (True, False):
        lat = structured(dims)
        lat.init_equilibrium(*fields, even=even)
        got = lat.read_populations(even).cpu().numpy()
        ref = oracle.init_field(*fields)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("even", [True, False])
def test_esoteric_twist_array_layout(cuda, even):
    """Raw array layout of VirtualFluids' pointer map
    (KernelUtilities.h:66-125, pinned by KernelUtilitiesTests.cpp:44-120): in an
    even step direction d of node k sits in plane d, in an odd step in the plane
    of the opposite direction, at node k shifted one step along every axis
    where c_d < 0."""
    dims = (6, 5, 4)
    nx, ny, nz = dims
    N = nx * ny * nz
    rng = np.random.default_rng(1)
    F = rng.normal(size=(N, 27))
    lat = structured(dims)
    lat.write_populations(F, even=even)
    raw = lat.f.cpu().numpy().reshape(27, N)
    for k in range(N):
        x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
        for d in range(27):
            xs = (x + (1 if cases.CX[d] < 0 else 0)) % nx
            ys = (y + (1 if cases.CY[d] < 0 else 0)) % ny
            zs = (z + (1 if cases.CZ[d] < 0 else 0)) % nz
            plane = d if even else cases.INV[d]
            assert raw[plane, xs + nx * (ys + ny * zs)] == F[k, d]


def test_sparse_layout_matches_oracle(cuda):
    """VirtualFluids sparse addressing (neighbour arrays, permuted node order,
    dummy node 0) gives the same populations as the oracle."""
    dims = (12, 10, 8)
    omega = 1.62
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, seed=5)
    fields = cases.random_flow(dims, seed=7)
    F0 = oracle.init_field(*fields)
    Fs = np.zeros((sidx.size + 1, 27))
    Fs[sidx] = F0
    lat = vfx.Lattice(storage="fp64", omega=omega, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
    lat.write_populations(Fs, even=True)
    st = vfx.Stepper(lat, "k17")
    st.step(12)
    got = st.populations().cpu().numpy()[sidx]
    ref = oracle.run(F0, dims, 12, omega)
    assert rel_err(got, ref) <= REL


@pytest.mark.parametrize("collision", ["bgk", "k17"])
def test_sparse_node_types_select_like_fluid_list(cuda, collision):
    """VirtualFluids' typeOfGridNode instead of the fluid-node list (the way
    B92CompressibleNavierStokes_Device selects its nodes): every node that is
    neither GEO_SOLID (15) nor GEO_VOID (16) collides -- fluid (19) and
    boundary-tagged nodes alike -- and the result equals the fluid-list run and
    the oracle (periodic box, permuted order, dummy node 0 typed solid)."""
    dims = (12, 10, 8)
    omega = 1.62 if collision == "k17" else 1.3
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, seed=5)
    types = np.full(sidx.size + 1, 19, np.uint32)
    types[0] = 15
    rng = np.random.default_rng(2)
    types[rng.choice(np.arange(1, sidx.size + 1), 40, replace=False)] = 7  # a non-fluid, non-solid tag
    fields = cases.random_flow(dims, seed=7)
    F0 = oracle.init_field(*fields)
    Fs = np.zeros((sidx.size + 1, 27))
    Fs[sidx] = F0
    a = vfx.Lattice(storage="fp64", omega=omega, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
    b = vfx.Lattice(storage="fp64", omega=omega, neighbors=(nbx, nby, nbz), node_types=types)
    assert b.n_fluid == sidx.size
    for lat in (a, b):
        lat.write_populations(Fs, even=True)
        vfx.Stepper(lat, collision).step(9)
    torch.cuda.synchronize()
    assert torch.equal(a.f, b.f)
    got = b.read_populations(False).cpu().numpy()[sidx]
    ref = oracle.run(F0, dims, 9, omega, collision=collision)
    assert rel_err(got, ref) <= REL


def _cavity_lattice(n, omega, storage="fp64", cpu_float=True):
    solid, groups = cases.cavity(n)
    m = n + 2
    lat = structured((m, m, m), storage=storage, omega=omega, solid=solid)
    bcs = [vfx.BoundaryCondition(0, nodes, q, u[0], u[1], u[2], cpu_float_velocity=cpu_float)
           for nodes, q, u in groups]
    obcs = [oracle.BCGroup(nodes, q, u[0], u[1], u[2], cpu_float_velocity=cpu_float) for nodes, q, u in groups]
    return lat, bcs, obcs, solid


@pytest.mark.parametrize("cpu_float", [True, False])
def test_cavity_fp64_short(cuda, cpu_float):
    n, omega, steps = 12, 1.6778523, 60
    print(f"[cavity n={n} cpu_float={cpu_float}]", flush=True)
    lat, bcs, obcs, solid = _cavity_lattice(n, omega, cpu_float=cpu_float)
    m = n + 2
    lat.init_equilibrium(np.zeros(m ** 3), np.zeros(m ** 3), np.zeros(m ** 3), np.zeros(m ** 3))
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    got = st.populations().cpu().numpy()
    ref = oracle.run(np.zeros((m ** 3, 27)), (m, m, m), steps, omega, solid=solid, bcs=obcs)
    fl = ~solid
    assert rel_err(got[fl], ref[fl]) <= REL


def test_cavity_64_1000_steps(cuda):
    """Lid-driven cavity, 64^3 fluid nodes, Re = 100 on the lid length, 1000
    steps, CPU-rounded wall velocity: every population vs the oracle."""
    n, steps = 64, 1000
    nu = 0.05 * n / 100.0
    omega = 1.0 / (3.0 * nu + 0.5)
    print(f"[cavity 64^3 x {steps}]", flush=True)
    lat, bcs, obcs, solid = _cavity_lattice(n, omega, cpu_float=True)
    m = n + 2
    z = np.zeros(m ** 3)
    lat.init_equilibrium(z, z, z, z)
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    got = st.populations().cpu().numpy()
    ref = oracle.run(np.zeros((m ** 3, 27)), (m, m, m), steps, omega, solid=solid, bcs=obcs)
    fl = ~solid
    err = rel_err(got[fl], ref[fl])
    print(f"cavity 64^3 max rel err {err:.3e} exact={np.array_equal(got[fl], ref[fl])}")
    assert err <= REL


@pytest.mark.parametrize("storage,variant", [("fp32", "default"), ("fp16", "default"), ("fp32", "plain"),
                                             ("fp16", "plain")])
def test_reduced_precision_close_to_fp64(cuda, storage, variant):
    """FP32 arithmetic on FP32 / scaled-FP16 storage stays close to the FP64
    oracle on a smooth flow (velocity L2 relative error), for the reduced-operation
    K17 ("default") and the reference-order one ("plain")."""
    dims = (32, 32, 32)
    omega = 1.8
    fields = cases.taylor_green(32, u0=0.04)
    lat = structured(dims, storage=storage, omega=omega, variant=variant)
    lat.init_equilibrium(*fields)
    st = vfx.Stepper(lat, "k17")
    st.step(200)
    r, ux, uy, uz = (t.cpu().numpy() for t in lat.macroscopic(st.even))
    F0 = oracle.init_field(*fields)
    ref, mac = oracle.run(F0, dims, 200, omega, want_macro=True)
    ro = np.array([oracle.macroscopic(f) for f in ref])
    du = np.sqrt(((ux - ro[:, 1]) ** 2 + (uy - ro[:, 2]) ** 2 + (uz - ro[:, 3]) ** 2).sum())
    nu_ = np.sqrt((ro[:, 1] ** 2 + ro[:, 2] ** 2 + ro[:, 3] ** 2).sum())
    l2 = du / nu_
    print(f"{storage}/{variant}: velocity L2 rel err {l2:.3e}")
    assert l2 < (3e-5 if storage == "fp32" else 1e-3)


def test_invalid_arguments_raise(cuda):
    lat = structured((8, 8, 8))
    lat.omega = 2.5
    with pytest.raises(vfx.VfxError):
        lat.collide_stream("k17", True)


def test_fast_and_plain_fp32_agree_closely(cuda):
    """Reduced-operation and reference-order FP32 K17 differ only by rounding."""
    dims = (20, 16, 12)
    fields = cases.random_flow(dims, seed=21)
    out = {}
    for variant in ("default", "plain"):
        lat = structured(dims, storage="fp32", omega=1.9, variant=variant)
        lat.init_equilibrium(*fields)
        st = vfx.Stepper(lat, "k17")
        st.step(30)
        out[variant] = st.populations().cpu().numpy()
    err = rel_err(out["default"], out["plain"])
    print(f"fast vs plain FP32 max rel err {err:.3e}")
    assert err < 1e-5


def test_body_force_template_matches_oracle(cuda):
    """ApplyBodyForce with a random per-node body force: one step equals the
    oracle bit for bit (FP64), the body force is 0 afterwards, and the next
    steps run on the global force alone; FP32 / FP16 stay close."""
    dims = (16, 12, 10)
    n = int(np.prod(dims))
    F0 = oracle.init_field(*cases.random_flow(dims, seed=19))
    rng = np.random.default_rng(20)
    bf = rng.uniform(-5e-5, 5e-5, (3, n))
    g, ff = (1e-5, 0.0, -2e-6), 0.5
    ref, rbf = oracle.run_body_force(F0, dims, 4, 1.85, g, ff, bf)
    for storage in ("fp64", "fp32", "fp16"):
        lat = structured(dims, storage=storage, omega=1.85)
        lat.write_populations(F0, even=True)
        lat.set_body_force(*bf, forces=g, force_factor=ff)
        st = vfx.Stepper(lat, "k17", apply_body_force=True)
        st.step(4)
        got = st.populations().cpu().numpy()
        err = rel_err(got, ref)
        print(f"[body force {storage}] rel err {err:.2e}", flush=True)
        assert not lat.body_force.any()
        assert err <= {"fp64": REL, "fp32": 2e-5, "fp16": 2e-2}[storage]
