"""GPU boundary conditions (libvfx) against the oracle, FP64, every population
after N steps within 1e-12 relative (observed: bit-exact), for every kind:
velocity bounce-back, interpolated no-slip / velocity / slip and the
non-equilibrium pressure outlet, on the structured and the sparse layout."""
import numpy as np
import pytest

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu
REL = 1e-12


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def run_both(dims, solid, groups_spec, steps, omega, F0, collision="k17", force=(0.0, 0.0, 0.0)):
    """groups_spec: list of dicts with kind, nodes, q, and optional ux/uy/uz,
    rho, nbr, normal.  Returns (gpu populations, oracle populations)."""
    lat = vfx.Lattice(dims, storage="fp64", omega=omega, solid=solid, force=force)
    lat.write_populations(F0)
    bcs, obcs = [], []
    for g in groups_spec:
        bcs.append(vfx.BoundaryCondition(g["kind"], g["nodes"], g.get("q"), g.get("ux"), g.get("uy"), g.get("uz"),
                                         rho=g.get("rho"), neighbors=g.get("nbr"), normal=g.get("normal", 0)))
        obcs.append(oracle.BCGroup(g["nodes"], g.get("q"), g.get("ux"), g.get("uy"), g.get("uz"), kind=g["kind"],
                                   rho=g.get("rho"), nbr=g.get("nbr"), normal=g.get("normal", 0)))
    st = vfx.Stepper(lat, collision, bcs)
    st.step(steps)
    got = st.populations().cpu().numpy()
    ref = oracle.run(F0, dims, steps, omega, collision=collision, solid=solid, bcs=obcs, force=force)
    return got, ref


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("collision", ["k17", "bgk"])
def test_interpolated_plates(cuda, kind, collision):
    dims = (6, 12, 5)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.3, q_top=0.65)
    N = np.prod(dims)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=kind, u_amp=0.02))
    spec = []
    for name, (nodes, q) in groups.items():
        g = {"kind": kind, "nodes": nodes, "q": q}
        if kind == 2:
            g["ux"] = 0.01 if name == "top" else 0.0
            g["uz"] = 0.003
        spec.append(g)
    got, ref = run_both(dims, solid, spec, 40, 1.85, F0, collision, force=(1e-6, 0, 0))
    fl = ~solid
    err = rel_err(got[fl], ref[fl])
    print(f"kind {kind} {collision}: max rel err {err:.2e}")
    assert err <= REL


def test_inflow_pressure_channel(cuda):
    dims = (20, 9, 8)
    solid, groups, onodes, onbrs = cases.inflow_channel(*dims)
    N = np.prod(dims)
    rng = np.random.default_rng(4)
    F0 = oracle.init_field(np.zeros(N), np.full(N, 0.02), np.zeros(N), np.zeros(N))
    inlet_nodes, inlet_q = groups["inlet"]
    spec = [{"kind": 4, "nodes": onodes, "rho": 1e-3 * rng.uniform(-1, 1, onodes.size), "nbr": onbrs, "normal": 1},
            {"kind": 1, "nodes": groups["wall"][0], "q": groups["wall"][1]},
            {"kind": 2, "nodes": inlet_nodes, "q": inlet_q, "ux": 0.02 * (1 + 0.01 * rng.uniform(-1, 1, inlet_nodes.size))}]
    got, ref = run_both(dims, solid, spec, 60, 1.9, F0)
    fl = ~solid
    assert rel_err(got[fl], ref[fl]) <= REL


def test_velocity_bounce_back_gpu_formula(cuda):
    """Moving-wall bounce-back with the GPU wall term (FP64 6 w c.u)."""
    dims = (8, 10, 6)
    solid, groups = cases.plate_channel(*dims)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=9, u_amp=0.02))
    spec = [{"kind": 0, "nodes": n, "q": q, "ux": (0.02 if name == "top" else 0.0), "uy": 0.0, "uz": 0.004}
            for name, (n, q) in groups.items()]
    got, ref = run_both(dims, solid, spec, 30, 1.8, F0)
    assert rel_err(got[~solid], ref[~solid]) <= REL


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_reduced_precision_channel_runs_close(cuda, storage):
    """FP32/FP16 forms with interpolated walls stay close to FP64."""
    dims = (16, 12, 8)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.4, q_top=0.6)
    fields = cases.random_flow(dims, seed=3, u_amp=0.03)
    out = {}
    for s in ("fp64", storage):
        lat = vfx.Lattice(dims, storage=s, omega=1.9, solid=solid)
        lat.init_equilibrium(*fields)
        bcs = [vfx.BoundaryCondition(1, n, q) for n, q in groups.values()]
        st = vfx.Stepper(lat, "k17", bcs)
        st.step(100)
        _, ux, uy, uz = (t.cpu().numpy() for t in lat.macroscopic(st.even))
        out[s] = np.stack([ux, uy, uz])[:, ~solid]
    l2 = np.sqrt(((out[storage] - out["fp64"]) ** 2).sum() / (out["fp64"] ** 2).sum())
    print(f"{storage} channel velocity L2 {l2:.2e}")
    assert l2 < (1e-4 if storage == "fp32" else 2e-3)


def test_sparse_layout_with_boundaries(cuda):
    """The sparse (neighbour-array) addressing with interpolated walls equals the oracle."""
    dims = (6, 10, 5)
    solid, groups = cases.plate_channel(*dims, q_bottom=0.35, q_top=0.5)
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, solid=solid, seed=3)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=12, u_amp=0.02))
    Fs = np.zeros((sidx.size + 1, 27))
    Fs[sidx] = F0
    lat = vfx.Lattice(storage="fp64", omega=1.85, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
    lat.write_populations(Fs)
    bcs = [vfx.BoundaryCondition(1, sidx[n], q) for n, q in groups.values()]
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(25)
    got = st.populations().cpu().numpy()[sidx]
    obcs = [oracle.BCGroup(n, q, kind=1) for n, q in groups.values()]
    ref = oracle.run(F0, dims, 25, 1.85, solid=solid, bcs=obcs)
    assert rel_err(got[~solid], ref[~solid]) <= REL


# ------------------------------------------------------------- wall models --
def _wall_channel():
    """Periodic in x and y, solid layers at z = 0 and z = nz-1; the wall model on
    the links into the floor (q = 0.5, sampling node 2 above), bounce-back on
    the lid links."""
    nx, ny, nz = 6, 5, 12
    dims = (nx, ny, nz)
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = solid[-1] = True
    fz, fy, fx = np.nonzero(~solid)
    floor, lid = {}, {}
    for z, y, x in zip(fz, fy, fx):
        k = x + nx * (y + ny * z)
        for d in range(1, 27):
            tz = z + cases.CZ[d]
            if tz == 0:
                floor.setdefault(k, np.full(27, -1.0, np.float32))[d] = 0.5
            elif tz == nz - 1:
                lid.setdefault(k, np.full(27, -1.0, np.float32))[d] = 0.5
    fn = np.array(sorted(floor), np.int64)
    ln = np.array(sorted(lid), np.int64)
    fq = np.stack([floor[k] for k in fn], axis=1)
    lq = np.stack([lid[k] for k in ln], axis=1)
    sample = fn + 2 * nx * ny
    return dims, solid.ravel(), (fn, fq, sample), (ln, lq)


@pytest.mark.parametrize("kind", [9, 10, 11])
def test_wall_model_channel_matches_oracle(cuda, kind):
    print(f"[wall model kind {kind}]", flush=True)
    dims, solid, (fn, fq, sample), (ln, lq) = _wall_channel()
    N = int(np.prod(dims))
    drho, ux, uy, uz = cases.random_flow(dims, seed=kind, u_amp=0.01)
    F0 = oracle.init_field(drho, ux + 0.04, uy, uz)
    force = (2e-6, 0.0, 0.0)
    normal = np.array([0.0, 0.0, 1.0])[:, None]
    ex0 = np.array([0.04, 0.0, 0.0])[:, None]
    lat = vfx.Lattice(dims, storage="fp64", omega=1.9, solid=solid, force=force)
    lat.write_populations(F0)
    wall = vfx.BoundaryCondition(kind, fn, fq, neighbors=sample, wall_normal=normal, sampling_offset=2, z0=0.1,
                                 exchange_velocity=ex0, wall_velocity=ex0)
    st = vfx.Stepper(lat, "k17", [wall, vfx.BoundaryCondition(0, ln, lq)])
    st.step(40)
    got = st.populations().cpu().numpy()
    og = oracle.BCGroup(fn, fq, kind=kind, nbr=sample, wall_normal=normal, sampling_offset=2, z0=0.1,
                        exchange_velocity=ex0, wall_velocity=ex0)
    ref = oracle.run(F0, dims, 40, 1.9, solid=solid, bcs=[og, oracle.BCGroup(ln, lq)], force=force)
    fl = ~solid
    errs = [rel_err(got[fl], ref[fl]),
            rel_err(wall.wm["exchange"].cpu().numpy(), og.exchange_velocity),
            rel_err(wall.wm["wall"].cpu().numpy(), og.wall_velocity),
            rel_err(wall.wm["u_star"].cpu().numpy(), og.u_star),
            rel_err(wall.wm["force"].cpu().numpy(), og.force)]
    print(f"  populations / exchange / wall / u* / force: {' '.join(f'{e:.1e}' for e in errs)}", flush=True)
    assert og.u_star.max() > 0
    assert max(errs) <= REL
