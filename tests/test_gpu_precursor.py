"""GPU parity of the precursor inflows against the oracle: a channel with the
precursor inlet on its first fluid layer (VFX_BC_PRECURSOR_VELOCITY on the
inlet links, PrecursorNonReflectiveCompressible; or
VFX_BC_PRECURSOR_DISTRIBUTIONS, PrecursorDistributions), interpolated no-slip
walls and the non-reflecting outflow on the last fluid layer.

FP64: every fluid population after 60 steps within 1e-12 relative of the
oracle and identical to it bit for bit; sparse layout equal to the direct one
bit for bit; plane cycling through PrecursorPlanes' reader (both kinds) equal
to the oracle stepped with the closed-form schedule (last = plane(mT), current
= plane((m+1)T), ratio (t - mT)/T), bit for bit; FP32 / FP16 storage close to
FP64 (velocity L2 below 1e-4 / 1e-3).  Stencils
mix random weights with first weights >= 1e6 (the 0PP node alone).
"""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu
REL = 1e-12
DIMS = (20, 9, 8)
N_PRE = 40
U = 0.03
PDIRS = [1, 7, 9, 11, 13, 19, 21, 23, 25]


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _setup(kind, seed=3):
    solid, groups, onodes, onbrs = cases.inflow_channel(*DIMS)
    inlet, iq = groups["inlet"]
    n = inlet.size
    rng = np.random.default_rng(seed)
    nb = rng.integers(0, N_PRE, (4, n))
    w = rng.uniform(0.1, 2.0, (4, n))
    w[0, :5] = 2e6
    if kind == 12:
        planes = [rng.normal(0, 2e-3, (3, N_PRE)) for _ in range(2)]
        ux = 1.25 * U * (1 + 0.01 * rng.uniform(-1, 1, n))
    else:
        base = oracle.init_field(np.zeros(1), np.full(1, U), np.zeros(1), np.zeros(1))[0][PDIRS]
        planes = [base[:, None] + rng.normal(0, 1e-4, (9, N_PRE)) for _ in range(2)]
        ux = None
    drho, vx, vy, vz = cases.random_flow(DIMS, seed=9, u_amp=0.01)
    F0 = oracle.init_field(drho, vx + U, vy, vz)
    # solid nodes hold zero populations, as a fresh lattice does there
    # (write_populations leaves them alone); the velocity inlet's second state
    # reads them
    F0[solid] = 0.0
    return dict(kind=kind, solid=solid, wall=groups["wall"], inlet=inlet, iq=iq, onodes=onodes, onbrs=onbrs, nb=nb,
                w=w, planes=planes, ux=ux, F0=F0)


def _oracle_bcs(c, tr, planes=None):
    last, cur = planes or c["planes"]
    pre = dict(neighbors=c["nb"], weights=c["w"], last=last, current=cur, time_ratio=tr, velocity_ratio=1.25)
    return [oracle.BCGroup(c["onodes"], kind=6, nbr=c["onbrs"], normal=1),
            oracle.BCGroup(*c["wall"], kind=1),
            oracle.BCGroup(c["inlet"], c["iq"] if c["kind"] == 12 else None, c["ux"], kind=c["kind"], precursor=pre)]


def _gpu_run(c, storage, steps, tr=0.35, index=None, layout=None, reader=None, T=3, n_pre=N_PRE):
    real = torch.float64 if storage == "fp64" else torch.float32
    m = (lambda a: a) if index is None else (lambda a: index[a])
    comps = 3 if c["kind"] == 12 else 9
    if index is None:
        lat = vfx.Lattice(DIMS, storage=storage, omega=1.9, solid=c["solid"])
        lat.write_populations(c["F0"])
    else:
        nbx, nby, nbz, fluid = layout
        lat = vfx.Lattice(storage=storage, omega=1.9, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
        Fs = np.zeros((index.size + 1, 27))
        Fs[index] = c["F0"]
        lat.write_populations(Fs)
    pc = vfx.PrecursorPlanes(c["nb"], c["w"], n_pre, comps, reader=reader, time_steps_between_reads=T,
                             velocity_ratio=1.25, real=real)
    if reader is None:
        pc.last, pc.current = (pc.plane(p) for p in c["planes"])
        pc.time_ratio = tr
    bcs = [vfx.BoundaryCondition(6, m(c["onodes"]), neighbors=m(c["onbrs"]), normal=1),
           vfx.BoundaryCondition(1, m(c["wall"][0]), c["wall"][1]),
           vfx.BoundaryCondition(c["kind"], m(c["inlet"]), c["iq"] if c["kind"] == 12 else None, c["ux"],
                                 precursor=pc)]
    lat.apply_bc(bcs[0], even=True)  # defines what the outflow's second step reads
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    torch.cuda.synchronize()
    return lat, st


@pytest.mark.parametrize("kind", [12, 13])
def test_precursor_channel_fp64_matches_oracle(cuda, kind):
    print(f"[precursor channel kind {kind}]", flush=True)
    c = _setup(kind)
    _, st = _gpu_run(c, "fp64", 60)
    got = st.populations().cpu().numpy()
    ref = oracle.run(c["F0"], DIMS, 60, 1.9, solid=c["solid"], bcs=_oracle_bcs(c, 0.35))
    fl = ~c["solid"]
    err = rel_err(got[fl], ref[fl])
    print(f"  max rel err {err:.2e} exact {np.array_equal(got[fl], ref[fl])}", flush=True)
    assert err <= REL
    assert np.array_equal(got[fl], ref[fl])


@pytest.mark.parametrize("kind", [12, 13])
def test_precursor_sparse_equals_direct(cuda, kind):
    print(f"[precursor sparse kind {kind}]", flush=True)
    c = _setup(kind, seed=5)
    _, sa = _gpu_run(c, "fp64", 40)
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(DIMS, solid=c["solid"], seed=4)
    _, sb = _gpu_run(c, "fp64", 40, index=sidx, layout=(nbx, nby, nbz, fluid))
    fl = ~c["solid"]
    pa = sa.populations().cpu().numpy()[fl]
    pb = sb.populations().cpu().numpy()[sidx][fl]
    print(f"  identical {np.array_equal(pa, pb)}", flush=True)
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("kind", [12, 13])
def test_precursor_plane_cycling_matches_oracle(cuda, kind):
    """Both kinds with a reader: ten steps through the cycle with three steps
    between reads, against the oracle stepped one state at a time.  The
    velocity kind's two-step lag (its value enters the node two collisions
    later) is in the comparison."""
    print(f"[precursor cycling kind {kind}]", flush=True)
    c = _setup(kind, seed=7)
    base = c["planes"][0]

    def plane(t):
        return base + 1e-4 * np.sin(0.7 * t + np.arange(N_PRE))[None, :]

    T, steps = 3, 10
    _, st = _gpu_run(c, "fp64", steps, reader=plane, T=T)
    got = st.populations().cpu().numpy()
    # the oracle in one run with the closed-form cycle (planes at 0, T, 2T, ...)
    pre = dict(neighbors=c["nb"], weights=c["w"], series=[plane(m * T) for m in range(steps // T + 2)], period=T,
               t0=1, velocity_ratio=1.25)
    bcs = [oracle.BCGroup(c["onodes"], kind=6, nbr=c["onbrs"], normal=1), oracle.BCGroup(*c["wall"], kind=1),
           oracle.BCGroup(c["inlet"], c["iq"] if kind == 12 else None, c["ux"], kind=kind, precursor=pre)]
    F = oracle.run(c["F0"], DIMS, steps, 1.9, solid=c["solid"], bcs=bcs)
    fl = ~c["solid"]
    err = rel_err(got[fl], F[fl])
    print(f"  max rel err {err:.2e}", flush=True)
    assert err <= REL
    assert np.array_equal(got[fl], F[fl])


@pytest.mark.parametrize("kind,storage,bar", [(12, "fp32", 1e-4), (13, "fp32", 1e-4), (12, "fp16", 1e-3),
                                              (13, "fp16", 1e-3)])
def test_precursor_reduced_precision_close_to_fp64(cuda, kind, storage, bar):
    c = _setup(kind, seed=11)
    out = {}
    for s in ("fp64", storage):
        lat, st = _gpu_run(c, s, 200)
        _, ux, uy, uz = (t.cpu().numpy() for t in lat.macroscopic(st.even))
        out[s] = np.stack([ux, uy, uz])[:, ~c["solid"]]
    l2 = np.sqrt(((out[storage] - out["fp64"]) ** 2).sum() / (out["fp64"] ** 2).sum())
    print(f"[precursor {storage} kind {kind}] velocity L2 vs FP64 {l2:.2e}", flush=True)
    assert np.isfinite(out[storage]).all()
    assert l2 < bar


@pytest.mark.parametrize("kind", [12, 13])
def test_precursor_planes_from_files(cuda, kind, tmp_path):
    """A precursor series in VirtualFluids' file format (PrecursorWriter's
    .bin.vti planes, one point per inlet node, records every T steps) drives
    the inflow through precursor_from_files: the run equals the same stencil
    fed with the same planes from memory, bit for bit."""
    from virtualfluids_amd import precursor_files as pf

    print(f"[precursor from files kind {kind}]", flush=True)
    c = _setup(kind, seed=11)
    nx, ny, nz = DIMS
    y = (c["inlet"] // nx) % ny
    z = c["inlet"] // (nx * ny)
    comps = 3 if kind == 12 else 9
    T, steps, rec = 3, 10, 3
    py, pz = ny - 2, nz - 2  # plane points at y = 1 .. ny-2, z = 1 .. nz-2
    gy, gz = np.meshgrid(np.arange(1, ny - 1), np.arange(1, nz - 1), indexing="xy")
    base = c["planes"][0][:, :1]

    def values(t):
        return base + 1e-4 * np.sin(0.3 * t + 0.5 * gy.ravel() + gz.ravel())[None, :]

    names = ["vx", "vy", "vz"] if kind == 12 else ["f%d" % j for j in range(9)]
    prefix = str(tmp_path / "inflow")
    for part in range(3):
        data = np.concatenate([values(T * (rec * part + r)) for r in range(rec)], axis=1)
        pf.write_vti(pf.file_name(prefix, 0, 0, part), names, data, (0, py - 1, 0, pz - 1, rec * part,
                     rec * part + rec - 1), (1.0, 1.0, 0.0), (1.0, 1.0, float(T)))
    nb, w, n, q, reader = pf.precursor_from_files(prefix, 0, y.astype(np.float64), z.astype(np.float64),
                                                  real=np.float64)
    assert q == comps and n == c["inlet"].size and np.all(w[0] == 1e6)
    c["nb"], c["w"] = nb, w
    _, a = _gpu_run(c, "fp64", steps, reader=reader, T=T, n_pre=n)
    order = (y - 1) + py * (z - 1)  # the plane point of each inlet node, in slot order

    def memory(t):
        return values(t)[:, order]

    _, b = _gpu_run(c, "fp64", steps, reader=memory, T=T, n_pre=n)
    assert np.array_equal(a.populations().cpu().numpy(), b.populations().cpu().numpy())
