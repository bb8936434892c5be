"""GPU parity of the non-reflecting outflow boundaries
(OutflowNonReflecting / OutflowNonReflectingPressureCorrection) and the
velocity inlet with pressure (VelocityWithPressureInterpolatedCompressible)
against the oracle: a channel with that inlet, interpolated no-slip walls and
the outflow on the last fluid layer.

FP64: every fluid population after 60 steps within 1e-12 relative of the
oracle (expected exact); sparse layout equal to the direct one bit for bit;
FP32 / FP16 storage close to FP64 (velocity L2 below 1e-4 / 2e-3).

The outflow kernels write the populations the collision after next reads, so
each run applies them once with the starting parity before the first step
(the oracle's outflow then holds from the second step on, as in the GPU run).
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
U = 0.03


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _spec(kind):
    solid, groups, onodes, onbrs = cases.inflow_channel(*DIMS)
    rng = np.random.default_rng(kind)
    inlet_nodes, inlet_q = groups["inlet"]
    spec = [{"kind": kind, "nodes": onodes, "nbr": onbrs, "normal": 1, "corr": 0.3 if kind == 7 else 0.0},
            {"kind": 1, "nodes": groups["wall"][0], "q": groups["wall"][1]},
            {"kind": 5, "nodes": inlet_nodes, "q": inlet_q,
             "ux": U * (1 + 0.01 * rng.uniform(-1, 1, inlet_nodes.size))}]
    return solid, spec


def _initial():
    N = int(np.prod(DIMS))
    drho, ux, uy, uz = cases.random_flow(DIMS, seed=9, u_amp=0.01)
    return oracle.init_field(drho, ux + U, uy, uz)


def _gpu_bcs(spec, index=None):
    m = (lambda a: a) if index is None else (lambda a: index[a])
    return [vfx.BoundaryCondition(g["kind"], m(g["nodes"]), g.get("q"), g.get("ux"), neighbors=None if "nbr" not in g
                                  else m(g["nbr"]), normal=g.get("normal", 0),
                                  pressure_correction_factor=g.get("corr", 0.0)) for g in spec]


def _run_gpu(lat, bcs, steps):
    # the Stepper applies the outflow kinds once with the starting parity
    # before its first step (prime_outflow): that defines what the second step reads
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    torch.cuda.synchronize()
    return st


@pytest.mark.parametrize("kind", [6, 7])
def test_outflow_channel_fp64_matches_oracle(cuda, kind):
    print(f"[outflow channel kind {kind}]", flush=True)
    solid, spec = _spec(kind)
    F0 = _initial()
    lat = vfx.Lattice(DIMS, storage="fp64", omega=1.9, solid=solid)
    lat.write_populations(F0)
    st = _run_gpu(lat, _gpu_bcs(spec), 60)
    got = st.populations().cpu().numpy()
    obcs = [oracle.BCGroup(g["nodes"], g.get("q"), g.get("ux"), kind=g["kind"], nbr=g.get("nbr"),
                           normal=g.get("normal", 0), correction=g.get("corr", 0.0)) for g in spec]
    ref = oracle.run(F0, DIMS, 60, 1.9, solid=solid, bcs=obcs)
    fl = ~solid
    err = rel_err(got[fl], ref[fl])
    print(f"  max rel err {err:.2e} exact {np.array_equal(got[fl], ref[fl])}", flush=True)
    assert err <= REL


@pytest.mark.parametrize("kind", [6, 7])
def test_outflow_sparse_equals_direct(cuda, kind):
    print(f"[outflow sparse kind {kind}]", flush=True)
    solid, spec = _spec(kind)
    F0 = _initial()
    a = vfx.Lattice(DIMS, storage="fp64", omega=1.9, solid=solid)
    a.write_populations(F0)
    sa = _run_gpu(a, _gpu_bcs(spec), 40)
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(DIMS, solid=solid, seed=4)
    Fs = np.zeros((sidx.size + 1, 27))
    Fs[sidx] = F0
    b = vfx.Lattice(storage="fp64", omega=1.9, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
    b.write_populations(Fs)
    sb = _run_gpu(b, _gpu_bcs(spec, sidx), 40)
    fl = ~solid
    pa = sa.populations().cpu().numpy()[fl]
    pb = sb.populations().cpu().numpy()[sidx][fl]
    print(f"  identical {np.array_equal(pa, pb)}", flush=True)
    assert np.array_equal(pa, pb)


@pytest.mark.parametrize("storage,bar", [("fp32", 1e-4), ("fp16", 2e-3)])
def test_outflow_reduced_precision(cuda, storage, bar):
    solid, spec = _spec(6)
    F0 = _initial()
    out = {}
    for s in ("fp64", storage):
        ref = (U, 0.0, 0.0) if s == "fp16" else (0.0, 0.0, 0.0)
        lat = vfx.Lattice(DIMS, storage=s, omega=1.9, solid=solid, reference_velocity=ref)
        lat.write_populations(F0)
        st = _run_gpu(lat, _gpu_bcs(spec), 300)
        _, ux, uy, uz = (t.cpu().numpy() for t in lat.macroscopic(st.even))
        out[s] = np.stack([ux, uy, uz])[:, ~solid]
    l2 = np.sqrt(((out[storage] - out["fp64"]) ** 2).sum() / (out["fp64"] ** 2).sum())
    print(f"[outflow {storage}] velocity L2 vs FP64 {l2:.2e}", flush=True)
    assert l2 < bar
