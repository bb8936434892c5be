"""GPU side of the output path: vfx_macroscopic_quantities (the reference's
calculateMacroscopicQuantitiesCompressible) against the oracle, and a whole
output step written with FileWriter read back.

FP64: rho, velocities and pressure of every fluid node equal the oracle's
restatement of getDensity / getCompressibleVelocity / getPressure (pinned
against the compiled reference functions) bit for bit; every other node is 0.
"""
import os

import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases
from virtualfluids_amd import output as out

pytestmark = pytest.mark.gpu


def _cavity_state(storage="fp64", steps=9):
    n = 10
    solid, groups = cases.cavity(n, lid=(0.08, 0.0, 0.0))
    m = n + 2
    fields = cases.random_flow((m, m, m), seed=41, u_amp=0.04)
    lat = vfx.Lattice((m, m, m), storage=storage, omega=1.9, solid=solid)
    lat.init_equilibrium(*fields)
    st = vfx.Stepper(lat, "k17", [vfx.BoundaryCondition(0, nd, q, u[0], u[1], u[2]) for nd, q, u in groups])
    st.step(steps)
    torch.cuda.synchronize()
    return lat, st, solid


def test_macroscopic_quantities_match_oracle(cuda):
    print("[macroscopic quantities fp64]", flush=True)
    lat, st, solid = _cavity_state()
    got = np.stack([t.cpu().numpy() for t in lat.macroscopic_quantities(st.even)], axis=1)
    F = st.populations().cpu().numpy()
    fl = ~solid
    want = np.array([oracle.macroscopic_quantities(f) for f in F[fl]])
    print(f"  identical {np.array_equal(got[fl], want)}", flush=True)
    assert np.array_equal(got[fl], want)
    assert not got[~fl].any()


def test_macroscopic_quantities_sparse_node_selection(cuda):
    """Sparse layout: with typeOfGridNode the types 19, 5, 6, 7 are computed and
    everything else is 0 (isValidFluidNode); with a fluid-node list the listed
    nodes are."""
    dims = (8, 6, 5)
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, seed=6)
    F0 = oracle.init_field(*cases.random_flow(dims, seed=8)) + \
        np.random.default_rng(2).normal(0, 1e-3, (int(np.prod(dims)), 27)) * cases.W
    Fs = np.zeros((sidx.size + 1, 27))
    Fs[sidx] = F0
    rng = np.random.default_rng(3)
    types = rng.choice(np.array([19, 5, 6, 7, 15, 8], np.uint32), sidx.size + 1)
    types[0] = 15
    a = vfx.Lattice(storage="fp64", omega=1.8, neighbors=(nbx, nby, nbz), node_types=types)
    a.write_populations(Fs)
    got = np.stack([t.cpu().numpy() for t in a.macroscopic_quantities(True)], axis=1)
    valid = np.isin(types, [19, 5, 6, 7])
    want = np.array([oracle.macroscopic_quantities(f) for f in Fs])
    assert np.array_equal(got[valid], want[valid]) and not got[~valid].any()
    listed = np.sort(rng.choice(np.arange(1, sidx.size + 1), 60, replace=False)).astype(np.uint32)
    b = vfx.Lattice(storage="fp64", omega=1.8, neighbors=(nbx, nby, nbz), fluid_nodes=listed)
    b.write_populations(Fs)
    got = np.stack([t.cpu().numpy() for t in b.macroscopic_quantities(True)], axis=1)
    mask = np.zeros(sidx.size + 1, bool)
    mask[listed] = True
    assert np.array_equal(got[mask], want[mask]) and not got[~mask].any()


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_file_writer_output_step(cuda, tmp_path, storage):
    """Two output steps of the cavity: parts, .pvtu and .pvd; the point data
    read back are the reference's scaled fields (press and rho / 3), geo the
    node types; voxels only between fluid nodes."""
    print(f"[output step {storage}]", flush=True)
    lat, st, solid = _cavity_state(storage)
    w = out.FileWriter(str(tmp_path / "cav"), density_ratio=1.0, velocity_ratio=1.0)
    w.write_timestep([out.structured_level(lat, st.even)], 9)
    lev = out.structured_level(lat, st.even)
    path = w.write_timestep([lev], 10, last=True)
    files = sorted(os.listdir(tmp_path))
    assert "cav_bin_ID_0.pvd" in files and "cav_bin_ID_0_t_10.vtk.pvtu" in files
    r = out.read_vtu(str(tmp_path / "cav_bin_lev_0_ID_0_Part_1_t_10.vtk.bin.vtu"))
    fl = ~solid
    assert np.array_equal(r["data"]["press"][fl], np.asarray(lev["press"], np.float64)[fl] / 3.0)
    assert np.array_equal(r["data"]["rho"][fl], np.asarray(lev["rho"], np.float64)[fl] / 3.0)
    assert np.array_equal(r["data"]["vx1"][fl], np.asarray(lev["vx"], np.float64)[fl])
    assert np.all(r["data"]["geo"][fl] == 19)
    assert len(r["connectivity"]) == 9 ** 3  # octs of the 10^3 fluid nodes
    if storage == "fp64":
        F = st.populations().cpu().numpy()
        want = np.array([oracle.macroscopic_quantities(f) for f in F[fl]])
        assert np.array_equal(r["data"]["vx3"][fl], want[:, 3])
    pvd = open(str(tmp_path / "cav_bin_ID_0.pvd")).read()
    assert 'timestep="9"' in pvd and 'timestep="10"' in pvd
    assert open(path).read().count("<Piece Source=") == 1


def test_mean_values_match_oracle(cuda):
    """MeanValues (calculateMeanCompressible at every step) on the cavity:
    FP64 sums equal the oracle's mean_add over the populations of each step
    bit for bit; solid nodes hold zero sums; calcMean is sum / (real)tdiff and
    resetMean zeroes the sums."""
    print("[mean values fp64]", flush=True)
    lat, st, solid = _cavity_state(steps=3)
    mv = vfx.MeanValues(lat)
    fl = ~solid
    acc = np.zeros((int(fl.sum()), 5))
    for _ in range(4):
        mv.accumulate(st.even)
        F = st.populations().cpu().numpy()[fl]
        acc = np.array([oracle.mean_add(f, a) for f, a in zip(F, acc)])
        st.step(1)
    torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in mv.sums], axis=1)  # vx, vy, vz, rho, press
    want = acc[:, [1, 2, 3, 0, 4]]
    print(f"  identical {np.array_equal(got[fl], want)}", flush=True)
    assert np.array_equal(got[fl], want)
    assert not got[~fl].any()
    mean = mv.mean(4)
    assert torch.equal(mean["rho"], mv.sums[3] / 4.0)
    mv.reset()
    torch.cuda.synchronize()
    assert not any(bool(t.any()) for t in mv.sums)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_mean_values_reduced_storage(cuda, storage):
    """FP32 sums of an FP32 / FP16 lattice against the FP64 oracle sums of
    the same populations (read back in double)."""
    print(f"[mean values {storage}]", flush=True)
    lat, st, solid = _cavity_state(storage, steps=3)
    mv = vfx.MeanValues(lat)
    assert mv.sums[0].dtype == torch.float32
    fl = ~solid
    acc = np.zeros((int(fl.sum()), 5))
    for _ in range(3):
        mv.accumulate(st.even)
        F = st.populations().cpu().numpy()[fl]
        acc = np.array([oracle.mean_add(f, a) for f, a in zip(F, acc)])
        st.step(1)
    got = np.stack([t.cpu().numpy().astype(np.float64) for t in mv.sums], axis=1)[fl]
    want = acc[:, [1, 2, 3, 0, 4]]
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < 1e-5


def test_mean_output_step(cuda, tmp_path):
    """An output step with means: the _bin_mean part holds the means of the
    interval (sums / tdiff) scaled as writeUnstructuredGridMeanLT, next to the
    instantaneous part; the mean .pvtu lists it."""
    print("[mean output step]", flush=True)
    lat, st, solid = _cavity_state(steps=2)
    mv = vfx.MeanValues(lat)
    for _ in range(5):
        st.step(1)
        mv.accumulate(st.even)
    mean = {k: v.cpu().numpy() for k, v in mv.mean(5).items()}
    w = out.FileWriter(str(tmp_path / "cav"))
    w.write_timestep([out.structured_level(lat, st.even)], 7, means=[mean])
    files = sorted(os.listdir(tmp_path))
    assert "cav_bin_mean_lev_0_ID_0_Part_1_t_7.vtk.bin.vtu" in files and "cav_bin_mean_ID_0_t_7.vtk.pvtu" in files
    r = out.read_vtu(str(tmp_path / "cav_bin_mean_lev_0_ID_0_Part_1_t_7.vtk.bin.vtu"))
    fl = ~solid
    assert np.array_equal(r["data"]["rhoMed"][fl], (mean["rho"] / 3.0)[fl])
    assert np.array_equal(r["data"]["vx1Med"][fl], mean["vx"][fl])
    assert "cav_bin_mean_lev_0_ID_0_Part_1_t_7.vtk.bin.vtu" in open(tmp_path / "cav_bin_mean_ID_0_t_7.vtk.pvtu").read()
