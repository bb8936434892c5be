"""Output in VirtualFluids' VTK format (virtualfluids_amd/output.py and the
C++ FileWriter of the host layer), without a GPU:

* FilePartCalculator and WriterUtilities give the answers of the reference's
  unit tests (tests/unit-tests/gpu/core/Output/FilePartCalculatorTest.cpp,
  WriterUtilitiesTest.cpp);
* a written part reads back to the points, voxels and point data it was given:
  a voxel for every fluid node whose oct is fluid and not wrapped around a
  periodic boundary -- (nx-1)(ny-1)(nz-1) in a periodic box, the interior
  octs inside a solid layer;
* the Python and the C++ writers produce byte-identical .bin.vtu, .pvtu and
  .pvd files from the same input, identical to the files VirtualFluids' own
  writer (WbWriterVtkXmlBinary, compiled from the reference sources) wrote
  from that input (tests/golden/vtk, tests/golden/make_vtk_golden.py);
* the _bin_mean files (FileWriter::writeUnstructuredGridMeanLT) scale the
  means in the real type, and the two writers agree on them;
* parts above 30 million nodes keep the reference's resize behaviour
  (checked with a lowered part limit).
"""
import os
import subprocess

import numpy as np
import pytest

from tests.test_host_abi import build_host_test
from virtualfluids_amd import output as out


def test_file_part_calculator_reference_answers():
    L = out.LIMIT_OF_NODES_FOR_VTK
    assert [out.number_of_parts(v) for v in (2 * L, L, L - 1, 1)] == [3, 2, 1, 1]
    assert out.nodes_in_part(L + 13, 0) == L and out.nodes_in_part(L + 13, 1) == 13
    with pytest.raises(RuntimeError):
        out.nodes_in_part(L + 13, 2)
    assert out.start_of_part(0) == 0 and out.start_of_part(1) == L


def test_writer_utilities_reference_answers():
    c = [np.array([0.0, 1.0]) for _ in range(3)]
    assert not out.is_periodic_cell(c, 1.0, 0, 1) and not out.is_periodic_cell(c, 1.0, 1, 0)
    for axes in range(1, 8):
        cc = [a.copy() for a in c]
        for i in range(3):
            if axes >> i & 1:
                cc[i][1] = 2.0
        assert out.is_periodic_cell(cc, 1.0, 0, 1) and out.is_periodic_cell(cc, 1.0, 1, 0)
    idx = np.arange(8)[::-1]
    nbx, nby, nbz = (np.zeros(8, np.int64) for _ in range(3))
    for i, cur in enumerate(idx):
        nbz[cur] = idx[i + 4] if i < 4 else 99
        nby[cur] = idx[i + 3] if i in (0, 4) else idx[i + 1] if i in (1, 5) else 999
        nbx[cur] = idx[i + 1] if i in (0, 4) else idx[i - 1] if i in (3, 7) else 9999
    assert np.array_equal(out.oct_indices(idx[0], nbx, nby, nbz), idx)


def _box(dims, solid_layer=False, seed=0):
    nx, ny, nz = dims
    k = np.arange(nx * ny * nz)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
    nbx = ((x + 1) % nx) + nx * (y + ny * z)
    nby = x + nx * (((y + 1) % ny) + ny * z)
    nbz = x + nx * (y + ny * ((z + 1) % nz))
    types = np.full(k.size, out.GEO_FLUID, np.uint32)
    if solid_layer:
        wall = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)
        types[wall] = 15
    rng = np.random.default_rng(seed)
    data = [rng.normal(size=k.size) for _ in range(5)] + [types.astype(np.float64)]
    coords = (0.5 + 0.25 * x, -1.0 + 0.25 * y, 2.0 + 0.25 * z)
    return coords, types, (nbx, nby, nbz), 0.25, data


def vtk_golden_case():
    """The input of the VTK golden files (tests/golden/make_vtk_golden.py)."""
    coords, types, nbs, dx, data = _box((9, 6, 5), solid_layer=True, seed=3)
    data.append(np.abs(data[0]) * 1e-3)  # nut
    return {"parts": list(out.unstructured_grid_parts(coords, types, nbs, dx, data)),
            "names": ["press", "rho", "vx1", "vx2", "vx3", "geo", "nut"],
            "mean_names": ["pressMed", "rhoMed", "vx1Med", "vx2Med", "vx3Med", "geo"], "t": 400}


@pytest.mark.parametrize("solid_layer", [False, True])
def test_written_part_reads_back(tmp_path, solid_layer):
    dims = (7, 5, 4)
    coords, types, nbs, dx, data = _box(dims, solid_layer)
    names = ["press", "rho", "vx1", "vx2", "vx3", "geo"]
    (points, cells, pdata), = list(out.unstructured_grid_parts(coords, types, nbs, dx, data))
    path = out.write_octs_with_node_data(str(tmp_path / "p"), points, cells, names, pdata)
    r = out.read_vtu(path)
    nx, ny, nz = dims
    fl = types == out.GEO_FLUID
    want_cells = (nx - 3) * (ny - 3) * (nz - 3) if solid_layer else (nx - 1) * (ny - 1) * (nz - 1)
    assert len(r["connectivity"]) == want_cells
    assert np.all(r["types"] == 11) and np.array_equal(r["offsets"], 8 * np.arange(1, want_cells + 1))
    xyz = np.stack(coords, axis=1).astype(np.float32)
    assert np.array_equal(r["points"][fl], xyz[fl]) and not r["points"][~fl].any()
    for name, d in zip(names, data):
        assert np.array_equal(r["data"][name][fl], d[fl]) and not r["data"][name][~fl].any()
    # every voxel is a unit cube of fluid nodes in VTK_VOXEL corner order
    p = r["points"][r["connectivity"]]
    steps = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]]) * dx
    assert np.allclose(p - p[:, :1], steps[None], atol=1e-6)
    assert np.all(types[r["connectivity"]] == out.GEO_FLUID)


def test_python_and_cpp_writers_are_identical(tmp_path):
    exe = build_host_test()  # against the current library (the host classes' layout changes with it)
    dims = (9, 6, 5)
    coords, types, nbs, dx, data = _box(dims, solid_layer=True, seed=3)
    data.append(np.abs(data[0]) * 1e-3)  # nut
    n, t = types.size, 400
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as fh:
        fh.write(np.int32(n).tobytes())
        for c in coords:
            fh.write(np.asarray(c, np.float64).tobytes())
        for a in (types, *nbs):
            fh.write(np.asarray(a, np.uint32).tobytes())
        fh.write(np.float64(dx).tobytes())
        fh.write(np.int32(len(data)).tobytes())
        for d in data:
            fh.write(np.asarray(d, np.float64).tobytes())
        fh.write(np.int32(t).tobytes())
    cpp = str(tmp_path / "cpp" / "run")
    r = subprocess.run([exe, "vtu", str(inp), cpp], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    py = str(tmp_path / "py" / "run")
    names = ["press", "rho", "vx1", "vx2", "vx3", "geo", "nut"]
    files = []
    for i, (points, cells, pdata) in enumerate(out.unstructured_grid_parts(coords, types, nbs, dx, data), start=1):
        files.append(os.path.basename(out.write_octs_with_node_data(out.part_file_name(py, 0, 0, i, t), points, cells,
                                                                    names, pdata)))
    out.write_parallel_file(out.collection_file_name(py, 0, t), files, names)
    out.write_collection_for_time_series(out.pvd_file_name(py, 0), {t: files, 1000000: files})
    a = sorted(os.listdir(tmp_path / "cpp"))
    b = sorted(os.listdir(tmp_path / "py"))
    assert a == b == sorted([files[0], "run_bin_ID_0_t_400.vtk.pvtu", "run_bin_ID_0.pvd"])
    for f in a:
        assert (tmp_path / "cpp" / f).read_bytes() == (tmp_path / "py" / f).read_bytes(), f
        # this input is the golden case: the reference writer's bytes
        assert (tmp_path / "cpp" / f).read_bytes() == open(os.path.join(GOLDEN_VTK, f), "rb").read(), f
    assert b'timestep="1e+06"' in (tmp_path / "py" / "run_bin_ID_0.pvd").read_bytes()


GOLDEN_VTK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtk")


def _python_golden_files(dest):
    case = vtk_golden_case()
    for prefix, mean in (("run", False), ("mean", True)):
        names = case["mean_names"] if mean else case["names"]
        pieces = []
        for i, (points, cells, data) in enumerate(case["parts"], start=1):
            name = out.part_file_name(os.path.join(dest, prefix), 0, 0, i, case["t"], mean=mean)
            pieces.append(os.path.basename(out.write_octs_with_node_data(name, points, cells, names,
                                                                         data[:len(names)])))
        out.write_parallel_file(out.collection_file_name(os.path.join(dest, prefix), 0, case["t"], mean=mean), pieces,
                                names)
        if not mean:
            out.write_collection_for_time_series(out.pvd_file_name(os.path.join(dest, prefix), 0),
                                                 {case["t"]: pieces, 1000000: pieces})


def test_python_writer_matches_reference_writer(tmp_path):
    """Byte identity with the files VirtualFluids' WbWriterVtkXmlBinary wrote
    from the same points, voxels and data (writeOctsWithNodeData,
    writeParallelFile, writeCollectionForTimeSeries)."""
    _python_golden_files(str(tmp_path))
    golden = sorted(os.listdir(GOLDEN_VTK))
    assert len(golden) == 5 and sorted(os.listdir(tmp_path)) == golden
    for f in golden:
        assert (tmp_path / f).read_bytes() == open(os.path.join(GOLDEN_VTK, f), "rb").read(), f


def test_reference_writer_still_writes_the_golden_files(tmp_path):
    """Where the reference tree exists: its writer, run now, writes the
    committed golden files."""
    sys_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    import sys

    sys.path.insert(0, sys_path)
    try:
        import make_vtk_golden as g
    finally:
        sys.path.remove(sys_path)
    if not os.path.exists(g.EXE):
        pytest.skip("oracle/_ref/ref_vtk not built (no reference tree here)")
    files = g.reference_files(str(tmp_path))
    assert files == sorted(os.listdir(GOLDEN_VTK))
    for f in files:
        assert (tmp_path / f).read_bytes() == open(os.path.join(GOLDEN_VTK, f), "rb").read(), f


def test_mean_files_scale_in_the_real_type(tmp_path):
    """_bin_mean parts: meanPressureOut / 3.0f * densityRatio * velocityRatio^2
    and meanVelocityOut * velocityRatio evaluated in float for a float lattice
    (the ratios rounded to float first), in double for a double one; the C++
    host writer writes the same bytes for the double case."""
    coords, types, nbs, dx, data = _box((6, 5, 4), solid_layer=True, seed=11)
    rng = np.random.default_rng(12)
    dr, vr = 1.2345678901, 0.0987654321
    for real in (np.float32, np.float64):
        mean = {k: rng.normal(size=types.size).astype(real) for k in ("vx", "vy", "vz", "rho", "press")}
        lev = {"coords": coords, "types": types, "neighbors": nbs, "grid_spacing": dx, "rho": mean["rho"]}
        w = out.FileWriter(str(tmp_path / real.__name__ / "m"), density_ratio=dr, velocity_ratio=vr)
        f, = w.write_level_mean(lev, mean, 0, 7)
        r = out.read_vtu(f)
        assert list(r["data"]) == out.FileWriter.mean_node_data_names()
        fl = types == out.GEO_FLUID
        R = real
        want = (mean["press"] / R(3.0)) * R(dr) * R(vr) * R(vr)
        assert np.array_equal(r["data"]["pressMed"][fl], want[fl].astype(np.float64))
        assert np.array_equal(r["data"]["vx2Med"][fl], (mean["vy"] * R(vr))[fl].astype(np.float64))
        if real is np.float32:  # a double-rounded ratio would differ
            assert not np.array_equal(r["data"]["pressMed"][fl], (mean["press"].astype(np.float64) / 3.0 * dr * vr * vr)[fl])
    exe = build_host_test()
    n, t = types.size, 7
    inp = tmp_path / "mean.bin"
    with open(inp, "wb") as fh:
        fh.write(np.int32(n).tobytes())
        for c in coords:
            fh.write(np.asarray(c, np.float64).tobytes())
        for a in (types, *nbs):
            fh.write(np.asarray(a, np.uint32).tobytes())
        fh.write(np.float64(dx).tobytes())
        for k in ("vx", "vy", "vz", "rho", "press"):
            fh.write(np.asarray(mean[k], np.float64).tobytes())
        fh.write(np.int32(t).tobytes())
        fh.write(np.float64(dr).tobytes())
        fh.write(np.float64(vr).tobytes())
    cpp = str(tmp_path / "cpp" / "m")
    r = subprocess.run([exe, "vtumean", str(inp), cpp], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    py = tmp_path / "float64"
    out.write_parallel_file(out.collection_file_name(str(py / "m"), 0, t, mean=True),
                            [os.path.basename(out.part_file_name("m", 0, 0, 1, t, mean=True)) + ".bin.vtu"],
                            out.FileWriter.mean_node_data_names())
    a = sorted(os.listdir(tmp_path / "cpp"))
    assert a == sorted(os.listdir(py)) == ["m_bin_mean_ID_0_t_7.vtk.pvtu", "m_bin_mean_lev_0_ID_0_Part_1_t_7.vtk.bin.vtu"]
    for f in a:
        assert (tmp_path / "cpp" / f).read_bytes() == (py / f).read_bytes(), f


def test_parts_keep_the_reference_resize_behaviour(monkeypatch):
    """Split a 7x5x4 box into parts of 50 nodes: a non-fluid slot of a later part
    keeps the point and data the previous part left there, and an oct may use
    the first node after its part."""
    monkeypatch.setattr(out, "LIMIT_OF_NODES_FOR_VTK", 50)
    coords, types, nbs, dx, data = _box((7, 5, 4), solid_layer=True, seed=5)
    parts = list(out.unstructured_grid_parts(coords, types, nbs, dx, data))
    assert len(parts) == 3 and [p[0].shape[0] for p in parts] == [50, 50, 40]
    p0, p1 = parts[0], parts[1]
    fl1 = types[50:100] == out.GEO_FLUID
    assert np.array_equal(p1[0][~fl1], p0[0][~fl1]) and np.array_equal(p1[2][0][~fl1], p0[2][0][~fl1])
    assert all((c.max(initial=-1) <= 50) for _, c, _ in parts)


def test_run_metadata(tmp_path):
    """PerformanceMeasurement's node updates per second count level l 2^l
    times; the record keeps the reference's keys in its order."""
    import yaml

    assert out.nups(100, [1000, 500], 2.0) == 100 * (1000 + 2 * 500) / 2.0
    meta = out.metadata("cube", [1000, 500], 100, 2.0, reynolds_number=4000.0, gpus=["MI355X"])
    path = out.write_yaml(meta, tmp_path / "run" / "cube.yaml")
    back = yaml.safe_load(open(path))
    assert list(back) == ["name", "number_of_processes", "number_of_threads", "vf_hardware", "simulation", "world",
                          "discretization", "build_info", "gpus"]
    assert list(back["simulation"]) == ["startDateTime", "runtimeSeconds", "nups", "numberOfTimeSteps",
                                        "collisionKernel", "reynoldsNumber", "lb_velocity", "lb_viscosity",
                                        "quadricLimiter"]
    assert back["simulation"]["nups"] == 100000.0
    assert back["discretization"]["numberOfNodesPerLevel"] == [1000, 500]
