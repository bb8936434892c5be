"""Precursor inflow files (virtualfluids_amd.precursor_files), without a GPU:

* the writer produces the bytes VirtualFluids' own image writer wrote for the
  same planes (golden files in tests/golden/precursor/, written by
  WbWriterVtkXmlImageBinary compiled from the reference);
* the reader takes the header and values back from those files;
* the stencil of boundary nodes on plane points is the exact match (the
  point alone, weight 1e6) and off the points the four surrounding points with
  inverse squared distances (closed form), each point in one plane slot;
* the planes come from the record nearest in time, from the second file once
  the first file's time range is left, and reading past the last file fails;
* a NaN in the first record at a point a stencil reads is refused.

VirtualFluids' reader itself (TransientBCSetter.cpp) needs spdlog, which the
reference tree does not hold, so it is not compiled here; these checks and
the reference-written files pin the restatement.
"""
import os

import numpy as np
import pytest

from virtualfluids_amd import precursor_files as pf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "precursor")
NY, NZ, REC = 6, 5, 4          # plane points along y and z, records per file
Y0, Z0, H, DT = 0.25, 0.5, 0.5, 0.125


def field(y, z, t):
    return np.stack([0.05 + 0.01 * np.sin(y) * np.cos(z) + 0.002 * t,
                     0.003 * np.cos(y + z) - 0.001 * t,
                     0.002 * np.sin(2 * z) + 0.0005 * y * t])


def precursor_case():
    """Two files of four records each, the plane's x = simulation y."""
    iy, iz = np.meshgrid(np.arange(NY), np.arange(NZ), indexing="xy")
    y = (Y0 + H * iy).ravel().astype(np.float64)
    z = (Z0 + H * iz).ravel().astype(np.float64)
    parts = []
    for part in range(2):
        recs = [field(y, z, DT * (REC * part + r)) for r in range(REC)]
        data = np.concatenate(recs, axis=1)  # record-major: x fastest, then y, then time
        parts.append(dict(part=part, data=data, extent=(0, NY - 1, 0, NZ - 1, REC * part, REC * part + REC - 1)))
    return dict(prefix="inflow", names=["vx", "vy", "vz"], origin=(Y0, Z0, 0.0), spacing=(H, H, DT), precision=8,
                parts=parts, y=y, z=z)


def _write(tmp_path, case=None):
    case = case or precursor_case()
    for p in case["parts"]:
        pf.write_vti(pf.file_name(tmp_path / case["prefix"], 0, 0, p["part"]), case["names"], p["data"],
                     p["extent"], case["origin"], case["spacing"], case["precision"])
    return str(tmp_path / case["prefix"])


def test_writer_bytes_are_the_reference_writers(tmp_path):
    _write(tmp_path)
    for part in (0, 1):
        name = f"inflow_lev_0_ID_0_File_{part}.bin.vti"
        assert (tmp_path / name).read_bytes() == open(os.path.join(GOLDEN, name), "rb").read(), name


def test_reader_takes_header_and_values_from_reference_files():
    case = precursor_case()
    for p in case["parts"]:
        f = pf.VTKFile(os.path.join(GOLDEN, f"inflow_lev_0_ID_0_File_{p['part']}.bin.vti"))
        assert (f.nx, f.ny, f.nz) == (NY, NZ, REC)
        assert [q[0] for q in f.quantities] == case["names"]
        assert f.min_x == np.float32(Y0) and f.min_y == np.float32(Z0)
        assert f.min_z == np.float32(np.float32(DT) * np.float32(REC * p["part"]))
        assert f.max_x == np.float32(np.float32(NY - 1) * np.float32(H) + np.float32(Y0))
        f.load()
        for j in range(3):
            assert np.array_equal(f.values[j], p["data"][j])


def test_exact_match_stencil_and_planes_through_the_files(tmp_path):
    case = precursor_case()
    prefix = _write(tmp_path)
    sel = np.array([0, 3, 7, 12, 29])  # boundary nodes on plane points, in plane order
    nb, w, n, q, reader = pf.precursor_from_files(prefix, 0, case["y"][sel], case["z"][sel])
    assert q == 3 and n == sel.size
    assert np.array_equal(nb, np.tile(np.arange(sel.size, dtype=np.uint32), (4, 1)))
    assert np.array_equal(w, np.tile(np.array([[1e6], [0], [0], [0]], np.float32), (1, sel.size)))
    for t, part, rec in ((0.0, 0, 0), (0.125, 0, 1), (0.3, 0, 2), (0.375, 0, 3), (0.5, 1, 0), (0.86, 1, 3)):
        want = case["parts"][part]["data"][:, rec * NY * NZ + sel].astype(np.float32)
        assert np.array_equal(reader(t), want), t
    with pytest.raises(RuntimeError):
        reader(0.93)  # past the last record


def test_off_grid_stencil_is_the_four_surrounding_points(tmp_path):
    case = precursor_case()
    prefix = _write(tmp_path)
    rng = np.random.default_rng(3)
    iy, iz = rng.integers(0, NY - 1, 6), rng.integers(0, NZ - 1, 6)
    fy, fz = rng.uniform(0.1, 0.9, 6), rng.uniform(0.1, 0.9, 6)
    y, z = Y0 + H * (iy + fy), Z0 + H * (iz + fz)
    rd = pf.VTKReader(pf.VTKFileCollection(prefix), 0)
    rd.fill_arrays(y, z)
    nb, w = rd.stencil()
    assert not rd.perfect_match
    read = np.array(rd.read_idx[0][0])
    corners = {0: (1, 1), 1: (1, 0), 2: (0, 1), 3: (0, 0)}  # 0PP, 0PM, 0MP, 0MM
    for s, (dy, dz) in corners.items():
        lin = (iy + dy) + NY * (iz + dz)
        assert np.array_equal(read[nb[s]], lin), s
        d2 = (Y0 + H * (iy + dy) - y) ** 2 + (Z0 + H * (iz + dz) - z) ** 2
        assert np.allclose(w[s], 1.0 / (d2 + 1e-7), rtol=1e-5), s
    # the plane holds each stencil point's value
    data = np.zeros(3 * rd.n_points_read, np.float32)
    rd.get_next_data(data, rd.n_points_read, 0.125)
    rec = case["parts"][0]["data"][:, NY * NZ: 2 * NY * NZ]
    assert np.array_equal(data.reshape(3, -1), rec[:, read].astype(np.float32))


def test_nan_at_a_read_point_is_refused(tmp_path):
    case = precursor_case()
    case["parts"][0]["data"][0, 1] = np.nan  # first record, plane point 1
    prefix = _write(tmp_path, case)
    with pytest.raises(RuntimeError):
        pf.precursor_from_files(prefix, 0, case["y"][:4], case["z"][:4])
    # a NaN at a point no stencil reads is allowed
    case = precursor_case()
    case["parts"][0]["data"][0, 25] = np.nan
    prefix = _write(tmp_path, case)
    pf.precursor_from_files(prefix, 0, case["y"][:4], case["z"][:4])


def test_no_files_is_an_error(tmp_path):
    with pytest.raises(RuntimeError):
        pf.VTKFileCollection(str(tmp_path / "missing"))


def test_shared_stencil_points_get_one_slot(tmp_path):
    """Neighbouring off-grid nodes share corner points: one slot per file
    point, every slot's read index distinct."""
    prefix = _write(tmp_path)
    y = Y0 + H * np.array([0.5, 1.5, 2.5, 0.5, 1.5])
    z = Z0 + H * np.array([0.5, 0.5, 0.5, 1.5, 1.5])
    rd = pf.VTKReader(pf.VTKFileCollection(prefix), 0)
    rd.fill_arrays(y, z)
    read = rd.read_idx[0][0]
    # corners: y 0-3 at z 0 and 1 (first row of cells), y 0-2 at z 2 (second row)
    assert len(read) == len(set(read)) == rd.n_points_read == 11


def test_host_precursor_reader_equals_python(tmp_path):
    """The C++ host layer's VTKReader (double `real`) on the same files gives
    the Python reader's stencils, weights and planes, on and off the grid."""
    import subprocess
    from tests.test_host_abi import build_host_test

    exe = build_host_test()
    case = precursor_case()
    prefix = _write(tmp_path)
    rng = np.random.default_rng(8)
    y = np.concatenate([case["y"][[2, 9, 17]], Y0 + H * rng.uniform(0.05, NY - 1.05, 4)])
    z = np.concatenate([case["z"][[2, 9, 17]], Z0 + H * rng.uniform(0.05, NZ - 1.05, 4)])
    times = [0.0, 0.2, 0.37, 0.51, 0.8]
    args = [exe, "precursor", prefix, str(y.size)] + [repr(float(v)) for v in y] + [repr(float(v)) for v in z] + \
        [repr(t) for t in times]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    rd = pf.VTKReader(pf.VTKFileCollection(prefix, np.float64), 0, np.float64)
    rd.fill_arrays(y, z)
    nb, w = rd.stencil()
    assert lines[0] == f"points {rd.n_points_read} quantities 3"
    for s in range(4):
        assert [int(v) for v in lines[1 + 2 * s].split()[1:]] == nb[s].tolist(), s
        assert np.array_equal(np.array([float(v) for v in lines[2 + 2 * s].split()[1:]]), w[s]), s
    for j, t in enumerate(times):
        data = np.zeros(3 * rd.n_points_read)
        rd.get_next_data(data, rd.n_points_read, t)
        assert np.array_equal(np.array([float(v) for v in lines[9 + j].split()[1:]]), data), t
