"""VirtualFluids grid folders (virtualfluids_amd.gridfiles), without a GPU:

* a lid-driven cavity written in SimulationFileWriter's layout (ASCII and
  binary coordinate / neighbour / geo files, side Qs and values) is read back
  by this reader exactly as VirtualFluids' own readers (CoordNeighborGeoV,
  BoundaryQs, BoundaryValues, compiled from the reference sources into
  oracle/_ref/ref_gridfiles) read it -- same levels, sizes, neighbours,
  coordinates, indices, q and values;
* the boundary tables follow GridReader: velocity sides joined in side order
  and divided by the velocity ratio, pressure values over the pressure factor
  with their neighbours, the geometry q of d000 set to 0, the reader's
  end-of-line and single-zero behaviour.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import cases
from virtualfluids_amd import gridfiles as gf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "ref_gridfiles")


def cavity_grid(n=6, lid=(0.05, 0.02, 0.0), seed=4):
    """The cases.cavity box as a one-level grid folder: every box node a sparse
    node (walls GEO_STOPPER 16, fluid 19) in a shuffled order, periodic
    neighbours, the lid links on the top side (velocity), the wall links on the
    bottom side (noSlip).  Returns (GridLevel, sides, sparse index of box node)."""
    solid, groups = cases.cavity(n, lid=lid)
    m = n + 2
    dims = (m, m, m)
    nbx, nby, nbz, _, sidx = cases.sparse_layout(dims, solid, seed=seed)
    wsb = cases.sparse_inverse(dims, sidx)
    N = m ** 3
    k = np.arange(N)
    coords = []
    for c in (k % m, (k // m) % m, k // (m * m)):
        a = np.zeros(N + 1)
        a[sidx] = 0.25 * c - 1.0
        coords.append(a)
    geo = np.full(N + 1, gf.GEO_STOPPER, np.uint32)
    geo[sidx[~solid]] = gf.GEO_FLUID
    level = gf.GridLevel(tuple(coords), (nbx, nby, nbz), wsb, geo)
    (lid_nodes, lid_q, u), (wall_nodes, wall_q, _) = groups
    vel = np.tile(np.array(u, np.float64)[:, None], (1, lid_nodes.size))
    sides = {"top": ("velocity", [(sidx[lid_nodes], lid_q, vel)]),
             "bottom": ("noSlip", [(sidx[wall_nodes], wall_q, np.zeros((3, wall_nodes.size)))])}
    return level, sides, sidx


def _ref(*args):
    r = subprocess.run([REF_EXE, *map(str, args)], capture_output=True, text=True, check=True)
    return r.stdout.split()


needs_ref = pytest.mark.skipif(not os.path.exists(REF_EXE), reason="oracle/_ref/ref_gridfiles not built")


@needs_ref
@pytest.mark.parametrize("binary", [False, True])
def test_coordinates_neighbours_geo_read_as_the_reference_reads_them(tmp_path, binary):
    level, sides, _ = cavity_grid()
    gf.write_grid(tmp_path, [level], sides, binary=binary)
    files = [("coordX", True, level.coords[0]), ("coordZ", True, level.coords[2]),
             ("neighborX", False, level.neighbors[0]), ("neighborY", False, level.neighbors[1]),
             ("neighborWSB", False, level.neighbor_wsb), ("geoVec", False, level.geo)]
    for name, coord, want in files:
        path = tmp_path / f"{name}.dat"
        ours = gf.read_coord_neighbor_geo(path, binary, coord)
        t = _ref("cng", path, int(binary), int(coord))
        assert int(t[0]) == 0 and int(t[1]) == want.size - 1
        ref = np.array([float(v) for v in t[2:]]) if coord else np.array([int(v) for v in t[2:]])
        assert len(ours) == 1 and np.array_equal(ours[0], ref), name
        if binary or not coord:
            assert np.array_equal(ours[0], want), name
        else:  # ASCII coordinates carry six significant digits, as written
            assert np.allclose(ours[0], want, rtol=1e-6)


@needs_ref
def test_boundary_qs_and_values_read_as_the_reference_reads_them(tmp_path):
    level, sides, _ = cavity_grid()
    gf.write_grid(tmp_path, [level], sides)
    for side in ("top", "bottom", "inlet"):
        path = tmp_path / f"{side}BoundaryQs.dat"
        (idx, q), = gf.read_boundary_qs(path)
        t = _ref("qs", path)
        assert int(t[0]) == 0 and int(t[1]) == idx.size
        rows = np.array([float(v) for v in t[2:]]).reshape(-1, 28) if idx.size else np.zeros((0, 28))
        assert np.array_equal(rows[:, 0].astype(np.int64), idx)
        assert np.array_equal(rows[:, 1:].T, q)
        name, per = gf.read_boundary_values(tmp_path / f"{side}BoundaryValues.dat")
        t = _ref("values", tmp_path / f"{side}BoundaryValues.dat")
        assert t[0] == name and int(t[1]) == len(per) - 1
        pos = 3
        ncols = int(t[2])
        (vidx, vals), = per
        for c in range(ncols):
            cnt = int(t[pos])
            col = np.array([float(v) for v in t[pos + 1:pos + 1 + cnt]])
            pos += 1 + cnt
            assert np.array_equal(col, vals[c]) if c < vals.shape[0] else cnt == 0
    # the written tables come back
    (idx, q), = gf.read_boundary_qs(tmp_path / "topBoundaryQs.dat")
    k, q0, _ = sides["top"][1][0]
    assert np.array_equal(idx, k) and np.array_equal(q, np.where(q0 > 0, q0, -1.0))


def test_reader_skips_the_rest_of_a_q_line_and_reads_single_zero_values(tmp_path):
    """BoundaryQs::init reads the q of the key's bits and drops the rest of the
    line; a value file of an unknown condition is one zero entry."""
    (tmp_path / "q.dat").write_text("0\n2\n5 3 0.25 0.5 9.9 extra\n7 4 0.75\n")
    (idx, q), = gf.read_boundary_qs(tmp_path / "q.dat")
    assert idx.tolist() == [5, 7]
    assert q[26, 0] == 0.25 and q[25, 0] == 0.5 and q[24, 1] == 0.75
    assert (q[:24, 0] == -1).all() and q[26, 1] == -1
    (tmp_path / "v.dat").write_text("slip\n0\n3\n1 0 0 0\n")
    name, per = gf.read_boundary_values(tmp_path / "v.dat")
    assert name == "slip" and len(per) == 1 and per[0][0].tolist() == [0] and per[0][1].tolist() == [[0.0]]


def test_grid_reader_boundary_tables(tmp_path):
    level, sides, sidx = cavity_grid()
    k_top, q_top, v_top = sides["top"][1][0]
    # a second velocity side, a pressure side and geometry links
    k_in = sidx[[10, 20, 30]]
    q_in = np.full((27, 3), -1.0)
    q_in[2] = 0.5
    v_in = np.array([[0.1, 0.2, 0.3], [0.0, 0.0, 0.0], [0.0, 0.0, 0.01]])
    k_out = sidx[[40, 50]]
    q_out = np.full((27, 2), -1.0)
    q_out[1] = 0.5
    p_out = np.array([[0.3, 0.6], [k_out[0] - 1, k_out[1] - 1]])
    q_geo = np.full((27, 2), -1.0)
    q_geo[[0, 5]] = 0.5
    sides = dict(sides, inlet=("velocity", [(k_in, q_in, v_in)]), outlet=("pressure", [(k_out, q_out, p_out)]))
    gf.write_grid(tmp_path, [level], sides, geometry=[(sidx[[60, 61]], q_geo)])
    lev, = gf.read_grid(tmp_path, velocity_ratio=2.0, pressure_factor=3.0)
    assert np.array_equal(lev.geo, level.geo) and np.array_equal(lev.neighbors[2], level.neighbors[2])
    k, q, vx, vy, vz = lev.boundaries["velocity"]
    # side order inlet, outlet, front, back, top, bottom
    assert np.array_equal(k, np.concatenate([k_in, k_top]))
    assert np.array_equal(vx, np.concatenate([v_in[0], v_top[0]]) / 2.0)
    assert np.array_equal(vz, np.concatenate([v_in[2], v_top[2]]) / 2.0)
    assert np.array_equal(q[:, :3], q_in) and np.array_equal(q[:, 3:], np.where(q_top > 0, q_top, -1.0))
    kp, qp, rho, kn = lev.boundaries["pressure"]
    assert np.array_equal(kp, k_out) and np.array_equal(rho, p_out[0] / 3.0) and np.array_equal(kn, k_out - 1)
    kg, qg = lev.boundaries["geometry"]
    assert np.array_equal(kg, sidx[[60, 61]]) and (qg[0] == 0).all() and (qg[5] == 0.5).all()
    assert "noSlip" in lev.boundaries and np.array_equal(lev.boundaries["noSlip"][0], sides["bottom"][1][0][0])
