"""VirtualFluids' grid files: the folder a GridGenerator run writes
(SimulationFileWriter, src/gpu/GridGenerator/io/SimulationFileWriter/
SimulationFileWriter.cpp:61-789, names SimulationFileNames.cpp:36-80) and the
GPU solver reads with GridReader (src/gpu/core/DataStructureInitializer/
GridReaderFiles/GridReader.cpp:48-929).

Formats (ASCII; ``binary`` as in the reference for the coordinate, neighbour
and geo files only):

* coordX/Y/Z, neighborX/Y/Z, neighborWSB, geoVec (``.dat``):
  ``maxLevel``, then per level ``size`` and size + 1 values -- entry 0 the
  dummy node (coordinate 0, neighbour 0, geo 16), entries 1..size the sparse
  nodes; neighbour indices are 1-based with 0 the dummy node
  (CoordNeighborGeoV.cpp:63-150);
* <side>BoundaryQs (inlet, outlet, top, bottom, front, back, geom):
  ``maxLevel``, per level ``size`` and per entry ``index key q...`` -- bit
  (26 - d) of key set for every direction d with a link, the q of those
  links from d = 26 down to 0 (BoundaryQs.cpp:98-170);
* <side>BoundaryValues: the condition name (velocity, noSlip, pressure,
  outflow, slip, ...), ``maxLevel``, per level ``size`` and per entry the index
  and the condition's columns -- 3 (vx vy vz) for velocity and noSlip, 2
  (rho, neighbour) for pressure (BoundaryValues.cpp:130-200).

GridReader turns them into the solver's arrays: every side whose condition is
``velocity`` contributes its entries (in side order inlet, outlet, front,
back, top, bottom) to one velocity boundary, velocities divided by the
velocity ratio; a ``pressure`` side fills the pressure boundary (RhoBC / the
pressure factor, kN); ``noSlip`` sides the no-slip boundary; geomBoundaryQs
(always ASCII) the geometry boundary with q of d000 set to 0
(GridReader.cpp:124-160, 237-350, 562-720).  Quirks kept: a velocity side
contributes only on levels with more than one entry (:272, :302), several
pressure or no-slip sides overwrite each other (the last one in side order
is kept).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

SIDES = ("inlet", "outlet", "front", "back", "top", "bottom")  # GridReader.cpp:62-67
# the order SimulationFileWriter writes the side files in (sides MX, PX, PZ, MZ, MY, PY, GEOMETRY)
WRITER_SIDES = ("inlet", "outlet", "top", "bottom", "front", "back", "geom")
GEO_FLUID = 19
GEO_STOPPER = 16
Q_COLUMNS = 27


def _tokens(path):
    with open(path, "r") as fh:
        return fh.read().split()


class _Stream:
    """The two std::istream reads the readers use: operator>> (skip white
    space, read a token) and getline (the rest of the current line)."""

    def __init__(self, path):
        with open(path, "r") as fh:
            self.text = fh.read()
        self.pos = 0

    def read(self) -> str:
        t, n = self.text, len(self.text)
        while self.pos < n and t[self.pos].isspace():
            self.pos += 1
        start = self.pos
        while self.pos < n and not t[self.pos].isspace():
            self.pos += 1
        if start == self.pos:
            raise ValueError("unexpected end of file")
        return t[start:self.pos]

    def getline(self):
        end = self.text.find("\n", self.pos)
        self.pos = len(self.text) if end < 0 else end + 1


# ------------------------------------------------------------------ readers --
def read_coord_neighbor_geo(path, binary=False, coord=False):
    """Per level the size + 1 values of a coordinate (float64) or neighbour /
    geo (uint32) file."""
    dtype = np.float64 if coord else np.uint32
    if not binary:
        t = _tokens(path)
        pos = 0
        max_level = int(t[pos])
        pos += 1
        out = []
        for _ in range(max_level + 1):
            size = int(t[pos])
            pos += 1
            vals = t[pos:pos + size + 1]
            if len(vals) != size + 1:
                raise ValueError(f"{path}: level with {size} nodes is cut short")
            pos += size + 1
            out.append(np.array([float(v) for v in vals], dtype) if coord else
                       np.array([int(v) for v in vals], dtype=np.int64).astype(dtype))
        return out
    # binary (CoordNeighborGeoV::init_Binary): "maxLevel\n", "size\n", the
    # dummy entry as text, one separator byte, then size raw values; the
    # next level's size follows a separator byte
    raw = open(path, "rb").read()
    pos = 0

    def token():
        nonlocal pos
        while pos < len(raw) and raw[pos:pos + 1].isspace():
            pos += 1
        start = pos
        while pos < len(raw) and not raw[pos:pos + 1].isspace():
            pos += 1
        return raw[start:pos].decode()

    max_level = int(token())
    size = int(token())
    out = []
    item = 8 if coord else 4
    for level in range(max_level + 1):
        first = float(token()) if coord else int(token())
        pos += 1  # skipSpace: one byte
        vals = np.frombuffer(raw, "<f8" if coord else "<u4", size, pos)
        pos += size * item
        out.append(np.concatenate([np.array([first], dtype), vals.astype(dtype)]))
        if level == max_level:
            break
        pos += 1
        size = int(token())
    return out


def read_boundary_qs(path):
    """Per level (indices int64, q27 (27, n) float64, -1 where no link).
    After an entry's q values the rest of its line is skipped."""
    st = _Stream(path)
    max_level = int(st.read())
    out = []
    for _ in range(max_level + 1):
        n = int(st.read())
        idx = np.zeros(n, np.int64)
        q = np.full((Q_COLUMNS, n), -1.0)
        for e in range(n):
            idx[e] = int(st.read())
            key = int(st.read())
            column = 26
            while key:
                if key & 1:
                    q[column, e] = float(st.read())
                key >>= 1
                column -= 1
            st.getline()
        out.append((idx, q))
    return out


def _columns(name):
    return {"velocity": 3, "noSlip": 3, "pressure": 2, "processor": 0, "concentration": 0}.get(name, -1)


def read_boundary_values(path):
    """(condition name, per level (indices int64, values (columns, n)))."""
    st = _Stream(path)
    name = st.read()
    max_level = int(st.read())
    cols = _columns(name)
    if cols == -1:
        # initalVectorsWithSingleZero: one entry, index 0, value 0
        return name, [(np.zeros(1, np.int64), np.zeros((1, 1)))]
    out = []
    for _ in range(max_level + 1):
        n = int(st.read())
        idx = np.zeros(n, np.int64)
        vals = np.zeros((cols, n))
        if n == 0:
            st.getline()
        for e in range(n):
            idx[e] = int(st.read())
            vals[:, e] = [float(st.read()) for _ in range(cols)]
        out.append((idx, vals))
    return name, out


# ------------------------------------------------------------- the grid ----
@dataclass
class GridLevel:
    coords: tuple          # (x, y, z), size + 1 each
    neighbors: tuple       # (x, y, z), 1-based, 0 = dummy node
    neighbor_wsb: np.ndarray
    geo: np.ndarray        # typeOfGridNode
    boundaries: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.geo.size)


def read_grid(folder, binary=False, velocity_ratio=1.0, pressure_factor=1.0):
    """The levels of a grid folder with the boundary tables GridReader builds:
    ``boundaries`` maps "velocity" to (k, q27, vx, vy, vz), "pressure" to
    (k, q27, rho, kN), "outflow" to (k, q27, rho, kN), "noSlip" to (k, q27),
    "geometry" to (k, q27) -- each only where the files give entries."""
    f = lambda name: os.path.join(folder, name)  # noqa: E731
    cx = read_coord_neighbor_geo(f("coordX.dat"), binary, True)
    cy = read_coord_neighbor_geo(f("coordY.dat"), binary, True)
    cz = read_coord_neighbor_geo(f("coordZ.dat"), binary, True)
    nx = read_coord_neighbor_geo(f("neighborX.dat"), binary)
    ny = read_coord_neighbor_geo(f("neighborY.dat"), binary)
    nz = read_coord_neighbor_geo(f("neighborZ.dat"), binary)
    nw = read_coord_neighbor_geo(f("neighborWSB.dat"), binary)
    geo = read_coord_neighbor_geo(f("geoVec.dat"), binary)
    levels = [GridLevel((cx[l], cy[l], cz[l]), (nx[l], ny[l], nz[l]), nw[l], geo[l]) for l in range(len(cx))]
    # channel sides (GridReader::allocArrays_BoundaryValues / allocArrays_BoundaryQs)
    values = {s: read_boundary_values(f(f"{s}BoundaryValues.dat")) for s in SIDES}
    qs = {s: read_boundary_qs(f(f"{s}BoundaryQs.dat")) for s in SIDES}
    for lev, L in enumerate(levels):
        vel_k, vel_q, vel_v = [], [], []
        for s in SIDES:
            name, vals = values[s]
            if lev >= len(vals):
                continue
            if name == "velocity":
                idx_v, v = vals[lev]
                if idx_v.size > 1:
                    vel_v.append(v / velocity_ratio)
                if lev < len(qs[s]) and qs[s][lev][0].size > 0:
                    vel_k.append(qs[s][lev][0])
                    vel_q.append(qs[s][lev][1])
            elif name in ("pressure", "outflow"):
                idx_v, v = vals[lev]
                if idx_v.size > 0 and lev < len(qs[s]) and qs[s][lev][0].size > 0:
                    rho = v[0] / pressure_factor
                    if name == "outflow":
                        rho = rho * 0.0
                    L.boundaries[name] = (qs[s][lev][0], qs[s][lev][1], rho, v[1].astype(np.int64))
            elif name == "noSlip":
                if lev < len(qs[s]) and qs[s][lev][0].size > 0:
                    L.boundaries["noSlip"] = (qs[s][lev][0], qs[s][lev][1])
        if vel_k:
            k = np.concatenate(vel_k)
            v = np.concatenate(vel_v, axis=1) if vel_v else np.zeros((3, 0))
            if v.shape[1] == k.size and k.size > 1:
                L.boundaries["velocity"] = (k, np.concatenate(vel_q, axis=1), v[0], v[1], v[2])
    gpath = f("geomBoundaryQs.dat")
    if os.path.exists(gpath):
        for lev, (idx, q) in enumerate(read_boundary_qs(gpath)):
            if lev < len(levels) and idx.size > 0:
                q = q.copy()
                q[0] = 0.0  # modifyQElement: q of d000 set to 0
                levels[lev].boundaries["geometry"] = (idx, q)
    return levels


def lattice(level: GridLevel, storage="fp64", omega=1.0, device="cuda", **kw):
    """A sparse-layout Lattice over one level of a grid folder (neighbours and
    typeOfGridNode as read; node 0 the dummy node)."""
    from .lattice import Lattice

    nbx, nby, nbz = level.neighbors
    return Lattice(storage=storage, omega=omega, neighbors=(nbx, nby, nbz), node_types=level.geo, device=device, **kw)


# ------------------------------------------------------------------ writer --
def _write_cng(path, levels, binary, coord, dummy):
    with open(path, "wb") as fh:
        fh.write(f"{len(levels) - 1}\n".encode())
        for a in levels:
            size = a.size - 1
            if binary:
                fh.write(f"{size}\n{dummy} ".encode())
                fh.write(np.ascontiguousarray(a[1:], "<f8" if coord else "<u4").tobytes())
                fh.write(b"\n")
            else:
                fh.write(f"{size}\n{dummy} \n".encode())
                body = "".join(f"{float(v):g}\n" if coord else f"{int(v)}\n" for v in a[1:])
                fh.write(body.encode())
                fh.write(b"\n")


def _write_qs(fh, idx, q):
    fh.write(f"{idx.size}\n")
    for e in range(idx.size):
        key = 0
        vals = []
        for d in range(26, -1, -1):
            if q[d, e] > 0:
                key += 1 << (26 - d)
                vals.append(f"{q[d, e]:.16f}")
        fh.write(f"{int(idx[e])} {key} " + "".join(v + " " for v in vals) + "\n")


def write_grid(folder, levels, sides=None, geometry=None, binary=False):
    """A grid folder in SimulationFileWriter's layout.

    levels: GridLevel-like objects (coords, neighbors, neighbor_wsb, geo with
    entry 0 the dummy node); sides: {side: (name, [per level (k, q27, values
    (columns, n))])} for the six channel sides (missing: noSlip with no
    entries); geometry: [per level (k, q27)] for geomBoundaryQs."""
    os.makedirs(folder, exist_ok=True)
    f = lambda name: os.path.join(folder, name)  # noqa: E731
    for i, name in enumerate(("coordX", "coordY", "coordZ")):
        _write_cng(f(f"{name}.dat"), [L.coords[i] for L in levels], binary, True, 0)
    for i, name in enumerate(("neighborX", "neighborY", "neighborZ")):
        _write_cng(f(f"{name}.dat"), [L.neighbors[i] for L in levels], binary, False, 0)
    _write_cng(f("neighborWSB.dat"), [L.neighbor_wsb for L in levels], binary, False, 0)
    _write_cng(f("geoVec.dat"), [L.geo for L in levels], binary, False, GEO_STOPPER)
    sides = dict(sides or {})
    nl = len(levels)
    for s in WRITER_SIDES:
        if s == "geom":
            name, per = "noSlip", [(k, q, np.zeros((3, len(k)))) for k, q in (geometry or [])]
        else:
            name, per = sides.get(s, ("noSlip", []))
        with open(f(f"{s}BoundaryQs.dat"), "w") as fq, open(f(f"{s}BoundaryValues.dat"), "w") as fv:
            fv.write(f"{name}\n{nl - 1}\n")
            fq.write(f"{nl - 1}\n")
            for lev in range(nl):
                if lev >= len(per):
                    fv.write("0\n")
                    fq.write("0\n")
                    continue
                k, q, vals = per[lev]
                k = np.asarray(k, np.int64)
                _write_qs(fq, k, np.asarray(q, np.float64))
                fv.write(f"{k.size}\n")
                vals = np.asarray(vals, np.float64).reshape(-1, k.size)
                for e in range(k.size):
                    fv.write(f"{int(k[e])} " + "".join(f"{float(v):g} " for v in vals[:, e]) + "\n")
