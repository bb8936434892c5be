"""Precursor inflow planes in VirtualFluids' file format.

A precursor run records the inflow plane of a simulation over time
(PrecursorWriter::write, src/gpu/core/PreCollisionInteractor/PrecursorWriter.cu:305-343)
as VTK image data (WbWriterVtkXmlImageBinary::writeData,
src/basics/writer/WbWriterVtkXmlImageBinary.cpp:258-360): the image's x and y
axes are the simulation's y and z, its z axis is time (extent = time step
indices, spacing = time between records), one raw Float64 array per quantity
(vx, vy, vz or the nine positive-x distributions) appended after the XML
header.  Files are named <prefix>_lev_<level>_ID_<process>_File_<part>.bin.vti.

The reading side restates TransientBCSetter
(src/gpu/GridGenerator/TransientBCSetter/TransientBCSetter.{h,cpp}): VTKFile
(header parsed line by line as VTKFile::readHeader does, data loaded on first
use), VTKFileCollection (file discovery), VTKReader.fill_arrays (the 0PP,
0PM, 0MP, 0MM stencil and inverse-distance weights of each boundary node, the
read/write index lists) and VTKReader.get_next_data (the plane at a physical
time: nearest record of the file whose time range holds it, one file switch
per call).  Arithmetic runs in VirtualFluids' `real` (float32 by default,
float64 as with VF_DOUBLE_ACCURACY); the header's origin and spacing are read
as float like readStringToVector<float>.

Three changes.  Two are where the reference's lists go wrong off the
precursor's grid (for a plane whose points coincide with the boundary nodes,
the usual case, the results are the same): the reference appends the 0PM
neighbour to the 0PP list (TransientBCSetter.cpp:386), which leaves 0PM short
and makes the neighbour copy run past its array, here it goes to the 0PM
list; and VTKReader::getWriteIndex (:430-441) looks the file index up among
the write indices, so a point that two stencils share can resolve to another
point's slot, here it is looked up among the read indices.  The third: the
NaN check (VTKFile::markNANs, :163-175) reads into a vector it only reserved,
so it never reports; here the first record's values at the read indices are
checked.
"""
from __future__ import annotations

import os

import numpy as np

VTI_SUFFIX = ".bin.vti"


# ------------------------------------------------------------------ writer --
def _num(v, precision: int) -> str:
    """A float as std::ostream prints it with `precision` (general format)."""
    return format(float(np.float32(v)), f".{precision}g")


def write_vti(path, names, data, extent, origin, spacing, precision=8) -> str:
    """WbWriterVtkXmlImageBinary::writeData with point data only, whole extent =
    piece extent, as PrecursorWriter::write calls it.  data: (len(names),
    n_points) values; extent: 6 ints; origin, spacing: 3 floats."""
    data = np.ascontiguousarray(data, dtype=np.float64).reshape(len(names), -1)
    n = data.shape[1]
    ext = " ".join(str(int(e)) for e in extent)
    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" >\n',
            f'   <ImageData WholeExtent="{ext}" Origin="{" ".join(_num(o, precision) for o in origin)}" '
            f'Spacing="{" ".join(_num(s, precision) for s in spacing)}"> \n',
            f'      <Piece Extent="{ext}">\n']
    if names:
        head.append("         <PointData>\n")
        off = 0
        for name in names:
            head.append(f'            <DataArray type="Float64" Name="{name}" format="appended" offset="{off}" /> \n')
            off += 4 + 8 * n
        head.append("         </PointData>\n")
    head += ["      </Piece>\n", "   </ImageData>\n", '   <AppendedData encoding="raw">\n', "_"]
    with open(path, "wb") as fh:
        fh.write("".join(head).encode())
        for j in range(len(names)):
            fh.write(np.int32(8 * n).tobytes())
            fh.write(data[j].tobytes())
        fh.write(b"\n   </AppendedData>\n</VTKFile>\n")
    return str(path)


def file_name(prefix, level, pid, part) -> str:
    """PrecursorWriter::makeFileName + the writer's extension."""
    return f"{prefix}_lev_{level}_ID_{pid}_File_{part}{VTI_SUFFIX}"


# ------------------------------------------------------------------ reader --
def _element(line: str) -> str:
    start = line.find("<") + 1
    end = line.find(" ", start)
    return line[start:] if end < 0 else line[start:end]


def _attribute(line: str, name: str) -> str:
    start = line.find(name) + len(name) + 2  # skip '="'
    return line[start:line.find('"', start)]


def _numbers(text: str, to) -> list:
    # readStringToVector reads every number as float, then converts
    return [to(np.float32(v)) for v in text.split()]


class VTKFile:
    """One precursor file (VTKFile, TransientBCSetter.h:66-150)."""

    def __init__(self, filename, real=np.float32):
        self.filename = str(filename)
        self.real = np.dtype(real).type
        self.values = None
        self._read_header()

    def _read_header(self):
        R = self.real
        with open(self.filename, "rb") as fh:
            line = fh.readline().decode(errors="replace")  # VTKFile or the xml declaration
            if line[1:2] == "?":
                fh.readline()  # VTKFile
            line = fh.readline().decode(errors="replace")  # ImageData
            origin = _numbers(_attribute(line, "Origin"), np.float32)
            spacing = _numbers(_attribute(line, "Spacing"), np.float32)
            line = fh.readline().decode(errors="replace")  # Piece
            extent = _numbers(_attribute(line, "Extent"), int)
            fh.readline()  # PointData
            line = fh.readline().decode(errors="replace")
            self.quantities = []
            while _element(line.rstrip("\n")) == "DataArray":
                self.quantities.append([_attribute(line, "Name"), int(_attribute(line, "offset"))])
                line = fh.readline().decode(errors="replace")
            fh.readline()  # </Piece>
            fh.readline()  # </ImageData>
            fh.readline()  # <AppendedData>
            base = fh.tell() + 1 + 4  # the underscore and the first array's byte count
        for q in self.quantities:
            q[1] += base
        self.dx, self.dy, self.dz = R(spacing[0]), R(spacing[1]), R(spacing[2])
        self.nx = extent[1] - extent[0] + 1
        self.ny = extent[3] - extent[2] + 1
        self.nz = extent[5] - extent[4] + 1
        self.min_x = R(R(origin[0]) + self.dx * R(extent[0]))
        self.min_y = R(R(origin[1]) + self.dy * R(extent[2]))
        self.min_z = R(R(origin[2]) + self.dz * R(extent[4]))
        self.max_x = R(R(self.nx - 1) * self.dx + self.min_x)
        self.max_y = R(R(self.ny - 1) * self.dy + self.min_y)
        self.max_z = R(R(self.nz - 1) * self.dz + self.min_z)

    # geometry (TransientBCSetter.h:79-111)
    @property
    def n_points(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_quantities(self) -> int:
        return len(self.quantities)

    def in_x(self, x):
        return self.min_x <= x <= self.max_x

    def in_y(self, y):
        return self.min_y <= y <= self.max_y

    def in_z(self, z):
        return self.min_z <= z <= self.max_z

    def in_box(self, x, y, z):
        return self.in_x(x) and self.in_y(y) and self.in_z(z)

    def _idx(self, pos, lo, d) -> int:
        return int(np.trunc(self.real(self.real(pos) - lo) / d))

    def linear(self, ix, iy, iz) -> int:
        return ix + self.nx * (iy + self.ny * iz)

    def neighbor(self, x, y, z, px, py, pz) -> int:
        """findNeighborMMM ... PPP: the point px/py/pz steps above the cell
        corner below (x, y, z); -1 outside the file's points."""
        i = self.linear(self._idx(x, self.min_x, self.dx) + px, self._idx(y, self.min_y, self.dy) + py,
                        self._idx(z, self.min_z, self.dz) + pz)
        return i if 0 <= i < self.n_points else -1

    def x_of(self, i):
        return self.real(self.real(i % self.nx) * self.dx + self.min_x)

    def y_of(self, i):
        return self.real(self.real((i // self.nx) % self.ny) * self.dy + self.min_y)

    def closest_z(self, z) -> int:
        v = float(self.real(self.real(z) - self.min_z) / self.dz)
        r = int(np.sign(v) * np.floor(abs(v) + 0.5))  # std::round: halves away from zero
        return self.nz if r > self.nz else (0 if r < 0 else r)

    # data
    def load(self):
        self.values = [np.fromfile(self.filename, dtype="<f8", count=self.n_points, offset=off)
                       for _, off in self.quantities]

    def unload(self):
        self.values = None

    @property
    def loaded(self) -> bool:
        return self.values is not None

    def has_nan(self, read_idx) -> bool:
        """A NaN in the first record of the first quantity at the read indices."""
        first = np.fromfile(self.filename, dtype="<f8", count=max(1, self.nx * self.ny), offset=self.quantities[0][1])
        idx = np.asarray(read_idx, dtype=np.int64)
        idx = idx[idx < first.size]
        return bool(np.isnan(first[idx]).any())

    def get_data(self, data, n_nodes, read_idx, write_idx, off_read, off_write):
        if not self.loaded:
            self.load()
        r = np.asarray(read_idx, dtype=np.int64) + off_read
        w = np.asarray(write_idx, dtype=np.int64) + off_write
        for j, vals in enumerate(self.values):
            data[j * n_nodes + w] = vals[r].astype(data.dtype)


class VTKFileCollection:
    """All files of a prefix, files[level][id][part] (VTKFileCollection,
    TransientBCSetter.h:168-199, findFiles TransientBCSetter.cpp:189-227)."""

    def __init__(self, prefix, real=np.float32):
        self.prefix = str(prefix)
        self.files = []
        while True:
            level = []
            while True:
                parts = []
                while os.path.exists(file_name(self.prefix, len(self.files), len(level), len(parts))):
                    parts.append(VTKFile(file_name(self.prefix, len(self.files), len(level), len(parts)), real))
                if not parts:
                    break
                level.append(parts)
            if not level:
                break
            self.files.append(level)
        if not self.files:
            raise RuntimeError("VTKFileCollection found no files!")

    @property
    def n_quantities(self) -> int:
        return self.files[0][0][0].n_quantities


class VTKReader:
    """TransientBCInputFileReader for VTK precursor files (VTKReader,
    TransientBCSetter.h:216-233, TransientBCSetter.cpp:231-475)."""

    def __init__(self, collection: VTKFileCollection, read_level=0, real=np.float32):
        self.collection = collection
        self.level = int(read_level)
        self.real = np.dtype(real).type
        self.n_quantities = collection.n_quantities
        self.n_points = 0
        self.n_points_read = 0
        self.writing_offset = 0
        nl = len(collection.files)
        self.read_idx = [[[] for _ in collection.files[lev]] for lev in range(nl)]
        self.write_idx = [[[] for _ in collection.files[lev]] for lev in range(nl)]
        self.n_file = [[0 for _ in collection.files[lev]] for lev in range(nl)]
        self.neighbors = [[], [], [], []]  # 0PP, 0PM, 0MP, 0MM
        self.weights = [[], [], [], []]
        self.perfect_match = True

    def _write_index(self, level, fid, linear) -> int:
        """The plane slot of file point `linear`, added on first use.  (The
        reference searches the write indices for the file index here; read
        indices are searched, so a point that two stencils share gets one slot
        and no stencil points at another point's slot.)"""
        ri = self.read_idx[level][fid]
        try:
            return self.write_idx[level][fid][ri.index(linear)]
        except ValueError:
            self.write_idx[level][fid].append(self.n_points_read)
            ri.append(linear)
            self.n_points_read += 1
            return self.n_points_read - 1

    def fill_arrays(self, coords_y, coords_z):
        """Stencil and weights of every boundary node at (y, z)."""
        R = self.real
        max_diff, eps = R(1e-4), R(1e-7)
        cy = np.asarray(coords_y, dtype=R)
        cz = np.asarray(coords_z, dtype=R)
        self.n_points = cy.size
        PP, PM, MP, MM = range(4)
        level = self.level
        files = self.collection.files[level]

        def weight(f, idx, y, z):
            dy = R(f.x_of(idx) - y)
            dz = R(f.y_of(idx) - z)
            return R(R(1.0) / R(R(dy * dy + dz * dz) + eps))

        for y, z in zip(cy, cz):
            found = [False] * 4
            for fid, parts in enumerate(files):
                f = parts[0]
                if not f.in_box(y, z, R(0)):
                    continue
                idx = f.neighbor(y, z, R(0), 0, 0, 0)
                if idx != -1:
                    if abs(R(y - f.x_of(idx))) < max_diff and abs(R(z - f.y_of(idx))) < max_diff:
                        for s, w in zip((PP, PM, MP, MM), (R(1e6), R(0), R(0), R(0))):
                            self.weights[s].append(w)
                        wi = self._write_index(level, fid, idx)
                        for s in (PP, PM, MP, MM):
                            self.neighbors[s].append(wi)
                        found = [True] * 4
                    else:
                        self.perfect_match = False
                    if not found[MM]:
                        found[MM] = True
                        self.weights[MM].append(weight(f, idx, y, z))
                        self.neighbors[MM].append(self._write_index(level, fid, idx))
                for s, (px, py) in ((PP, (1, 1)), (PM, (1, 0)), (MP, (0, 1))):
                    if found[s]:
                        continue
                    i = f.neighbor(y, z, R(0), px, py, 0)
                    if i != -1:
                        found[s] = True
                        self.weights[s].append(weight(f, i, y, z))
                        self.neighbors[s].append(self._write_index(level, fid, i))
                if all(found):
                    break
            if not all(found):
                raise RuntimeError(f"VTKReader.fill_arrays: no precursor neighbours for the point y={y}, z={z}")
        for lev, parts_of in enumerate(self.collection.files):
            for fid, parts in enumerate(parts_of):
                if parts[0].has_nan(self.read_idx[lev][fid]):
                    raise RuntimeError("Found a NAN in the precursor where a velocity is needed")

    def stencil(self):
        """(neighbors (4, n) uint32, weights (4, n)) in the order 0PP, 0PM, 0MP, 0MM."""
        nb = np.array(self.neighbors, dtype=np.uint32).reshape(4, -1)
        w = np.array(self.weights, dtype=self.real).reshape(4, -1)
        return nb, w

    def get_next_data(self, data, n_nodes, time):
        """Fill data[q * n_nodes + writing_offset + write index] with the plane
        at physical time `time`."""
        level = self.level
        t = self.real(time)
        for fid, parts in enumerate(self.collection.files[level]):
            k = self.n_file[level][fid]
            if not parts[k].in_z(t):
                k += 1
                if k == len(parts):
                    raise RuntimeError("Not enough Precursor Files to read")
                parts[k - 1].unload()
            f = parts[k]
            off = f.closest_z(t) * f.nx * f.ny
            f.get_data(data, n_nodes, self.read_idx[level][fid], self.write_idx[level][fid], off, self.writing_offset)
            self.n_file[level][fid] = k


def precursor_from_files(prefix, level, coords_y, coords_z, real=np.float32):
    """Stencil, weights, plane size, quantity count and a plane reader
    (reader(physical_time) -> (quantities, plane points) array) for
    PrecursorPlanes, from the precursor files of `prefix`."""
    rd = VTKReader(VTKFileCollection(prefix, real), level, real)
    rd.fill_arrays(coords_y, coords_z)
    nb, w = rd.stencil()
    n, q = rd.n_points_read, rd.n_quantities

    def reader(t):
        data = np.zeros(q * n, dtype=real)
        rd.get_next_data(data, n, t)
        return data.reshape(q, n)

    return nb, w, n, q, reader
