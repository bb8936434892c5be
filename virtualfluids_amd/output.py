"""VTK output in VirtualFluids' format (the GPU solver's FileWriter).

What a VirtualFluids GPU run writes at an output step
(src/gpu/core/Output/FileWriter.cpp:102-330):

* per level, the nodes split into parts of at most 30 million
  (FilePartCalculator.cpp:40-58); part i of level l, process ID and time step
  t is ``<prefix>_bin_lev_<l>_ID_<ID>_Part_<i>_t_<t>.vtk.bin.vtu``, an
  unstructured grid written by WbWriterVtkXmlBinary::writeOctsWithNodeData
  (src/basics/writer/WbWriterVtkXmlBinary.cpp:1306-1425): Float32 points,
  one voxel (VTK type 11) per fluid node whose oct of seven + neighbours is
  all GEO_FLUID, in this part and not wrapped around a periodic boundary
  (WriterUtilities.cpp:44-86), and Float64 point data press, rho, vx1, vx2,
  vx3, geo (+ nut with a turbulence model), raw appended binary;
* ``<prefix>_bin_ID_<ID>_t_<t>.vtk.pvtu`` listing the parts
  (writeParallelFile, :141-177);
* at the last step, ``<prefix>_bin_ID_<ID>.pvd`` with the collection files of
  every written step (writeCollectionForTimeSeries, :88-99);
* with time averages (Parameter::getCalcMean), ``<prefix>_bin_mean_...`` parts
  of pressMed, rhoMed, vx1Med, vx2Med, vx3Med, geo (writeUnstructuredGridMeanLT,
  FileWriter.cpp:332-431; the means scaled in the real type) and their .pvtu.

Quirks kept so the files match: a node that is not GEO_FLUID keeps the point
coordinates and data a previous part left at its slot (zero in the first
part), and an oct may reference the first node after its part (the bound
test is ``index > end``).  Both only matter for levels above 30 million nodes.
"""
from __future__ import annotations

import os
import struct

import numpy as np

LIMIT_OF_NODES_FOR_VTK = 30_000_000  # FilePartCalculator.h:53
GEO_FLUID = 19
VTK_VOXEL = 11


# ------------------------------------------------------- FilePartCalculator --
def number_of_parts(number_of_nodes: int) -> int:
    return number_of_nodes // LIMIT_OF_NODES_FOR_VTK + 1


def nodes_in_part(number_of_nodes: int, part: int) -> int:
    last = number_of_parts(number_of_nodes) - 1
    if part > last:
        raise RuntimeError("The number of nodes for a non-existing part can not be calculated")
    return number_of_nodes - part * LIMIT_OF_NODES_FOR_VTK if part == last else LIMIT_OF_NODES_FOR_VTK


def start_of_part(part: int) -> int:
    return part * LIMIT_OF_NODES_FOR_VTK


# -------------------------------------------------------------- file names --
def part_file_name(prefix, level, ID, part, timestep, mean=False) -> str:
    return f"{prefix}_bin{'_mean' if mean else ''}_lev_{level}_ID_{ID}_Part_{part}_t_{timestep}.vtk"


def collection_file_name(prefix, ID, timestep, mean=False) -> str:
    return f"{prefix}_bin{'_mean' if mean else ''}_ID_{ID}_t_{timestep}.vtk"


def pvd_file_name(prefix, ID) -> str:
    return f"{prefix}_bin_ID_{ID}"


def _open(path):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    return open(path, "wb")


def _g(x) -> str:
    """std::ostream's default formatting of a double (six significant digits)."""
    return f"{float(x):g}"


# --------------------------------------------------------- the VTK writers --
def write_octs_with_node_data(filename, nodes, cells, datanames, nodedata) -> str:
    """nodes: (n, 3) float32; cells: (m, 8) node indices in oct order (base,
    +x, +x+y, +y, +z, +x+z, +x+y+z, +y+z); nodedata: one length-n float64
    array per name.  Returns the file name (filename + '.bin.vtu')."""
    path = filename + ".bin.vtu"
    nodes = np.ascontiguousarray(nodes, dtype=np.float32).reshape(-1, 3)
    cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1, 8)
    n, m = nodes.shape[0], cells.shape[0]
    b_points, b_conn, b_off, b_types, b_scalar = 12 * n, 32 * m, 4 * m, m, 8 * n
    head = []
    offset = 0
    head.append('<?xml version="2.0"?>\n')
    head.append('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" >\n')
    head.append("   <UnstructuredGrid>\n")
    head.append(f'      <Piece NumberOfPoints="{n}" NumberOfCells="{m}">\n')
    head.append("         <Points>\n")
    head.append(f'            <DataArray type="Float32" NumberOfComponents="3" format="appended" offset="{offset}"  />\n')
    head.append("         </Points>\n")
    offset += 4 + b_points
    head.append("         <Cells>\n")
    head.append(f'            <DataArray type="Int32" Name="connectivity" format="appended" offset="{offset}" />\n')
    offset += 4 + b_conn
    head.append(f'            <DataArray type="Int32" Name="offsets" format="appended" offset="{offset}" />\n')
    offset += 4 + b_off
    head.append(f'            <DataArray type="UInt8" Name="types" format="appended" offset="{offset}" />\n ')
    offset += 4 + b_types
    head.append("         </Cells>\n")
    head.append("         <PointData>\n")
    for name in datanames:
        head.append(f'            <DataArray type="Float64" Name="{name}" format="appended" offset="{offset}" /> \n')
        offset += 4 + b_scalar
    head.append("         </PointData>\n")
    head.append("      </Piece>\n")
    head.append("   </UnstructuredGrid>\n")
    head.append('   <AppendedData encoding="raw">\n')
    head.append("_")
    i32 = lambda v: struct.pack("<i", int(v))  # noqa: E731
    with _open(path) as fh:
        fh.write("".join(head).encode())
        fh.write(i32(b_points))
        fh.write(nodes.tobytes())
        fh.write(i32(b_conn))
        fh.write(cells[:, [0, 1, 3, 2, 4, 5, 7, 6]].astype("<u4").tobytes())
        fh.write(i32(b_off))
        fh.write((8 * np.arange(1, m + 1, dtype=np.int64)).astype("<i4").tobytes())
        fh.write(i32(b_types))
        fh.write(np.full(m, VTK_VOXEL, np.uint8).tobytes())
        for arr in nodedata:
            fh.write(i32(b_scalar))
            fh.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
        fh.write(b"\n</AppendedData>\n</VTKFile>\n")
    return path


def write_parallel_file(filename, piece_sources, point_data_names, cell_data_names=()) -> str:
    path = filename + ".pvtu"
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="PUnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
           '  <PUnstructuredGrid GhostLevel="1">\n', "    <PPointData>\n"]
    out += [f'      <PDataArray type="Float64" Name="{s}"/>\n' for s in point_data_names]
    out.append("    </PPointData>\n")
    if cell_data_names:
        out.append("    <PCellData>\n")
        out += [f'      <PDataArray type="Float64" Name="{s}"/>\n' for s in cell_data_names]
        out.append("    </PCellData>\n")
    out += ["    <PPoints>\n", '      <PDataArray type="Float32" NumberOfComponents="3"/>\n', "    </PPoints>\n"]
    out += [f'    <Piece Source="{s}"/>\n' for s in piece_sources]
    out += ["  </PUnstructuredGrid>\n", "</VTKFile>\n"]
    with _open(path) as fh:
        fh.write("".join(out).encode())
    return path


def write_collection_for_time_series(filename, files_by_step, separate_groups=False) -> str:
    path = filename + ".pvd"
    out = ['<VTKFile type="Collection" version="0.1" byte_order="LittleEndian" >\n', "   <Collection>\n"]
    for step in sorted(files_by_step):
        group = part = 0
        for f in files_by_step[step]:
            out.append(f'       <DataSet timestep="{_g(step)}" group="{group}" part="{part}" file="{f}"/>\n')
            if separate_groups:
                group += 1
            else:
                part += 1
    out.append("   </Collection>\n</VTKFile>")
    with _open(path) as fh:
        fh.write("".join(out).encode())
    return path


# ---------------------------------------------------------- WriterUtilities --
def is_periodic_cell(coords, grid_spacing, base, other) -> bool:
    """WriterUtilities::isPeriodicCell: the oct's space diagonal is longer than
    1.01 times the lattice diagonal."""
    x, y, z = coords
    d = np.sqrt((x[other] - x[base]) ** 2 + (y[other] - y[base]) ** 2 + (z[other] - z[base]) ** 2)
    return d > 1.01 * np.sqrt(3.0 * grid_spacing ** 2)


def oct_indices(base, nbx, nby, nbz):
    """WriterUtilities::getIndicesOfAllNodesInOct (arrays of bases)."""
    i0 = np.asarray(base)
    i1 = nbx[i0]
    i2 = nby[i1]
    i3 = nby[i0]
    return np.stack([i0, i1, i2, i3, nbz[i0], nbz[i1], nbz[i2], nbz[i3]], axis=-1)


def unstructured_grid_parts(coords, types, neighbors, grid_spacing, node_data):
    """Points, cells and point data of every part of one level
    (FileWriter::writeUnstructuredGridLT, FileWriter.cpp:222-330).

    coords: (x, y, z) arrays; types: typeOfGridNode; neighbors: (nbx, nby, nbz);
    node_data: list of arrays already scaled to the written values.  Yields
    (points, cells, data) per part."""
    types = np.asarray(types, dtype=np.uint32)
    nbx, nby, nbz = (np.asarray(a, dtype=np.int64) for a in neighbors)
    x, y, z = (np.asarray(a) for a in coords)
    n_nodes = types.size
    nodes = np.zeros((0, 3), np.float32)
    data = [np.zeros(0) for _ in node_data]
    for part in range(number_of_parts(n_nodes)):
        size = nodes_in_part(n_nodes, part)
        start = start_of_part(part)
        end = start + size
        # std::vector::resize keeps what a previous part left
        grown = np.zeros((size, 3), np.float32)
        keep = min(size, nodes.shape[0])
        grown[:keep] = nodes[:keep]
        nodes = grown
        for i in range(len(data)):
            g = np.zeros(size)
            g[:keep] = data[i][:keep]
            data[i] = g
        pos = np.arange(start, end)
        fl = pos[types[start:end] == GEO_FLUID]
        rel = fl - start
        nodes[rel, 0] = np.asarray(x[fl], np.float32)
        nodes[rel, 1] = np.asarray(y[fl], np.float32)
        nodes[rel, 2] = np.asarray(z[fl], np.float32)
        for i, arr in enumerate(node_data):
            data[i][rel] = np.asarray(arr, np.float64)[fl]
        octs = oct_indices(fl, nbx, nby, nbz) if fl.size else np.zeros((0, 8), np.int64)
        if fl.size:
            a, b = octs[:, 0], octs[:, 6]
            diag = np.sqrt((x[b] - x[a]) ** 2 + (y[b] - y[a]) ** 2 + (z[b] - z[a]) ** 2)
            periodic = diag > 1.01 * np.sqrt(3.0 * grid_spacing ** 2)
            valid = np.all(types[octs] == GEO_FLUID, axis=1) & ~np.any(octs > end, axis=1)
            octs = octs[~periodic & valid]
        yield nodes.copy(), octs - start, [d.copy() for d in data]


class FileWriter:
    """FileWriter (FileWriter.cpp:102-220) over host copies of one or more
    levels.  Each level is a dict with coords (x, y, z), types, neighbors
    (nbx, nby, nbz), grid_spacing and the fields rho, vx, vy, vz, press (and
    nut when the run has a turbulence model), as the lattice computes them
    (``Lattice.macroscopic_quantities``)."""

    def __init__(self, prefix, ID=0, density_ratio=1.0, velocity_ratio=1.0, viscosity_ratio=1.0):
        self.prefix = str(prefix)
        self.ID = int(ID)
        self.density_ratio = float(density_ratio)
        self.velocity_ratio = float(velocity_ratio)
        self.viscosity_ratio = float(viscosity_ratio)
        self.collection = []
        self.collection_mean = []
        self.time_series = {}

    def node_data_names(self, turbulent):
        names = ["press", "rho", "vx1", "vx2", "vx3", "geo"]
        return names + ["nut"] if turbulent else names

    @staticmethod
    def mean_node_data_names():
        """FileWriter::getMeanNodeDataNames (FileWriter.cpp:197-210)."""
        return ["pressMed", "rhoMed", "vx1Med", "vx2Med", "vx3Med", "geo"]

    @staticmethod
    def _real(lev):
        """The level's real type (that of its fields): Parameter's ratios are
        real, so they are rounded to it before any use."""
        dt = np.asarray(lev["rho"]).dtype
        return dt.type if dt in (np.float32, np.float64) else np.float64

    def _scaled(self, lev):
        R = self._real(lev)
        dr, vr = float(R(self.density_ratio)), float(R(self.velocity_ratio))
        # (double)p / 3.0 * densityRatio * velocityRatio * velocityRatio, left to right
        out = [np.asarray(lev[k], np.float64) / 3.0 * dr * vr * vr for k in ("press", "rho")]
        out += [np.asarray(lev[k], np.float64) * vr for k in ("vx", "vy", "vz")]
        out.append(np.asarray(lev["types"], np.float64))
        if lev.get("nut") is not None:
            out.append(np.asarray(lev["nut"], np.float64) * float(R(lev.get("viscosity_ratio", self.viscosity_ratio))))
        return out

    def _scaled_mean(self, lev, mean):
        """writeUnstructuredGridMeanLT (FileWriter.cpp:332-431): the means in the
        real type, meanPressureOut / 3.0f * densityRatio * velocityRatio *
        velocityRatio and meanVelocityOut * velocityRatio, then to double."""
        R = self._real(lev)
        dr, vr = R(self.density_ratio), R(self.velocity_ratio)
        out = [((np.asarray(mean[k], R) / R(3.0)) * dr * vr * vr).astype(np.float64) for k in ("press", "rho")]
        out += [(np.asarray(mean[k], R) * vr).astype(np.float64) for k in ("vx", "vy", "vz")]
        out.append(np.asarray(lev["types"], np.float64))
        return out

    def write_level_mean(self, lev, mean, level, timestep):
        """The _bin_mean parts of one level; mean: rho, vx, vy, vz, press
        (MeanValues.mean)."""
        names = self.mean_node_data_names()
        files = []
        parts = unstructured_grid_parts(lev["coords"], lev["types"], lev["neighbors"], lev["grid_spacing"],
                                        self._scaled_mean(lev, mean))
        for i, (points, cells, data) in enumerate(parts, start=1):
            name = part_file_name(self.prefix, level, self.ID, i, timestep, mean=True)
            files.append(write_octs_with_node_data(name, points, cells, names, data))
        for f in files:
            self.collection_mean.append(os.path.basename(f))
        return files

    def write_level(self, lev, level, timestep):
        names = self.node_data_names(lev.get("nut") is not None)
        files = []
        parts = unstructured_grid_parts(lev["coords"], lev["types"], lev["neighbors"], lev["grid_spacing"],
                                        self._scaled(lev))
        for i, (points, cells, data) in enumerate(parts, start=1):
            name = part_file_name(self.prefix, level, self.ID, i, timestep)
            files.append(write_octs_with_node_data(name, points, cells, names, data))
        for f in files:
            self.collection.append(os.path.basename(f))
        return files

    def write_timestep(self, levels, timestep, last=False, means=None):
        """Every level's parts, the step's .pvtu and, with last=True, the .pvd;
        with means (one MeanValues.mean dict per level, Parameter::getCalcMean)
        also the _bin_mean parts and their .pvtu (FileWriter.cpp:102-140)."""
        turbulent = any(lev.get("nut") is not None for lev in levels)
        for level, lev in enumerate(levels):
            self.write_level(lev, level, timestep)
            if means is not None:
                self.write_level_mean(lev, means[level], level, timestep)
        self.time_series[timestep] = list(self.collection)
        if last:
            write_collection_for_time_series(pvd_file_name(self.prefix, self.ID), self.time_series)
        path = write_parallel_file(collection_file_name(self.prefix, self.ID, timestep), self.collection,
                                   self.node_data_names(turbulent))
        self.collection = []
        if means is not None:
            write_parallel_file(collection_file_name(self.prefix, self.ID, timestep, mean=True), self.collection_mean,
                                self.mean_node_data_names())
            self.collection_mean = []
        return path


def structured_level(lattice, even=True, origin=(0.0, 0.0, 0.0), spacing=1.0):
    """Host copies of a structured (direct) Lattice in the shape FileWriter
    takes: coordinates origin + index * spacing, typeOfGridNode GEO_FLUID for
    fluid nodes and GEO_SOLID (15) for solid ones, the periodic +x/+y/+z
    neighbours, and the output fields of ``Lattice.macroscopic_quantities``."""
    nx, ny, nz = lattice.dims
    k = np.arange(nx * ny * nz)
    xi, yi, zi = k % nx, (k // nx) % ny, k // (nx * ny)
    coords = tuple(o + spacing * i for o, i in zip(origin, (xi, yi, zi)))
    nbx = ((xi + 1) % nx) + nx * (yi + ny * zi)
    nby = xi + nx * (((yi + 1) % ny) + ny * zi)
    nbz = xi + nx * (yi + ny * ((zi + 1) % nz))
    types = np.full(k.size, GEO_FLUID, np.uint32)
    if lattice.solid_mask is not None:
        bits = lattice.solid_mask.cpu().numpy().view(np.uint32)
        solid = ((bits[k >> 5] >> (k & 31).astype(np.uint32)) & 1).astype(bool)
        types[solid] = 15
    rho, vx, vy, vz, press = (t.cpu().numpy() for t in lattice.macroscopic_quantities(even))
    lev = {"coords": coords, "types": types, "neighbors": (nbx, nby, nbz), "grid_spacing": spacing,
           "rho": rho, "vx": vx, "vy": vy, "vz": vz, "press": press}
    if getattr(lattice, "nut", None) is not None:
        lev["nut"] = lattice.nut.cpu().numpy()
    return lev


# ------------------------------------------------------------- run metadata --
def nups(number_of_time_steps, nodes_per_level, seconds) -> float:
    """PerformanceMeasurement::log (src/gpu/core/Output/PerformanceMeasurement.cpp:40-64):
    node updates per second, level l counted 2^l times per coarse step."""
    corrected = sum(n * 2.0 ** level for level, n in enumerate(nodes_per_level))
    return number_of_time_steps * corrected / seconds


def metadata(name, nodes_per_level, number_of_time_steps, runtime_seconds, collision_kernel="K17CompressibleNavierStokes",
             reynolds_number=0.0, lb_velocity=0.0, lb_viscosity=0.0, quadric_limiters=(0.01, 0.01, 0.01),
             world_length=1.0, world_velocity=0.0, dx=1.0, dt=1.0, start_date_time="", gpus=()) -> dict:
    """The run record VirtualFluids writes to <prefix>.yaml after a simulation
    (vf::basics::MetaData, src/basics/MetaData/MetaData.h:45-120; keys and
    order of YAML_MetaData.cpp:40-160; createMetaData, MetaDataCreator.cpp:40-90)."""
    return {
        "name": name,
        "number_of_processes": 1,
        "number_of_threads": 1,
        "vf_hardware": "GPU",
        "simulation": {
            "startDateTime": start_date_time,
            "runtimeSeconds": float(runtime_seconds),
            "nups": nups(number_of_time_steps, nodes_per_level, runtime_seconds),
            "numberOfTimeSteps": int(number_of_time_steps),
            "collisionKernel": collision_kernel,
            "reynoldsNumber": float(reynolds_number),
            "lb_velocity": float(lb_velocity),
            "lb_viscosity": float(lb_viscosity),
            "quadricLimiter": [float(x) for x in quadric_limiters],
        },
        "world": {"length": float(world_length), "velocity": float(world_velocity)},
        "discretization": {
            "dt": float(dt),
            "dx": float(dx),
            "totalNumberOfNodes": float(sum(nodes_per_level)),
            "numberOfLevels": len(nodes_per_level),
            "numberOfNodesPerLevel": [int(n) for n in nodes_per_level],
        },
        "build_info": {"precision": "mixed (FP64 / FP32 / FP16 storage)", "compiler": "hipcc", "build_type": "Release"},
        "gpus": [{"name": g, "compute_capability": "gfx950"} for g in gpus],
    }


def write_yaml(meta, filename) -> str:
    """vf::basics::writeYAML: the record as a YAML block mapping."""
    import yaml

    path = str(filename)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as fh:
        yaml.safe_dump(meta, fh, sort_keys=False, default_flow_style=False)
    return path


# ------------------------------------------------------- reading (tests) --
def read_vtu(path):
    """Points, connectivity (file order), offsets, types and point data of a
    file written by write_octs_with_node_data."""
    raw = open(path, "rb").read()
    head_end = raw.index(b'<AppendedData encoding="raw">\n_') + len(b'<AppendedData encoding="raw">\n_')
    head = raw[:head_end].decode()
    import re

    n = int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))
    m = int(re.search(r'NumberOfCells="(\d+)"', head).group(1))
    names = re.findall(r'DataArray type="Float64" Name="([^"]+)"', head)
    pos = head_end

    def block(dtype, count):
        nonlocal pos
        (nbytes,) = struct.unpack_from("<i", raw, pos)
        pos += 4
        arr = np.frombuffer(raw, dtype, count, pos)
        assert arr.nbytes == nbytes
        pos += nbytes
        return arr

    points = block("<f4", 3 * n).reshape(n, 3)
    conn = block("<i4", 8 * m).reshape(m, 8)
    offsets = block("<i4", m)
    types = block("u1", m)
    data = {name: block("<f8", n) for name in names}
    assert raw[pos:] == b"\n</AppendedData>\n</VTKFile>\n"
    return {"points": points, "connectivity": conn, "offsets": offsets, "types": types, "data": data}
