"""Writes the VTK output golden files with VirtualFluids' own writer
(WbWriterVtkXmlBinary, src/basics/writer/WbWriterVtkXmlBinary.cpp, compiled
from the reference sources into oracle/_ref/ref_vtk by `make -C oracle ref`).
Run once in a container that has the reference tree:

    python tests/golden/make_vtk_golden.py

Fixtures in tests/golden/vtk/ (reference writer output only):
  run_bin_lev_0_ID_0_Part_1_t_400.vtk.bin.vtu
                 writeOctsWithNodeData (:1306-1425) of the 9 x 6 x 5 box with a
                 solid outer layer of tests/test_output.py (_box, seed 3): the
                 points, voxels and point data press, rho, vx1, vx2, vx3, geo,
                 nut that virtualfluids_amd.output.unstructured_grid_parts makes
                 of it
  run_bin_ID_0_t_400.vtk.pvtu   writeParallelFile (:141-177) of that part
  run_bin_ID_0.pvd              writeCollectionForTimeSeries (:88-99) with steps
                                400 and 1000000
  mean_bin_mean_lev_0_ID_0_Part_1_t_400.vtk.bin.vtu, mean_bin_mean_ID_0_t_400.vtk.pvtu
                 the same writer with the mean-value names of
                 FileWriter::getMeanNodeDataNames (FileWriter.cpp:197-210) and
                 the first six arrays
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.test_output import vtk_golden_case  # noqa: E402
from virtualfluids_amd import output as out  # noqa: E402

OUT = os.path.join(HERE, "vtk")
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_vtk")


def write_input(path, points, cells, names, data):
    with open(path, "wb") as fh:
        fh.write(np.int32(points.shape[0]).tobytes())
        fh.write(np.ascontiguousarray(points, np.float32).tobytes())
        fh.write(np.int32(cells.shape[0]).tobytes())
        fh.write(np.ascontiguousarray(cells, np.uint32).tobytes())
        fh.write(np.int32(len(names)).tobytes())
        for name, d in zip(names, data):
            b = name.encode()
            fh.write(np.int32(len(b)).tobytes())
            fh.write(b)
            fh.write(np.ascontiguousarray(d, np.float64).tobytes())


def reference_files(dest, exe=EXE):
    """Run the reference writer on the golden case; returns the file names."""
    os.makedirs(dest, exist_ok=True)
    case = vtk_golden_case()
    files = []
    for prefix, mean in (("run", False), ("mean", True)):
        names = case["mean_names"] if mean else case["names"]
        pieces = []
        for i, (points, cells, data) in enumerate(case["parts"], start=1):
            inp = os.path.join(dest, f"_input_{prefix}_{i}.bin")
            write_input(inp, points, cells, names, data[:len(names)])
            name = out.part_file_name(os.path.join(dest, prefix), 0, 0, i, case["t"], mean=mean)
            r = subprocess.run([exe, "octs", inp, name], capture_output=True, text=True, check=True)
            os.remove(inp)
            pieces.append(os.path.basename(r.stdout.strip()))
        coll = out.collection_file_name(os.path.join(dest, prefix), 0, case["t"], mean=mean)
        subprocess.run([exe, "pvtu", coll, ",".join(names)] + pieces, check=True, capture_output=True)
        files += pieces + [os.path.basename(coll) + ".pvtu"]
        if not mean:
            pvd = out.pvd_file_name(os.path.join(dest, prefix), 0)
            subprocess.run([exe, "pvd", pvd, f"{case['t']}={','.join(pieces)}", f"1000000={','.join(pieces)}"],
                           check=True, capture_output=True)
            files.append(os.path.basename(pvd) + ".pvd")
    return sorted(files)


if __name__ == "__main__":
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} is missing: build it with `make -C oracle ref` where the reference tree exists")
    print("\n".join(reference_files(OUT)))
