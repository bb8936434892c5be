"""Writes the precursor-file golden series with VirtualFluids' own image
writer (WbWriterVtkXmlImageBinary::writeData,
src/basics/writer/WbWriterVtkXmlImageBinary.cpp:258-360, compiled from the
reference sources into oracle/_ref/ref_vtk by `make -C oracle ref`), called as
PrecursorWriter::write calls it (whole extent = piece extent, precision 8).
Run once in a container that has the reference tree:

    python tests/golden/make_precursor_golden.py

Fixtures in tests/golden/precursor/ (reference writer output only):
  inflow_lev_0_ID_0_File_0.bin.vti, inflow_lev_0_ID_0_File_1.bin.vti
      the plane of tests/test_precursor_files.py (precursor_case): vx, vy, vz
      on 6 x 5 points (y, z), four records per file, records 0-3 and 4-7
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.test_precursor_files import precursor_case  # noqa: E402
from virtualfluids_amd import precursor_files as pf  # noqa: E402

OUT = os.path.join(HERE, "precursor")
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_vtk")


def write_input(path, names, data, extent, origin, spacing, precision):
    with open(path, "wb") as fh:
        fh.write(np.int32(len(names)).tobytes())
        for name in names:
            b = name.encode()
            fh.write(np.int32(len(b)).tobytes())
            fh.write(b)
        fh.write(np.int32(data.shape[1]).tobytes())
        fh.write(np.ascontiguousarray(data, np.float64).tobytes())
        fh.write(np.asarray(extent, np.int32).tobytes())
        fh.write(np.asarray(origin, np.float32).tobytes())
        fh.write(np.asarray(spacing, np.float32).tobytes())
        fh.write(np.int32(precision).tobytes())


def reference_files(dest, exe=EXE):
    os.makedirs(dest, exist_ok=True)
    case = precursor_case()
    out = []
    for part in case["parts"]:
        name = pf.file_name(os.path.join(dest, case["prefix"]), 0, 0, part["part"])
        inp = os.path.join(dest, "_input.bin")
        write_input(inp, case["names"], part["data"], part["extent"], case["origin"], case["spacing"],
                    case["precision"])
        subprocess.run([exe, "vti", inp, name], check=True, capture_output=True)
        os.remove(inp)
        out.append(os.path.basename(name))
    return out


if __name__ == "__main__":
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} is missing: build it with `make -C oracle ref` where the reference tree exists")
    print("\n".join(reference_files(OUT)))
