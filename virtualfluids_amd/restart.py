"""VirtualFluids restart files (RestartObject, src/gpu/core/Restart/RestartObject.cpp).

A restart object holds one distribution array per grid level, coarse to fine:
the `distributions.f[0]` array of 27 * numberOfNodes `real` values at the
current parity, NaN written as 0 (`RestartObject::serialize`, :33-48).

* binary, `<name>.bin` (`BinaryRestartObject::serialize_internal`, :158-186):
  int32 number of levels, then per level an int32 value count followed by the
  values as raw `real` (native byte order; float32, or float64 when the
  reference is built with VF_DOUBLE_ACCURACY);
* ASCII, `<name>.txt` (`ASCIIRestartObject::serialize_internal`, :70-97): one
  line per level, every value followed by one space, written by std::ostream
  with its default formatting (six significant digits, %g style).  Like the
  reference's, this form rounds the populations.

`save` / `load` move the arrays of `Lattice` objects through these files, so a
run can stop and resume (bit for bit with the binary form at the lattice's own
precision) or start from a VirtualFluids checkpoint.
"""
from __future__ import annotations

import os

import numpy as np

EXTENSION = {"binary": ".bin", "ascii": ".txt"}


def _real(real):
    real = np.dtype(real)
    if real not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError("VirtualFluids' real is float32 or float64")
    return real


def write(name, arrays, fmt="binary", real=np.float32) -> str:
    """Write one array per level to `name` + extension; returns the path."""
    real = _real(real)
    if fmt not in EXTENSION:
        raise ValueError("fmt is 'binary' or 'ascii'")
    path = str(name) + EXTENSION[fmt]
    levels = []
    for a in arrays:
        a = np.asarray(a, dtype=real).ravel().copy()
        a[np.isnan(a)] = 0  # RestartObject::serialize
        levels.append(a)
    if fmt == "binary":
        with open(path, "wb") as f:
            f.write(np.int32(len(levels)).tobytes())
            for a in levels:
                f.write(np.int32(a.size).tobytes())
                f.write(a.tobytes())
    else:
        with open(path, "w") as f:
            for a in levels:
                f.write("".join(f"{float(v):g} " for v in a))
                f.write("\n")
    return path


def read(name, fmt="binary", real=np.float32) -> list:
    """Read the per-level arrays of `name` + extension."""
    real = _real(real)
    if fmt not in EXTENSION:
        raise ValueError("fmt is 'binary' or 'ascii'")
    path = str(name) + EXTENSION[fmt]
    if not os.path.exists(path):
        raise FileNotFoundError(f"couldn't open check point file {path}")
    out = []
    if fmt == "binary":
        with open(path, "rb") as f:
            data = f.read()
        pos = 0
        (n,) = np.frombuffer(data, np.int32, 1, pos)
        pos += 4
        for _ in range(int(n)):
            (size,) = np.frombuffer(data, np.int32, 1, pos)
            pos += 4
            nbytes = int(size) * real.itemsize
            if pos + nbytes > len(data):
                raise ValueError(f"{path}: truncated restart file")
            out.append(np.frombuffer(data, real, int(size), pos).copy())
            pos += nbytes
    else:
        with open(path) as f:
            for line in f.read().split("\n")[:-1]:
                out.append(np.array([float(t) for t in line.split()], dtype=real))
    return out


def save(name, levels, fmt="binary", real=np.float32) -> str:
    """levels: [(lattice, even), ...] coarse to fine.  Writes each lattice's
    distribution array at its parity in VirtualFluids' layout."""
    import torch

    dt = torch.float64 if _real(real) == np.float64 else torch.float32
    arrays = [lat.export_distributions(even, dtype=dt).cpu().numpy() for lat, even in levels]
    return write(name, arrays, fmt, real)


def load(name, levels, fmt="binary", real=np.float32) -> None:
    """levels: [(lattice, even), ...] coarse to fine, matching the file."""
    import torch

    arrays = read(name, fmt, real)
    if len(arrays) != len(levels):
        raise ValueError(f"restart file has {len(arrays)} levels, {len(levels)} given")
    for a, (lat, even) in zip(arrays, levels):
        lat.import_distributions(torch.from_numpy(a), even)
