"""Build recipe for libvfx.so (gfx950).  ``python -m virtualfluids_amd.build``.

Plain hipcc invocations, one object per translation unit, so each unit can
carry its own floating-point contract flags:

* ``step_fp64.hip`` : ``-ffp-contract=off`` (FP64 results reproduce the
  reference CPU arithmetic bit for bit);
* ``step_fp32.hip`` : default contraction (FP32 arithmetic on FP32 / FP16 storage);
* ``api.hip``, host adapters: host-side dispatch, no kernels.

Every HIP unit is compiled with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvfx.so")
ARCH = os.environ.get("VFX_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include")]

UNITS = {
    "step_fp64.hip": ["-ffp-contract=off"],
    "step_fp32.hip": [],
    "api.hip": ["-ffp-contract=off"],
}
# the C++ host layer once per VirtualFluids `real` (float, and double as with
# VF_DOUBLE_ACCURACY); its inline namespaces keep the two apart
HOST_UNITS = {
    "host/vfx_host.cpp": [],
    "host/vfx_host.cpp#double": ["-DVF_DOUBLE_ACCURACY"],
}
GXX = shutil.which("g++") or "g++"


def _compile(src: str, flags: list[str]) -> str:
    out = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    return out


def _compile_host(src: str, flags: list[str]) -> str:
    out = os.path.join(OBJ, src.replace("/", "_").replace("#", "_") + ".o")
    path = os.path.join(CSRC, src.split("#")[0])
    cmd = [GXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off"] + flags + ["-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{' '.join(cmd)}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = [(_compile, src, fl) for src, fl in UNITS.items()] + [(_compile_host, src, fl) for src, fl in HOST_UNITS.items()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: j[0](j[1], j[2]), jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
