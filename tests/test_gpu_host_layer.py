"""The C++ host layer (KernelFactoryImp, BoundaryConditionFactory,
UpdateGrid27 with QforBoundaryConditions / QforDirectionalBoundaryCondition /
QforPrecursorBoundaryConditions holding VirtualFluids `real` arrays) against
the oracle and against the same case stepped through the Python Stepper.

tests/cpp/host_update_grid.cpp is built twice: `real` = double
(VF_DOUBLE_ACCURACY) drives FP64 storage, `real` = float drives FP32 and FP16
storage.  Its boundary tables are `real` device arrays that go to the C ABI as
they are (q27[0], Vx/Vy/Vz, RhoBC, the precursor weights and planes).

* channel with 3 spheres (interpolated no-slip with ray-sphere q, velocity
  inlet, directional pressure outlet): FP64 equal to the oracle and to the
  Python stepper bit for bit; FP32 and FP16 equal to the Python stepper on
  the same float tables bit for bit, and close to FP64 (velocity L2);
* precursor inflows (both kinds) with a plane reader: getNextData is asked
  for the planes at 3T, 4T, ... and the populations equal the oracle stepped
  with the closed-form cycle and the Python PrecursorPlanes run, bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases
from tests.test_geometry import oracle_groups, small_channel
from virtualfluids_amd import geometry

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "host_update_grid.cpp")
LIBDIR = os.path.dirname(vfx._lib.LIB_PATH)
STORAGE = {"fp64": 0, "fp32": 1, "fp16": 2}


def harness(real):
    exe = os.path.join(ROOT, "tests", "cpp", f"host_update_grid_{real}")
    # the harness links the host classes' layout: rebuild after any change to
    # its source, the host header or the library
    deps = [SRC, os.path.join(ROOT, "virtualfluids_amd", "csrc", "host", "vfx_host.hpp"),
            os.path.join(LIBDIR, "libvfx.so")]
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(d) for d in deps):
        flags = ["-DVF_DOUBLE_ACCURACY"] if real == "f64" else []
        tmp = f"{exe}.{os.getpid()}"
        subprocess.run(["g++", "-O1", "-std=c++17"] + flags + [SRC, "-o", tmp, "-L" + LIBDIR, "-lvfx",
                        "-Wl,-rpath," + LIBDIR], check=True)
        os.replace(tmp, exe)
    return exe


class Case:
    """Writes a case directory in host_update_grid's format."""

    def __init__(self, path, dims, storage, omega, steps):
        self.path = str(path)
        os.makedirs(self.path, exist_ok=True)
        self.lines = [f"dims {dims[0]} {dims[1]} {dims[2]}", f"storage {STORAGE[storage]}", f"omega {omega!r}",
                      f"steps {steps}"]
        if storage == "fp16":
            self.lines.append("fp16_scale 1024")
        self.n = 0

    def arr(self, a, dtype):
        if a is None:
            return "-"
        self.n += 1
        name = f"a{self.n}.{'f64' if dtype == np.float64 else 'i32'}"
        np.ascontiguousarray(a, dtype=dtype).tofile(os.path.join(self.path, name))
        return name

    def solid(self, solid):
        mask = vfx.pack_mask(solid)
        self.lines.append(f"solid {self.arr(mask.view(np.int32), np.int32)} {mask.size}")

    def populations(self, F):
        self.lines.append(f"populations {self.arr(F, np.float64)}")

    def bc(self, group, kind, nodes, q27, vx=None, vy=None, vz=None):
        n = len(nodes)
        self.lines.append(f"bc {group} {kind} {n} {self.arr(nodes, np.int32)} {self.arr(q27, np.float64)} "
                          f"{self.arr(vx, np.float64)} {self.arr(vy, np.float64)} {self.arr(vz, np.float64)}")

    def pressure(self, kind, nodes, kn, rho, direction):
        self.lines.append(f"bc pressure {kind} {len(nodes)} {self.arr(nodes, np.int32)} {self.arr(kn, np.int32)} "
                          f"{self.arr(rho, np.float64)} {direction}")

    def precursor(self, kind, nodes, q27, n_pre, comps, T, vratio, u0, nb, w, planes):
        pn = " ".join(self.arr(nb[c], np.int32) for c in range(4))
        ws = " ".join(self.arr(w[c], np.float64) for c in range(4))
        self.lines.append(f"precursor {kind} {len(nodes)} {self.arr(nodes, np.int32)} {self.arr(q27, np.float64)} "
                          f"{n_pre} {comps} {T} {vratio!r} 0 {u0[0]!r} {u0[1]!r} {u0[2]!r} {pn} {ws} "
                          f"{self.arr(np.stack(planes), np.float64)} {len(planes)}")

    def run(self, real, N):
        with open(os.path.join(self.path, "case.txt"), "w") as f:
            f.write("\n".join(self.lines) + "\n")
        print(f"  host_update_grid_{real} {self.path}", flush=True)
        r = subprocess.run([harness(real), self.path], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out = np.fromfile(os.path.join(self.path, "out.f64"), dtype=np.float64).reshape(N, 27)
        lt = open(os.path.join(self.path, "load_times.txt")).read().split()
        return out, [float(t) for t in lt]


def channel_case(tmp_path, ch, storage, steps, F0):
    c = Case(tmp_path / f"channel_{storage}", ch.dims, storage, ch.omega, steps)
    c.solid(ch.solid)
    c.populations(F0)
    for b in ch.bcs:
        if b.kind == 4:
            c.pressure(4, b.nodes, b.neighbors, b.rho, b.normal)
        elif b.kind == 2:
            c.bc("velocity", 2, b.nodes, b.q27, *b.velocity)
        else:
            c.bc("noslip", b.kind, b.nodes, b.q27)
    return c


def python_channel(ch, storage, steps, F0):
    lat = vfx.Lattice(ch.dims, storage=storage, omega=ch.omega, solid=ch.solid)
    lat.write_populations(F0)
    st = vfx.Stepper(lat, "k17", geometry.boundary_conditions(ch.bcs))
    st.step(steps)
    torch.cuda.synchronize()
    return st.populations().cpu().numpy(), lat, st


def test_host_layer_channel_fp64_matches_oracle(cuda, tmp_path):
    print("[host layer channel fp64]", flush=True)
    ch = small_channel()
    steps = 40
    N = int(np.prod(ch.dims))
    drho, ux, uy, uz = cases.random_flow(ch.dims, seed=21, u_amp=0.01)
    F0 = oracle.init_field(drho, ux + ch.u_in, uy, uz)
    F0[ch.solid] = 0.0
    got, _ = channel_case(tmp_path, ch, "fp64", steps, F0).run("f64", N)
    ref = oracle.run(F0, ch.dims, steps, ch.omega, solid=ch.solid, bcs=oracle_groups(ch))
    py, _, _ = python_channel(ch, "fp64", steps, F0)
    fl = ~ch.solid
    print(f"  oracle exact {np.array_equal(got[fl], ref[fl])}, python exact {np.array_equal(got[fl], py[fl])}",
          flush=True)
    assert np.array_equal(got[fl], ref[fl])
    assert np.array_equal(got[fl], py[fl])


@pytest.mark.parametrize("storage,bar", [("fp32", 1e-4), ("fp16", 1e-2)])
def test_host_layer_channel_float_real(cuda, tmp_path, storage, bar):
    """`real` = float tables on FP32 / FP16 storage: the same populations as
    the Python stepper (same kernels, same float tables), close to FP64."""
    print(f"[host layer channel {storage}]", flush=True)
    ch = small_channel()
    steps = 40
    N = int(np.prod(ch.dims))
    drho, ux, uy, uz = cases.random_flow(ch.dims, seed=22, u_amp=0.01)
    F0 = oracle.init_field(drho, ux + ch.u_in, uy, uz)
    F0[ch.solid] = 0.0
    got, _ = channel_case(tmp_path, ch, storage, steps, F0).run("f32", N)
    py, _, _ = python_channel(ch, storage, steps, F0)
    ref, _, _ = python_channel(ch, "fp64", steps, F0)
    fl = ~ch.solid
    assert np.array_equal(got[fl], py[fl])
    mg = np.array([oracle.macroscopic(f) for f in got[fl]])
    mr = np.array([oracle.macroscopic(f) for f in ref[fl]])
    l2 = np.sqrt(((mg[:, 1:] - mr[:, 1:]) ** 2).sum() / (mr[:, 1:] ** 2).sum())
    print(f"  velocity L2 vs FP64 {l2:.2e}", flush=True)
    assert l2 <= bar


def test_host_layer_rejects_real_storage_mismatch(cuda, tmp_path):
    """The double-real host layer refuses FP32 storage at the first boundary condition."""
    ch = small_channel()
    N = int(np.prod(ch.dims))
    F0 = np.zeros((N, 27))
    c = channel_case(tmp_path, ch, "fp32", 1, F0)
    with open(os.path.join(c.path, "case.txt"), "w") as f:
        f.write("\n".join(c.lines) + "\n")
    r = subprocess.run([harness("f64"), c.path], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "VF_DOUBLE_ACCURACY needs VFX_FP64" in r.stderr


# ------------------------------------------------------------------ precursor
DIMS = (20, 9, 8)
N_PRE = 40
U = 0.03
PDIRS = [1, 7, 9, 11, 13, 19, 21, 23, 25]


@pytest.mark.parametrize("kind", [12, 13])
def test_host_layer_precursor_cycle(cuda, tmp_path, kind):
    """UpdateGrid27::runPrecursorBCKernelPost: planes read at 3T, 4T, ... and the
    populations equal to the oracle stepped with the closed-form cycle and to
    the Python PrecursorPlanes run, bit for bit."""
    print(f"[host layer precursor kind {kind}]", flush=True)
    T, steps = 3, 10
    solid, groups, onodes, onbrs = cases.inflow_channel(*DIMS)
    inlet, iq = groups["inlet"]
    n = inlet.size
    rng = np.random.default_rng(40 + kind)
    nb = rng.integers(0, N_PRE, (4, n))
    w = rng.uniform(0.1, 2.0, (4, n))
    w[0, :5] = 2e6
    comps = 3 if kind == 12 else 9
    if kind == 12:
        base = np.zeros((3, N_PRE))
        amp = 2e-3
    else:
        base = np.repeat(oracle.init_field(np.zeros(1), np.full(1, U), np.zeros(1), np.zeros(1))[0][PDIRS][:, None],
                         N_PRE, axis=1)
        amp = 1e-4
    planes = [base + amp * np.sin(0.7 * m * T + np.arange(comps * N_PRE).reshape(comps, N_PRE)) for m in range(7)]

    def plane(t):
        return planes[int(round(t / T))]

    u0 = (1.25 * U, 0.0, 0.0)
    drho, vx, vy, vz = cases.random_flow(DIMS, seed=9, u_amp=0.01)
    F0 = oracle.init_field(drho, vx + U, vy, vz)
    F0[solid] = 0.0
    N = int(np.prod(DIMS))
    # host layer, double real
    c = Case(tmp_path / f"precursor{kind}", DIMS, "fp64", 1.9, steps)
    c.solid(solid)
    c.populations(F0)
    c.pressure(6, onodes, onbrs, np.zeros(onodes.size), 1)
    c.bc("noslip", 1, *groups["wall"])
    c.precursor(kind, inlet, iq if kind == 12 else None, N_PRE, comps, T, 1.25, u0 if kind == 12 else (0, 0, 0),
                nb, w, planes)
    got, load_times = c.run("f64", N)
    print(f"  load times {load_times}", flush=True)
    assert load_times == [float(m * T) for m in range(3, 3 + len(load_times))] and len(load_times) == 3
    # oracle with the closed-form cycle: last = plane(mT), current = plane((m+1)T), ratio (t - mT)/T
    ux = np.full(n, u0[0]) if kind == 12 else None
    pre = dict(neighbors=nb, weights=w, series=planes, period=T, t0=1, velocity_ratio=1.25)
    bcs = [oracle.BCGroup(onodes, kind=6, nbr=onbrs, normal=1), oracle.BCGroup(*groups["wall"], kind=1),
           oracle.BCGroup(inlet, iq if kind == 12 else None, ux, kind=kind, precursor=pre)]
    F = oracle.run(F0, DIMS, steps, 1.9, solid=solid, bcs=bcs)
    fl = ~solid
    assert np.array_equal(got[fl], F[fl])
    # the Python stepper with PrecursorPlanes' reader
    lat = vfx.Lattice(DIMS, storage="fp64", omega=1.9, solid=solid)
    lat.write_populations(F0)
    pc = vfx.PrecursorPlanes(nb, w, N_PRE, comps, reader=plane, time_steps_between_reads=T, velocity_ratio=1.25)
    bcs = [vfx.BoundaryCondition(6, onodes, neighbors=onbrs, normal=1), vfx.BoundaryCondition(1, *groups["wall"]),
           vfx.BoundaryCondition(kind, inlet, iq if kind == 12 else None, ux, precursor=pc)]
    lat.apply_bc(bcs[0], even=True)
    st = vfx.Stepper(lat, "k17", bcs)
    st.step(steps)
    py = st.populations().cpu().numpy()
    assert np.array_equal(got[fl], py[fl])
