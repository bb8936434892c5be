"""The C ABI library and the C++ host layer, without a GPU: every entry point
include/vfx.h declares is exported by libvfx.so, argument validation fails
cleanly, and the C++ KernelFactory / BoundaryConditionFactory / UpdateGrid27
behave as VirtualFluids' (string keys, errors)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import virtualfluids_amd as vfx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vfx.h")
TEST_SRC = os.path.join(ROOT, "tests", "cpp", "host_adapters_test.cpp")
TEST_BIN = os.path.join(ROOT, "tests", "cpp", "host_adapters_test")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(vfx_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    lib = vfx.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_arguments_fail_without_device():
    L = vfx.lib()
    s = vfx._lib.VfxLattice()
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"NULL" in L.vfx_last_error()
    s.f = 1  # non-null but no nodes
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    s.n_nodes = 8
    s.omega = 1.5
    s.direct = 1
    s.dims[:] = [2, 2, 3]
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL  # dims do not match
    s.dims[:] = [2, 2, 2]
    assert L.vfx_collide_stream(C.byref(s), 7, 1, 0, None) == vfx._lib.VFX_EINVAL  # unknown collision
    s.omega = 2.0
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    s.omega = 1.5
    s.fp16_reference_velocity[:] = [0.05, 0.0, 0.0]  # FP16 storage only
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"fp16_reference_velocity" in L.vfx_last_error()
    s.fp16_reference_velocity[:] = [0.0, 0.0, 0.0]
    s.type_of_grid_node = 1  # typeOfGridNode belongs to the sparse layout
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"type_of_grid_node" in L.vfx_last_error()
    s.type_of_grid_node = None
    s.turbulence_model = 4  # unknown model
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"turbulence_model" in L.vfx_last_error()
    s.turbulence_model = vfx._lib.VFX_TURB_SMAGORINSKY  # no viscosity array
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"turbulent_viscosity" in L.vfx_last_error()
    s.turbulence_model = vfx._lib.VFX_TURB_AMD
    s.turbulent_viscosity = 1
    assert L.vfx_collide_stream(C.byref(s), 0, 1, 0, None) == vfx._lib.VFX_EINVAL  # AMD writes the velocities
    assert b"AMD" in L.vfx_last_error()
    assert L.vfx_turbulent_viscosity_amd(C.byref(s), None) == vfx._lib.VFX_EINVAL
    s.vx = s.vy = s.vz = 1
    s.direct = 0
    s.neighbor_x = s.neighbor_y = s.neighbor_z = 1
    assert L.vfx_turbulent_viscosity_amd(C.byref(s), None) == vfx._lib.VFX_EINVAL  # sparse needs the node types
    assert b"neighbor_inverse" in L.vfx_last_error()
    s.direct = 1
    s.neighbor_x = s.neighbor_y = s.neighbor_z = None
    s.vx = s.vy = s.vz = None
    s.turbulence_model = 0
    s.turbulent_viscosity = None
    assert L.vfx_export_distributions(C.byref(s), 1, 1, 9, None) == vfx._lib.VFX_EINVAL  # float or double only
    assert L.vfx_export_distributions(C.byref(s), 1, None, 1, None) == vfx._lib.VFX_EINVAL
    assert L.vfx_version() >= 1


def test_refinement_arguments_fail_without_device():
    L = vfx.lib()
    V = vfx._lib
    a = V.VfxLattice()
    b = V.VfxLattice()
    cells = V.VfxICells()
    assert L.vfx_scale_cf(C.byref(a), C.byref(b), C.byref(cells), 1, 1, None) == V.VFX_EINVAL  # f NULL
    for s in (a, b):
        s.f = 1
        s.n_nodes = 8
        s.omega = 1.9
        s.direct = 1
        s.dims[:] = [2, 2, 2]
    b.storage = 1
    assert L.vfx_scale_fc(C.byref(a), C.byref(b), C.byref(cells), 1, 1, None) == V.VFX_EINVAL
    assert b"storage differ" in L.vfx_last_error()
    b.storage = 0
    assert L.vfx_scale_cf(C.byref(a), C.byref(b), None, 1, 1, None) == V.VFX_EINVAL
    cells.n_cells = 3
    assert L.vfx_scale_cf(C.byref(a), C.byref(b), C.byref(cells), 1, 1, None) == V.VFX_EINVAL  # NULL index arrays


def test_batch_arguments_fail_without_device():
    L = vfx.lib()
    V = vfx._lib
    doms = (V.VfxDomain * 2)()
    for d in doms:
        d.dims[:] = [8, 8, 8]
        d.omega = 1.8
    doms[1].f_offset = 27 * 512
    h = C.c_void_p()
    assert L.vfx_batch_create(None, 27 * 1024, 2, 1024.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL  # NULL pool
    assert L.vfx_batch_create(1, 27 * 1024, 2, 1024.0, doms, 0, None, 0, C.byref(h)) == V.VFX_EINVAL  # no domains
    assert L.vfx_batch_create(1, 27 * 1024, 2, 0.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL  # FP16 scale
    assert L.vfx_batch_create(1, 27 * 1024, 2, 1024.0, doms, 2, None, 5, C.byref(h)) == V.VFX_EINVAL  # variant
    doms[1].dims[:] = [8, 2, 8]
    assert L.vfx_batch_create(1, 27 * 1024, 2, 1024.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL
    assert b"domain 1" in L.vfx_last_error()
    doms[1].dims[:] = [8, 8, 8]
    doms[0].omega = 2.0
    assert L.vfx_batch_create(1, 27 * 1024, 2, 1024.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL
    doms[0].omega = 1.8
    assert L.vfx_batch_create(1, 27 * 1024 - 1, 2, 1024.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL
    assert b"beyond the pool" in L.vfx_last_error()
    doms[1].f_offset = 27 * 511
    assert L.vfx_batch_create(1, 27 * 1024, 2, 1024.0, doms, 2, None, 0, C.byref(h)) == V.VFX_EINVAL
    assert b"overlap" in L.vfx_last_error()
    assert h.value is None
    assert L.vfx_batch_step(None, 1, None) == V.VFX_EINVAL
    assert L.vfx_graph_launch(None, None) == V.VFX_EINVAL
    assert L.vfx_capture_begin(None) == V.VFX_EINVAL  # the NULL stream cannot be captured
    assert L.vfx_stream_destroy(None) == V.VFX_EINVAL


def build_host_test():
    # built under a name of its own, then renamed into place: tests running in
    # parallel processes build and run it at the same time
    tmp = f"{TEST_BIN}.{os.getpid()}"
    cmd = ["g++", "-O1", "-std=c++17", "-DVF_DOUBLE_ACCURACY", TEST_SRC, "-o", tmp, "-L" + os.path.dirname(vfx._lib.LIB_PATH), "-lvfx",
           "-Wl,-rpath," + os.path.dirname(vfx._lib.LIB_PATH)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, TEST_BIN)
    return TEST_BIN


def test_host_layer_factories_cpu():
    exe = build_host_test()
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_host_layer_steps_like_c_abi(cuda):
    deps = [TEST_SRC, vfx._lib.LIB_PATH, os.path.join(ROOT, "virtualfluids_amd", "csrc", "host", "vfx_host.hpp")]
    fresh = os.path.exists(TEST_BIN) and os.path.getmtime(TEST_BIN) >= max(os.path.getmtime(d) for d in deps)
    exe = TEST_BIN if fresh else build_host_test()
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_launch_shape_arguments_fail_without_device():
    """tile_nodes and xcd_run of the structured and the sparse step are
    checked before any launch, by the step and by the autotuner."""
    L = vfx.lib()
    s = vfx._lib.VfxLattice()
    s.f = 1
    s.n_nodes = 8
    s.omega = 1.5
    s.direct = 1
    s.dims[:] = [2, 2, 2]
    s.tile_nodes = 100
    assert L.vfx_k17_step(C.byref(s), 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"tile_nodes" in L.vfx_last_error()
    s.tile_nodes = 128
    for run in (5000, 8192, 12):
        s.xcd_run = run
        assert L.vfx_k17_step(C.byref(s), 1, 0, None) == vfx._lib.VFX_EINVAL
        assert b"xcd_run" in L.vfx_last_error()
    t, r, ms = C.c_uint32(), C.c_uint32(), C.c_float()
    assert L.vfx_k17_autotune(C.byref(s), 1, 0, 4, None, C.byref(t), C.byref(r), C.byref(ms)) == vfx._lib.VFX_EINVAL
    assert b"xcd_run" in L.vfx_last_error()
    # the sparse step: 256-item workgroups, any power-of-two XCD run
    nb = (C.c_uint32 * 8)()
    s.direct = 0
    s.neighbor_x = s.neighbor_y = s.neighbor_z = C.cast(nb, C.c_void_p).value
    s.xcd_run = 16
    for tile in (64, 128):
        s.tile_nodes = tile
        assert L.vfx_k17_step(C.byref(s), 1, 0, None) == vfx._lib.VFX_EINVAL
        assert b"tile_nodes of the sparse step" in L.vfx_last_error()
        assert L.vfx_k17_autotune(C.byref(s), 1, 0, 4, None, C.byref(t), C.byref(r), C.byref(ms)) == vfx._lib.VFX_EINVAL
    s.tile_nodes = 256
    s.xcd_run = 12
    assert L.vfx_k17_step(C.byref(s), 1, 0, None) == vfx._lib.VFX_EINVAL
    assert b"xcd_run" in L.vfx_last_error()
    assert L.vfx_k17_autotune(None, 1, 0, 4, None, C.byref(t), C.byref(r), C.byref(ms)) == vfx._lib.VFX_EINVAL
