"""Regular wavefronts of the sparse layout (vfx_sparse_regular_groups): the
sparse step that forms the twist indices of regular wavefronts by arithmetic
leaves the same bytes as the step that reads every index, in FP64, FP32 and
FP16, over the fluid-node list and over typeOfGridNode, on an x-fastest
numbering with an obstacle (most wavefronts regular, those at the periodic
row ends and around the obstacle not) and on a shuffled numbering (none
regular)."""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu

DIMS = (256, 6, 5)


def _xfastest(dims, solid):
    nx, ny, nz = dims
    N = nx * ny * nz
    k = np.arange(N)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)

    def kof(xx, yy, zz):
        return (xx % nx) + nx * ((yy % ny) + ny * (zz % nz)) + 1

    nb = [np.zeros(N + 1, np.uint32) for _ in range(3)]
    nb[0][1:] = kof(x + 1, y, z)
    nb[1][1:] = kof(x, y + 1, z)
    nb[2][1:] = kof(x, y, z + 1)
    fluid = (np.nonzero(~solid)[0] + 1).astype(np.uint32)
    return nb, fluid, np.arange(N) + 1


def _obstacle(dims):
    nx, ny, nz = dims
    s = np.zeros((nz, ny, nx), bool)
    s[1:4, 2:4, 100:140] = True
    return s.ravel()


def _run(storage, layout, groups, steps=6):
    nb, fluid, sidx, types = layout
    N = nb[0].size - 1
    F0 = oracle.init_field(*cases.random_flow(DIMS, seed=4))
    Fs = np.zeros((N + 1, 27))
    Fs[sidx] = F0
    kw = dict(node_types=types) if types is not None else dict(fluid_nodes=fluid)
    lat = vfx.Lattice(storage=storage, omega=1.8, neighbors=tuple(nb), regular_groups=groups, **kw)
    lat.write_populations(Fs)
    for i in range(steps):
        lat.collide_stream("k17", i % 2 == 0)
    torch.cuda.synchronize()
    return lat


def _layouts():
    solid = _obstacle(DIMS)
    nb, fluid, sidx = _xfastest(DIMS, solid)
    types = np.full(nb[0].size, 19, np.uint32)
    types[0] = 16
    types[sidx[solid]] = 16
    nbs, nys, nzs, fluid_s, sidx_s = cases.sparse_layout(DIMS, solid, seed=2)
    return {"list": (nb, fluid, sidx, None), "types": (nb, fluid, sidx, types),
            "shuffled": ((nbs, nys, nzs), fluid_s, sidx_s, None)}


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
@pytest.mark.parametrize("name", ["list", "types", "shuffled"])
def test_regular_wavefronts_give_the_same_bytes(cuda, storage, name):
    print(f"[sparse regular groups {name} {storage}]", flush=True)
    layout = _layouts()[name]
    a = _run(storage, layout, False)
    b = _run(storage, layout, True)
    frac = b.regular_fraction
    print(f"  regular fraction {frac:.3f}", flush=True)
    if name == "shuffled":
        assert frac == 0.0
    else:
        assert 0.3 < frac < 0.95
    assert torch.equal(a.f.view(torch.uint8), b.f.view(torch.uint8))


def _run_shape(storage, layout, run, steps=6):
    nb, fluid, sidx, types = layout
    N = nb[0].size - 1
    F0 = oracle.init_field(*cases.random_flow(DIMS, seed=4))
    Fs = np.zeros((N + 1, 27))
    Fs[sidx] = F0
    lat = vfx.Lattice(storage=storage, omega=1.8, neighbors=tuple(nb), fluid_nodes=fluid)
    lat.xcd_run = run
    lat.write_populations(Fs)
    for i in range(steps):
        lat.collide_stream("k17", i % 2 == 0)
    torch.cuda.synchronize()
    return lat


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_sparse_xcd_runs_give_the_same_bytes(cuda, storage):
    """The XCD run of the sparse step only decides which workgroup takes which
    256 work items: every run leaves the built-in order's bytes."""
    print(f"[sparse xcd runs {storage}]", flush=True)
    layout = _layouts()["list"]
    ref = _run_shape(storage, layout, 0)
    for run in (1, 2, 16, 256):
        lat = _run_shape(storage, layout, run)
        assert torch.equal(lat.f.view(torch.uint8), ref.f.view(torch.uint8)), run


def test_sparse_autotune_keeps_parity_and_picks_a_run(cuda):
    print("[sparse autotune fp16]", flush=True)
    layout = _layouts()["list"]
    a, b = _run_shape("fp16", layout, 0, steps=0), _run_shape("fp16", layout, 0, steps=0)
    tile, run, ms = b.autotune(even=True, steps=2)
    print(f"  tile {tile} run {run} {ms:.4f} ms", flush=True)
    assert tile == 256 and run in (1, 4, 8, 16, 32, 64, 256) and ms > 0
    assert (b.tile_nodes, b.xcd_run) == (tile, run)
    # 7 trials of 2 warm-up + 2 timed steps: a continues from the same parity
    for i in range(28):
        a.collide_stream("k17", i % 2 == 0)
    for i in range(4):
        a.collide_stream("k17", i % 2 == 0)
        b.collide_stream("k17", i % 2 == 0)
    torch.cuda.synchronize()
    assert torch.equal(a.f.view(torch.uint8), b.f.view(torch.uint8))
