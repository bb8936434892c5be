"""Launch shapes of the structured K17 step (vfx_lattice.tile_nodes /
xcd_run) and its autotuner (vfx_k17_autotune): every shape leaves the same
bytes as the built-in one (the shape only decides which workgroup updates
which nodes), on a box whose x extent is no multiple of any tile width, with
solid nodes, in FP64, FP32 and FP16; the autotuner returns one of its shapes
and leaves the parity as it was."""
import numpy as np
import pytest
import torch

import virtualfluids_amd as vfx
from tests import cases

pytestmark = pytest.mark.gpu

DIMS = (100, 12, 10)


def _lattice(storage):
    rng = np.random.default_rng(5)
    solid = rng.random(int(np.prod(DIMS))) < 0.05
    lat = vfx.Lattice(DIMS, storage=storage, omega=1.8, solid=solid)
    drho, ux, uy, uz = cases.random_flow(DIMS, seed=3)
    lat.init_equilibrium(drho, ux, uy, uz)
    return lat


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_every_launch_shape_gives_the_same_bytes(cuda, storage):
    print(f"[launch shapes {storage}]", flush=True)
    ref = _lattice(storage)
    for i in range(6):
        ref.collide_stream("k17", i % 2 == 0)
    torch.cuda.synchronize()
    for tile in (64, 128, 256):
        for run in (1, 2, 16, 256):
            lat = _lattice(storage)
            lat.tile_nodes, lat.xcd_run = tile, run
            for i in range(6):
                lat.collide_stream("k17", i % 2 == 0)
            torch.cuda.synchronize()
            same = torch.equal(lat.f.view(torch.uint8), ref.f.view(torch.uint8))
            assert same, (tile, run)


def test_autotune_keeps_parity_and_picks_a_shape(cuda):
    print("[k17 autotune fp16]", flush=True)
    a, b = _lattice("fp16"), _lattice("fp16")
    tile, run, ms = b.autotune(even=True, steps=2)
    print(f"  tile {tile} run {run} {ms:.4f} ms", flush=True)
    assert tile in (64, 128, 256) and run in (1, 4, 8, 16, 32, 64, 256) and ms > 0
    assert (b.tile_nodes, b.xcd_run) == (tile, run)
    # 3 * 7 trials of 2 warm-up + 2 timed steps: an even number of steps, so
    # a continues from the same parity; b has the tuned shape
    for i in range(84):
        a.collide_stream("k17", i % 2 == 0)
    for i in range(4):
        a.collide_stream("k17", i % 2 == 0)
        b.collide_stream("k17", i % 2 == 0)
    torch.cuda.synchronize()
    assert torch.equal(a.f.view(torch.uint8), b.f.view(torch.uint8))
