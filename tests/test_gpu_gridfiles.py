"""A lid-driven cavity read from a VirtualFluids grid folder (the layout of
DrivenCavity, apps/gpu/DrivenCavity/DrivenCavity.cpp:98-146: velocity
bounce-back on the lid, no-slip bounce-back on the walls) runs on the sparse
layout exactly as the same cavity on the structured layout: FP64 populations
equal node for node after 30 steps, and equal to the oracle."""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases
from tests.test_gridfiles import cavity_grid
from virtualfluids_amd import gridfiles as gf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binary", [False, True])
def test_cavity_from_grid_folder(cuda, tmp_path, binary):
    print(f"[cavity from grid folder binary={binary}]", flush=True)
    n, steps, omega, lid = 6, 30, 1.7, (0.05, 0.02, 0.0)
    level, sides, sidx = cavity_grid(n, lid)
    gf.write_grid(tmp_path, [level], sides, binary=binary)
    lev, = gf.read_grid(tmp_path, binary=binary)
    m = n + 2
    solid, groups = cases.cavity(n, lid=lid)
    fields = cases.random_flow((m, m, m), seed=17, u_amp=0.02)
    F0 = oracle.init_field(*fields)
    F0[solid] = 0.0
    # structured reference run
    a = vfx.Lattice((m, m, m), storage="fp64", omega=omega, solid=solid)
    a.write_populations(F0)
    sa = vfx.Stepper(a, "k17", [vfx.BoundaryCondition(0, nd, q, u[0], u[1], u[2]) for nd, q, u in groups])
    sa.step(steps)
    # the grid folder's sparse lattice: the velocity table on the lid
    # (VelocityBounceBack), the no-slip table with zero wall velocity (NoSlipBounceBack)
    b = gf.lattice(lev, storage="fp64", omega=omega)
    Fs = np.zeros((lev.n_nodes, 27))
    Fs[sidx] = F0
    b.write_populations(Fs)
    k, q, vx, vy, vz = lev.boundaries["velocity"]
    kw, qw = lev.boundaries["noSlip"]
    sb = vfx.Stepper(b, "k17", [vfx.BoundaryCondition(0, k, q, vx, vy, vz),
                                vfx.BoundaryCondition(0, kw, qw, 0.0, 0.0, 0.0)])
    sb.step(steps)
    torch.cuda.synchronize()
    ga = sa.populations().cpu().numpy()
    gb = sb.populations().cpu().numpy()[sidx]
    fl = ~solid
    print(f"  identical {np.array_equal(ga[fl], gb[fl])}", flush=True)
    assert np.array_equal(ga[fl], gb[fl])
    ref = oracle.run(F0, (m, m, m), steps, omega, solid=solid,
                     bcs=[oracle.BCGroup(nd, q, *u, kind=0) for nd, q, u in groups])
    assert np.array_equal(gb[fl], ref[fl])
