"""Grid refinement on the GPU (vfx_scale_cf / vfx_scale_fc, RefinedStepper)
against the oracle.

FP64 bar: 1e-12 relative to the largest population (the kernels follow the
oracle's formulation; only the order of the five second-moment sums differs).
Reduced precision: velocity L2 of a shear wave through the nest vs FP64.
"""
import numpy as np
import pytest
import torch

import oracle
import virtualfluids_amd as vfx
from tests import cases
from tests.test_oracle_refinement import refined_oracle_run, shear_wave_init, shear_wave_nest

pytestmark = pytest.mark.gpu

REL = 1e-12


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def _random_state(dims, seed):
    return oracle.init_field(*cases.random_flow(dims, seed=seed)) + \
        np.random.default_rng(seed).normal(0, 1e-3, (int(np.prod(dims)), 27)) * cases.W


NEST = ((16, 14, 14), (3, 3, 4), (11, 10, 11))


@pytest.mark.parametrize("even_c,even_f", [(True, True), (False, True), (True, False), (False, False)])
def test_scale_single_call_direct(cuda, even_c, even_f):
    print(f"[scale direct even_c={even_c} even_f={even_f}]", flush=True)
    nest = vfx.NestedBox(*NEST)
    oc = 1.93
    of = vfx.fine_omega(oc)
    Fc = _random_state(nest.coarse_dims, 1)
    Ff = _random_state(nest.fine_dims, 2)
    C = vfx.Lattice(nest.coarse_dims, storage="fp64", omega=oc)
    F = vfx.Lattice(nest.fine_dims, storage="fp64", omega=of)
    C.write_populations(Fc, even=even_c)
    F.write_populations(Ff, even=even_f)
    fc = vfx.ICells(nest.fc_coarse, nest.fc_fine, "cuda")
    cf = vfx.ICells(nest.cf_coarse, nest.cf_fine, "cuda")
    vfx.scale_fc(C, F, fc, even_c, even_f)
    vfx.scale_cf(C, F, cf, even_c, even_f)
    torch.cuda.synchronize()
    got_c = C.read_populations(even_c).cpu().numpy()
    got_f = F.read_populations(even_f).cpu().numpy()
    ref_c = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, oc, of)
    ref_f = oracle.scale_cf(ref_c, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, oc, of)
    ec, ef = rel(got_c, ref_c), rel(got_f, ref_f)
    print(f"coarse {ec:.3e} fine {ef:.3e}")
    assert ec <= REL and ef <= REL


def test_scale_offsets_sparse_layout(cuda):
    """Sparse (neighbour-array) addressing on both levels, permuted node order,
    and per-cell offsets."""
    print("[scale sparse + offsets]", flush=True)
    nest = vfx.NestedBox(*NEST)
    oc = 1.9
    of = vfx.fine_omega(oc)
    rng = np.random.default_rng(3)
    Fc = _random_state(nest.coarse_dims, 4)
    Ff = _random_state(nest.fine_dims, 5)
    layouts = []
    for dims, F, seed, omega in ((nest.coarse_dims, Fc, 6, oc), (nest.fine_dims, Ff, 7, of)):
        nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, seed=seed)
        Fs = np.zeros((sidx.size + 1, 27))
        Fs[sidx] = F
        lat = vfx.Lattice(storage="fp64", omega=omega, neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
        lat.write_populations(Fs, even=True)
        layouts.append((lat, sidx))
    (C, sc), (F, sf) = layouts
    offs_fc = [rng.choice([-0.5, 0.0, 0.5], nest.fc_coarse.size) for _ in range(3)]
    offs_cf = [rng.choice([-0.5, 0.0, 0.5], nest.cf_coarse.size) for _ in range(3)]
    fc = vfx.ICells(sc[nest.fc_coarse], sf[nest.fc_fine], "cuda", offsets=offs_fc)
    cf = vfx.ICells(sc[nest.cf_coarse], sf[nest.cf_fine], "cuda", offsets=offs_cf)
    vfx.scale_fc(C, F, fc, True, True)
    vfx.scale_cf(C, F, cf, True, True)
    torch.cuda.synchronize()
    got_c = C.read_populations(True).cpu().numpy()[sc]
    got_f = F.read_populations(True).cpu().numpy()[sf]
    ref_c = oracle.scale_fc(Fc, Ff, nest.coarse_dims, nest.fine_dims, nest.fc_coarse, nest.fc_fine, oc, of, offs_fc)
    ref_f = oracle.scale_cf(ref_c, Ff, nest.coarse_dims, nest.fine_dims, nest.cf_coarse, nest.cf_fine, oc, of,
                            offs_cf)
    ec, ef = rel(got_c, ref_c), rel(got_f, ref_f)
    print(f"coarse {ec:.3e} fine {ef:.3e}")
    assert ec <= REL and ef <= REL


def test_refined_steps_fp64_match_oracle(cuda):
    """RefinedStepper (two fine steps, one coarse step, FC, CF per coarse step)
    vs the oracle's two-level run, 20 coarse steps of a shear wave plus noise:
    every fine node and every active coarse node."""
    steps = 20
    print(f"[refined fp64 x {steps}]", flush=True)
    nest = shear_wave_nest()
    oc = 1.9
    Fc, Ff, _ = shear_wave_init(nest, 0.02)
    rng = np.random.default_rng(8)
    Fc = Fc + rng.normal(0, 1e-4, Fc.shape) * cases.W
    Ff = Ff + rng.normal(0, 1e-4, Ff.shape) * cases.W
    st = vfx.RefinedStepper(nest, oc, storage="fp64")
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    st.step(steps)
    torch.cuda.synchronize()
    got_c = st.coarse.read_populations(st.even_c).cpu().numpy()
    got_f = st.fine.read_populations(st.even_f).cpu().numpy()
    ref_c, ref_f = refined_oracle_run(nest, Fc, Ff, steps, oc)
    act = nest.coarse_state == 0
    ec, ef = rel(got_c[act], ref_c[act]), rel(got_f, ref_f)
    print(f"coarse {ec:.3e} fine {ef:.3e}")
    assert ec <= REL and ef <= REL


@pytest.mark.parametrize("storage,bar", [("fp32", 1e-4), ("fp16", 1e-2)])
def test_refined_reduced_precision_shear_wave(cuda, storage, bar):
    """Shear wave (u = 0.02) through the nest, 200 coarse steps (400 fine):
    velocity L2 vs FP64 on both levels.  FP16 bar 1e-2: for this x/z-uniform
    wave the FP16 rounding errors are coherent and add up step by step; storing
    a single-level box in FP16 (CPU emulation, FP64 arithmetic otherwise)
    already gives 4.5e-3 after 200 and 8.3e-3 after 400 steps, so the bar
    checks that the interface adds nothing beyond storage rounding."""
    steps = 200
    print(f"[refined {storage} x {steps}]", flush=True)
    nest = shear_wave_nest()
    Fc, Ff, _ = shear_wave_init(nest, 0.02)
    runs = {}
    for s in ("fp64", storage):
        st = vfx.RefinedStepper(nest, 1.9, storage=s)
        st.coarse.write_populations(Fc, even=True)
        st.fine.write_populations(Ff, even=True)
        st.step(steps)
        torch.cuda.synchronize()
        runs[s] = (st.coarse.macroscopic(st.even_c), st.fine.macroscopic(st.even_f))
    act = torch.from_numpy(nest.coarse_state == 0).cuda()
    for lvl in (0, 1):
        ref = torch.stack(runs["fp64"][lvl][1:4])
        got = torch.stack(runs[storage][lvl][1:4])
        if lvl == 0:
            ref, got = ref[:, act], got[:, act]
        l2 = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
        print(f"level {lvl}: velocity L2 {l2:.3e}")
        assert l2 <= bar


def test_refined_graph_replay_matches_eager(cuda):
    """One HIP graph per 2 coarse steps (4 fine steps, 2 coarse steps, the
    interpolations), replayed 3 times with eager work in between, equals the
    eager launches bit for bit on both levels; a graph of an odd number of
    steps refuses to replay from the other parity."""
    print("[refined graph replay]", flush=True)
    nest = shear_wave_nest()
    Fc, Ff, _ = shear_wave_init(nest, 0.02)
    runs = []
    for use_graph in (False, True):
        st = vfx.RefinedStepper(nest, 1.9, storage="fp32")
        st.coarse.write_populations(Fc, even=True)
        st.fine.write_populations(Ff, even=True)
        g = st.capture(2) if use_graph else None
        for _ in range(3):
            if use_graph:
                st.replay(g, 2)
            else:
                st.step(2)
            torch.cuda.synchronize()
            _ = float(st.fine.f.abs().sum())  # eager work between replays
        runs.append((st.coarse.f.clone(), st.fine.f.clone(), st.even_c, st.even_f))
    assert runs[0][2:] == runs[1][2:]
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])
    odd = st.capture(1)
    st.replay(odd, 1)
    with pytest.raises(ValueError):
        st.replay(odd, 1)


def test_wall_mounted_cube_fp64_matches_oracle(cuda):
    """The nest standing on the floor with a cube resolved on the fine level:
    inlet/outlet/floor/slip-top boundaries on the coarse level, floor and cube
    on the fine level, coarse-to-fine cells next to the floor moved up with
    offset (0, 0, -1).  FP64, 15 coarse steps vs the oracle on every fine fluid
    node and every active coarse fluid node."""
    from tests.test_oracle_refinement import refined_oracle_run_walls, small_cube_case
    from virtualfluids_amd import geometry

    steps = 15
    print(f"[wall nest fp64 x {steps}]", flush=True)
    case = small_cube_case()
    nest = case.nest
    Nc, Nf = int(np.prod(nest.coarse_dims)), int(np.prod(nest.fine_dims))
    rng = np.random.default_rng(12)
    Fc = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nc, 1)) + rng.normal(0, 1e-4, (Nc, 27)) * cases.W
    Ff = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nf, 1)) + rng.normal(0, 1e-4, (Nf, 27)) * cases.W
    st = vfx.RefinedStepper(nest, case.omega, storage="fp64", coarse_solid=case.coarse_solid,
                            fine_solid=case.fine_solid, coarse_bcs=geometry.boundary_conditions(case.coarse_bcs),
                            fine_bcs=geometry.boundary_conditions(case.fine_bcs))
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    st.step(steps)
    torch.cuda.synchronize()
    got_c = st.coarse.read_populations(st.even_c).cpu().numpy()
    got_f = st.fine.read_populations(st.even_f).cpu().numpy()
    ref_c, ref_f = refined_oracle_run_walls(case, Fc, Ff, steps, case.omega)
    act = (nest.coarse_state == 0) & ~case.coarse_solid
    ffl = ~case.fine_solid
    ec, ef = rel(got_c[act], ref_c[act]), rel(got_f[ffl], ref_f[ffl])
    print(f"coarse {ec:.3e} fine {ef:.3e}")
    assert ec <= REL and ef <= REL


def _cube_stepper(storage, merged):
    """The small wall-mounted cube nest from a perturbed uniform flow."""
    from tests.test_oracle_refinement import small_cube_case
    from virtualfluids_amd import geometry

    case = small_cube_case()
    nest = case.nest
    Nc, Nf = int(np.prod(nest.coarse_dims)), int(np.prod(nest.fine_dims))
    rng = np.random.default_rng(12)
    Fc = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nc, 1)) + rng.normal(0, 1e-4, (Nc, 27)) * cases.W
    Ff = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nf, 1)) + rng.normal(0, 1e-4, (Nf, 27)) * cases.W
    uref = (case.u_ref, 0.0, 0.0) if storage == "fp16" else (0.0, 0.0, 0.0)
    st = vfx.RefinedStepper(nest, case.omega, storage=storage, coarse_solid=case.coarse_solid,
                            fine_solid=case.fine_solid, coarse_bcs=geometry.boundary_conditions(case.coarse_bcs),
                            fine_bcs=geometry.boundary_conditions(case.fine_bcs), reference_velocity=uref,
                            merged=merged)
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    return case, st, Fc, Ff


def _shear_stepper(storage, merged):
    nest = shear_wave_nest()
    Fc, Ff, _ = shear_wave_init(nest, 0.02)
    st = vfx.RefinedStepper(nest, 1.9, storage=storage, merged=merged)
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    return None, st, Fc, Ff


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
@pytest.mark.parametrize("make", [_cube_stepper, _shear_stepper], ids=["cube", "shear"])
def test_merged_schedule_equals_separate_launches(cuda, storage, make):
    """The merged coarse step (coarse and first fine step in one launch, each
    phase's boundary groups of both levels in one launch, fine-to-coarse and
    coarse-to-fine in one launch) leaves the same bytes as the launch-per-
    kernel schedule, on the wall-mounted cube nest (boundary sets) and the
    shear-wave nest (no boundary groups); eager and graph-replayed."""
    print(f"[merged refined schedule {make.__name__} {storage}]", flush=True)
    _, a, _, _ = make(storage, False)
    _, b, _, _ = make(storage, True)
    assert b._one_interface
    a.step(6)
    b.step(3)
    g = b.capture(3)
    b.replay(g, 3)
    torch.cuda.synchronize()
    assert (a.even_c, a.even_f) == (b.even_c, b.even_f)
    for la, lb in ((a.coarse, b.coarse), (a.fine, b.fine)):
        diff = int((la.f.view(torch.uint8) != lb.f.view(torch.uint8)).sum())
        print(f"  differing bytes {diff}", flush=True)
        assert diff == 0


def test_merged_schedule_fp64_matches_oracle(cuda):
    from tests.test_oracle_refinement import refined_oracle_run_walls

    steps = 15
    print(f"[merged refined fp64 vs oracle x {steps}]", flush=True)
    case, st, Fc, Ff = _cube_stepper("fp64", True)
    st.step(steps)
    torch.cuda.synchronize()
    got_c = st.coarse.read_populations(st.even_c).cpu().numpy()
    got_f = st.fine.read_populations(st.even_f).cpu().numpy()
    ref_c, ref_f = refined_oracle_run_walls(case, Fc, Ff, steps, case.omega)
    act = (case.nest.coarse_state == 0) & ~case.coarse_solid
    ffl = ~case.fine_solid
    ec, ef = rel(got_c[act], ref_c[act]), rel(got_f[ffl], ref_f[ffl])
    print(f"coarse {ec:.3e} fine {ef:.3e}")
    assert ec <= REL and ef <= REL


@pytest.mark.parametrize("merged", [False, True])
def test_capture_before_any_eager_step(cuda, merged):
    """A graph recorded straight after construction (boundary and interface
    tables first used inside the capture) replays like eager stepping."""
    print(f"[refined capture first merged={merged}]", flush=True)
    _, a, _, _ = _cube_stepper("fp32", merged)
    _, b, _, _ = _cube_stepper("fp32", merged)
    g = b.capture(2)
    b.replay(g, 2)
    a.step(2)
    torch.cuda.synchronize()
    assert torch.equal(a.coarse.f, b.coarse.f) and torch.equal(a.fine.f, b.fine.f)


@pytest.mark.parametrize("merged", [False, True])
def test_autotuned_levels_give_the_same_bytes(cuda, merged):
    """RefinedStepper.autotune picks each level's launch shape and restores the
    populations: stepping on gives the bytes of the untuned stepper."""
    print(f"[refined autotune merged={merged}]", flush=True)
    _, a, _, _ = _cube_stepper("fp16", merged)
    _, b, _, _ = _cube_stepper("fp16", merged)
    tuned = b.autotune(steps=2)
    print(f"  {tuned}", flush=True)
    assert set(tuned) == {"coarse", "fine"}
    a.step(3)
    b.step(3)
    torch.cuda.synchronize()
    assert torch.equal(a.coarse.f, b.coarse.f) and torch.equal(a.fine.f, b.fine.f)


@pytest.mark.parametrize("tiles", [(64, 64), (128, 128), (64, 256)])
def test_two_level_launch_in_the_levels_tile_width(cuda, tiles):
    """The two-level launch runs in the levels' tile width where they agree
    (else 256 nodes): same bytes as the built-in shape."""
    print(f"[two-level launch tiles {tiles}]", flush=True)
    _, a, _, _ = _cube_stepper("fp32", True)
    _, b, _, _ = _cube_stepper("fp32", True)
    b.coarse.tile_nodes, b.fine.tile_nodes = tiles
    b.coarse.xcd_run, b.fine.xcd_run = 2, 8
    a.step(3)
    b.step(3)
    torch.cuda.synchronize()
    assert torch.equal(a.coarse.f, b.coarse.f) and torch.equal(a.fine.f, b.fine.f)


def _cube_stepper_kw(storage, **kw):
    from tests.test_oracle_refinement import small_cube_case
    from virtualfluids_amd import geometry

    case = small_cube_case()
    nest = case.nest
    Nc, Nf = int(np.prod(nest.coarse_dims)), int(np.prod(nest.fine_dims))
    rng = np.random.default_rng(12)
    Fc = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nc, 1)) + rng.normal(0, 1e-4, (Nc, 27)) * cases.W
    Ff = np.tile(oracle.equilibrium(0.0, case.u_ref, 0.0, 0.0), (Nf, 1)) + rng.normal(0, 1e-4, (Nf, 27)) * cases.W
    uref = (case.u_ref, 0.0, 0.0) if storage == "fp16" else (0.0, 0.0, 0.0)
    st = vfx.RefinedStepper(nest, case.omega, storage=storage, coarse_solid=case.coarse_solid,
                            fine_solid=case.fine_solid, coarse_bcs=geometry.boundary_conditions(case.coarse_bcs),
                            fine_bcs=geometry.boundary_conditions(case.fine_bcs), reference_velocity=uref, **kw)
    st.coarse.write_populations(Fc, even=True)
    st.fine.write_populations(Ff, even=True)
    return st


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_fused_outlet_equals_separate_launches(cuda, storage):
    """The coarse outlet inside the coarse collide-stream kernel (taken at the
    next coarse gather, applied once more at the end of each step call) leaves
    the bytes of the separate pressure kernel; split calls, eager and
    graph-replayed."""
    print(f"[refined fused outlet {storage}]", flush=True)
    a = _cube_stepper_kw(storage)
    b = _cube_stepper_kw(storage, fuse_outlet=True)
    assert b._plan_c.pre is not None and b._plan_c.pre.bc.n > 0
    a.step(7)
    b.step(2)
    b.step(1)
    g = b.capture(2)
    b.replay(g, 2)
    b.step(2)
    torch.cuda.synchronize()
    assert (a.even_c, a.even_f) == (b.even_c, b.even_f)
    assert torch.equal(a.coarse.f.view(torch.uint8), b.coarse.f.view(torch.uint8))
    assert torch.equal(a.fine.f.view(torch.uint8), b.fine.f.view(torch.uint8))
