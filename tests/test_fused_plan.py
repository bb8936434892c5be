"""Host side of the fused boundary step (virtualfluids_amd.FusedPlan), without
a GPU: which entries the collide-stream kernel may apply itself, the node
bit mask / word base that find an entry from its node, and the split into
fused and residual boundaries."""
from types import SimpleNamespace

import numpy as np

import virtualfluids_amd as vfx
from tests.test_geometry import small_channel
from virtualfluids_amd import geometry
from virtualfluids_amd.lattice import fusable_post_entries, fusable_pressure_entries


def _lat(ch):
    return SimpleNamespace(dims=ch.dims, solid_host=ch.solid, n_nodes=int(np.prod(ch.dims)))


def _bcs(ch):
    return geometry.boundary_conditions(ch.bcs, device="cpu")


def test_group_index_finds_every_entry():
    rng = np.random.default_rng(5)
    n_nodes = 5000
    nodes = np.sort(rng.choice(n_nodes, 700, replace=False))
    bc = vfx.BoundaryCondition(1, nodes, np.full((27, nodes.size), -1.0), device="cpu")
    g = vfx.FusedGroup(bc, n_nodes)
    mask = g.node_mask.numpy().view(np.uint32)
    base = g.word_base.numpy().view(np.uint32)
    pos = {int(k): i for i, k in enumerate(nodes)}
    for k in range(n_nodes):
        w, b = k >> 5, k & 31
        member = bool((int(mask[w]) >> b) & 1)
        assert member == (k in pos)
        if member:
            below = int(mask[w]) & ((1 << b) - 1)
            assert int(base[w]) + bin(below).count("1") == pos[k]


def test_group_rejects_unsorted_or_repeated_nodes():
    for nodes in ([3, 1, 7], [1, 4, 4]):
        bc = vfx.BoundaryCondition(1, nodes, np.full((27, 3), -1.0), device="cpu")
        try:
            vfx.FusedGroup(bc, 100)
        except ValueError:
            continue
        raise AssertionError(f"{nodes} accepted")


def test_channel_entries_are_all_fusable():
    """Every wall link of the channel ends in a solid node and the outlet's
    inward neighbour sits in the same 64-node wavefront: all entries fuse."""
    ch = small_channel()
    lat = _lat(ch)
    bcs = _bcs(ch)
    for b in bcs:
        ok = fusable_pressure_entries(lat, b) if b.pre_collision else fusable_post_entries(lat, b)
        assert ok.all(), b.kind
    plan = vfx.FusedPlan.build(lat, bcs)
    assert len(plan.post) == 2 and plan.pre is not None
    assert plan.residual_post == [] and plan.residual_pre == []
    assert sorted(g.bc.kind for g in plan.post) == [1, 2]
    assert sum(g.bc.n for g in plan.post) == sum(b.n for b in bcs if not b.pre_collision)
    for g in plan.post + [plan.pre]:
        assert np.all(np.diff(g.bc.nodes.astype(np.int64)) > 0)


def test_link_into_fluid_is_not_fusable():
    """A wall link whose far node is fluid (an element another thread writes
    in the same launch) keeps its entry in a separate boundary kernel."""
    ch = small_channel()
    solid = ch.solid.copy()
    b = next(t for t in ch.bcs if t.kind == 1)
    i = b.nodes.size // 2
    k = int(b.nodes[i])
    d = int(np.flatnonzero(b.q27[:, i] >= 0)[0])
    nx, ny, nz = ch.dims
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)
    far = (x + vfx.CX[d]) % nx + nx * ((y + vfx.CY[d]) % ny + ny * ((z + vfx.CZ[d]) % nz))
    assert solid[far]
    solid[far] = False
    lat = SimpleNamespace(dims=ch.dims, solid_host=solid, n_nodes=solid.size)
    bc = vfx.BoundaryCondition(1, b.nodes, b.q27, device="cpu")
    ok = fusable_post_entries(lat, bc)
    assert not ok[i] and ok.sum() >= ok.size - 27
    plan = vfx.FusedPlan.build(lat, [bc])
    fused = plan.post[0].bc
    rest = plan.residual_post[0]
    assert fused.n + rest.n == bc.n and k in set(rest.nodes.tolist()) and k not in set(fused.nodes.tolist())
    # the split keeps each entry's q row
    j = int(np.flatnonzero(rest.nodes == k)[0])
    assert np.array_equal(rest._host["q"][:, j], b.q27[:, i].astype(np.float64))


def test_q_rounding_to_one_counts_as_link():
    """q slightly above 1 in float64 rounds to 1 in float32 (a wall link for
    the FP32/FP16 lattices): its far node must be solid as well."""
    dims = (8, 4, 4)
    solid = np.zeros(int(np.prod(dims)), bool)
    k = 1 + 8 * (1 + 4 * 1)
    q = np.full((27, 1), -1.0)
    q[1, 0] = 1.0 + 1e-12
    lat = SimpleNamespace(dims=dims, solid_host=solid, n_nodes=solid.size)
    bc = vfx.BoundaryCondition(1, [k], q, device="cpu")
    assert not fusable_post_entries(lat, bc)[0]
    solid[k + 1] = True
    assert fusable_post_entries(lat, bc)[0]


def test_pressure_wavefront_and_neighbour_rules():
    nx, ny, nz = 130, 4, 4
    n = nx * ny * nz
    lat = SimpleNamespace(dims=(nx, ny, nz), solid_host=np.zeros(n, bool), n_nodes=n)
    xs = np.array([0, 1, 63, 64, 65, 127, 128, 129])
    k = xs + nx * (1 + ny * 1)

    def one(normal, step):
        return [bool(fusable_pressure_entries(lat, vfx.BoundaryCondition(
            4, [kk], rho=[0.0], neighbors=[kk + step], normal=normal, device="cpu"))[0]) for kk in k]

    # +x outward normal (dP00): neighbour k - 1, lane != 0
    assert one(1, -1) == [False, True, True, False, True, True, False, True]
    # -x outward normal (dM00): neighbour k + 1, lane != 63, inside the row
    assert one(2, +1) == [True, True, False, True, True, False, True, False]
    p = vfx.BoundaryCondition(4, k[[1, 4, 7]], rho=np.zeros(3), neighbors=k[[1, 4, 7]] - 1, normal=1, device="cpu")
    assert fusable_pressure_entries(lat, p).all()
    # a neighbour that is not the row neighbour, or is solid, or another pressure node
    q = vfx.BoundaryCondition(4, k[[1, 2]], rho=np.zeros(2), neighbors=k[[1, 2]] - nx, normal=1, device="cpu")
    assert not fusable_pressure_entries(lat, q).any()
    lat.solid_host[k[4] - 1] = True
    assert fusable_pressure_entries(lat, p).tolist() == [True, False, True]
    lat.solid_host[:] = False
    r = vfx.BoundaryCondition(4, k[[1, 2]], rho=np.zeros(2), neighbors=k[[1, 2]] - 1, normal=1, device="cpu")
    assert fusable_pressure_entries(lat, r).all()
    kk = np.array([k[1], k[1] + 1])
    chain = vfx.BoundaryCondition(4, kk, rho=np.zeros(2), neighbors=kk - 1, normal=1, device="cpu")
    assert not fusable_pressure_entries(lat, chain).any()
    # other face normals are never fused
    y = vfx.BoundaryCondition(4, k[[1]], rho=np.zeros(1), neighbors=k[[1]] - nx, normal=3, device="cpu")
    assert not fusable_pressure_entries(lat, y).any()


def test_pressure_fused_only_as_the_sole_pre_collision_boundary():
    ch = small_channel()
    lat = _lat(ch)
    bcs = _bcs(ch)
    p = next(b for b in bcs if b.pre_collision)
    extra = vfx.BoundaryCondition(6, p.nodes[:2], neighbors=p.kn_host[:2], normal=1, device="cpu")
    plan = vfx.FusedPlan.build(lat, bcs + [extra])
    assert plan.pre is None and plan.residual_pre == [p, extra]


def test_pressure_only_plan():
    ch = small_channel()
    lat = _lat(ch)
    bcs = _bcs(ch)
    plan = vfx.FusedPlan.build(lat, bcs, post=False)
    assert plan.post == [] and plan.pre is not None
    assert plan.residual_post == [b for b in bcs if not b.pre_collision]
    u = plan.any_mask.numpy().view(np.uint32)
    assert np.array_equal(u, plan.pre.node_mask.numpy().view(np.uint32))
    full = vfx.FusedPlan.build(lat, bcs)
    u = full.any_mask.numpy().view(np.uint32)
    want = np.zeros_like(u)
    for g in full.post + [full.pre]:
        want |= g.node_mask.numpy().view(np.uint32)
    assert np.array_equal(u, want)
