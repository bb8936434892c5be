"""vfx_reduce_mass_momentum_noalloc (the reduction that allocates nothing,
partials in the library's own device array) against vfx_reduce_mass_momentum
(partials in a caller workspace): the same kernels, so the sums must be
identical bit for bit, on two calls in a row, the second on a smaller lattice,
in every storage form and on the sparse layout."""
import ctypes as C

import numpy as np
import pytest
import torch

import virtualfluids_amd as vfx
from tests import cases

V = vfx._lib


def both(lat, even):
    s = lat.struct()
    ws = torch.empty(max(1, V.lib().vfx_reduce_workspace_bytes(C.byref(s)) // 8), dtype=torch.float64,
                     device=lat.device)
    a = torch.full((4,), np.nan, dtype=torch.float64, device=lat.device)
    b = torch.full((4,), np.nan, dtype=torch.float64, device=lat.device)
    stream = torch.cuda.current_stream().cuda_stream
    V.check(V.lib().vfx_reduce_mass_momentum(C.byref(s), int(even), ws.data_ptr(), a.data_ptr(), stream),
            "vfx_reduce_mass_momentum")
    V.check(V.lib().vfx_reduce_mass_momentum_noalloc(C.byref(s), int(even), b.data_ptr(), stream),
            "vfx_reduce_mass_momentum_noalloc")
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def stepped(dims, storage, seed, steps=3, **kw):
    lat = vfx.Lattice(dims, storage=storage, omega=1.8, device="cuda:0", **kw)
    drho, ux, uy, uz = cases.random_flow(dims, seed=seed)
    lat.init_equilibrium(drho, ux, uy, uz)
    st = vfx.Stepper(lat, "k17")
    st.step(steps)
    return lat, st.even


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_noalloc_equals_workspace_reduction_twice(cuda, storage):
    # a large lattice first (many partial blocks), then a smaller one: the
    # second call must not see anything the first left in the array
    for dims, seed in (((96, 80, 72), 3), ((20, 12, 9), 4)):
        lat, even = stepped(dims, storage, seed)
        a, b = both(lat, even)
        print(f"noalloc {storage} {dims}: {b}", flush=True)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (a, b)


@pytest.mark.gpu
def test_noalloc_sparse_layout(cuda):
    dims = (24, 20, 16)
    nbx, nby, nbz, fluid, sidx = cases.sparse_layout(dims, seed=5)
    fields = cases.random_flow(dims, seed=5)
    full = []
    for f in fields:
        g = np.zeros(sidx.size + 1)
        g[sidx] = f
        full.append(g)
    lat = vfx.Lattice(storage="fp64", omega=1.8, device="cuda:0", neighbors=(nbx, nby, nbz), fluid_nodes=fluid)
    lat.init_equilibrium(*full)
    st = vfx.Stepper(lat, "k17")
    st.step(2)
    a, b = both(lat, st.even)
    assert np.array_equal(a, b), (a, b)
    # then a smaller direct lattice on the same GPU
    lat2, even2 = stepped((10, 9, 8), "fp64", 6)
    a2, b2 = both(lat2, even2)
    assert np.array_equal(a2, b2), (a2, b2)


@pytest.mark.gpu
def test_noalloc_rejects_oversize_lattice(cuda):
    # rejected on the host, before anything is launched
    s = V.VfxLattice()
    s.f = 1
    s.storage = V.VFX_FP16_SCALED
    s.fp16_scale = 1024.0
    s.omega = 1.5
    s.direct = 0
    s.neighbor_x = s.neighbor_y = s.neighbor_z = 1
    s.fluid_nodes = 1
    limit = V.lib().vfx_reduce_noalloc_max_items()
    s.n_nodes = limit + 10
    s.n_fluid = limit + 1
    out = torch.zeros(4, dtype=torch.float64, device="cuda:0")
    assert V.lib().vfx_reduce_mass_momentum_noalloc(C.byref(s), 1, out.data_ptr(), None) == V.VFX_EINVAL
    assert b"maximum" in V.lib().vfx_last_error()
