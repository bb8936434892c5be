"""Ensembles of independent box domains advanced by one launch (libvfx batch API).

Each domain is a box whose outer layer is wall (moving-wall halfway bounce-back
fused into the collide-stream kernel, lid on the z = nz-1 face) with its own
relaxation rate; all domains share one device pool of populations.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as V
from .lattice import BYTES, STORAGE, STORAGE_DTYPE, Lattice, _stream, pack_mask


def cavity_ensemble(n_domains, seed=7, lmin=16, lmax=96, u_lid=0.05, re=(50.0, 1000.0)):
    """Ensemble of lid-driven cavities: edge length L log-uniform in [lmin, lmax]
    (rounded), aspect ratio 1, 1:1:2 or 2:1:1 (x:y:z), Re uniform in `re`,
    viscosity u_lid * L / Re.  Returns (fluid dims list, omegas, lids)."""
    rng = np.random.default_rng(seed)
    aspects = [(1, 1, 1), (1, 1, 2), (2, 1, 1)]
    dims, omegas, lids = [], [], []
    for _ in range(n_domains):
        L = int(round(np.exp(rng.uniform(np.log(lmin), np.log(lmax)))))
        a = aspects[int(rng.integers(0, 3))]
        Re = float(rng.uniform(*re))
        nu = u_lid * L / Re
        dims.append((L * a[0], L * a[1], L * a[2]))
        omegas.append(1.0 / (3.0 * nu + 0.5))
        lids.append((u_lid, 0.0, 0.0))
    return dims, np.array(omegas), np.array(lids)


class Graph:
    """An instantiated HIP graph of captured steps.  `parity` records the time
    step parity the recording started from: a graph of an odd number of steps
    replays correctly only from that parity."""

    def __init__(self, handle, parity=None):
        self.handle = handle
        self.parity = parity

    def launch(self, stream=None):
        V.check(V.lib().vfx_graph_launch(self.handle, _stream(stream)), "vfx_graph_launch")

    def __del__(self):
        if getattr(self, "handle", None):
            V.lib().vfx_graph_destroy(self.handle)
            self.handle = None


def capture(record) -> Graph:
    """Record the libvfx launches that `record(stream)` issues on a private
    stream into a HIP graph (nothing executes while recording)."""
    L = V.lib()
    s = C.c_void_p()
    V.check(L.vfx_stream_create(C.byref(s)), "vfx_stream_create")
    g = C.c_void_p()
    try:
        V.check(L.vfx_capture_begin(s), "vfx_capture_begin")
        try:
            record(s.value)
        finally:
            rc = L.vfx_capture_end(s, C.byref(g))
        V.check(rc, "vfx_capture_end")
    finally:
        L.vfx_stream_destroy(s)
    return Graph(g.value)


class Ensemble:
    def __init__(self, dims, omegas, lids, storage="fp16", fp16_scale=2.0 ** 10, variant="default", device="cuda",
                 limiter=(0.01, 0.01, 0.01)):
        dims = [tuple(int(v) for v in d) for d in dims]
        n = len(dims)
        if not (len(omegas) == n and len(lids) == n):
            raise ValueError("one omega and one lid velocity per domain")
        self.storage = storage
        self.dims = dims
        self.device = torch.device(device)
        self.sizes = np.array([d[0] * d[1] * d[2] for d in dims], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(27 * self.sizes)[:-1]]).astype(np.int64)
        self.pool_elements = int(27 * self.sizes.sum())
        self.pool = torch.zeros(self.pool_elements, dtype=STORAGE_DTYPE[storage], device=self.device)
        arr = (V.VfxDomain * n)()
        for i, d in enumerate(dims):
            arr[i].f_offset = int(self.offsets[i])
            arr[i].dims[:] = list(d)
            arr[i].omega = float(omegas[i])
            arr[i].lid_velocity[:] = [float(x) for x in lids[i]]
        self.omegas = np.asarray(omegas, dtype=np.float64)
        self.lids = np.asarray(lids, dtype=np.float64)
        lim = (C.c_double * 3)(*limiter)
        h = C.c_void_p()
        V.check(V.lib().vfx_batch_create(self.pool.data_ptr(), self.pool_elements, STORAGE[storage], float(fp16_scale), arr, n,
                                         C.cast(lim, C.c_void_p), 0 if variant == "default" else 1, C.byref(h)),
                "vfx_batch_create")
        self.handle = h.value
        self.n = n
        self.even = True
        # interior (colliding) nodes: the LUPs of one step
        self.fluid_nodes = int(sum((d[0] - 2) * (d[1] - 2) * (d[2] - 2) for d in dims))

    def __del__(self):
        if getattr(self, "handle", None):
            V.lib().vfx_batch_destroy(self.handle)
            self.handle = None

    @property
    def bytes_per_lup(self):
        return 54 * BYTES[self.storage]

    def init_rest(self, stream=None):
        V.check(V.lib().vfx_batch_init_rest(self.handle, _stream(stream)), "vfx_batch_init_rest")
        self.even = True

    def set_launch(self, block, grid):
        V.check(V.lib().vfx_batch_set_launch(self.handle, int(block), int(grid)), "vfx_batch_set_launch")

    def launch_shape(self):
        b, g = C.c_uint32(), C.c_uint32()
        V.check(V.lib().vfx_batch_get_launch(self.handle, C.byref(b), C.byref(g)), "vfx_batch_get_launch")
        return b.value, g.value

    def autotune(self, steps=4, stream=None):
        b, g, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        V.check(V.lib().vfx_batch_autotune(self.handle, int(self.even), int(steps), _stream(stream), C.byref(b), C.byref(g),
                                           C.byref(ms)), "vfx_batch_autotune")
        return b.value, g.value, ms.value

    @property
    def xcd_run(self) -> int:
        """Tiles per XCD run of the batched launch (1 = plain order)."""
        r = C.c_uint32()
        V.check(V.lib().vfx_batch_get_xcd_run(self.handle, C.byref(r)), "vfx_batch_get_xcd_run")
        return r.value

    @xcd_run.setter
    def xcd_run(self, run: int):
        V.check(V.lib().vfx_batch_set_xcd_run(self.handle, int(run)), "vfx_batch_set_xcd_run")

    def step(self, n=1, stream=None):
        for _ in range(n):
            V.check(V.lib().vfx_batch_step(self.handle, int(self.even), _stream(stream)), "vfx_batch_step")
            self.even = not self.even

    def step_per_domain(self, n=1, stream=None):
        """The same update as one launch per domain (the per-domain baseline)."""
        L = V.lib()
        s = _stream(stream)
        for _ in range(n):
            for d in range(self.n):
                V.check(L.vfx_batch_step_domain(self.handle, d, int(self.even), s), "vfx_batch_step_domain")
            self.even = not self.even

    def capture(self, n_steps, per_domain=False) -> Graph:
        g = C.c_void_p()
        V.check(V.lib().vfx_batch_capture(self.handle, int(n_steps), int(self.even), int(per_domain), C.byref(g)),
                "vfx_batch_capture")
        return Graph(g.value, parity=self.even)

    def replay(self, graph: Graph, n_steps, stream=None):
        """Launch a graph captured from the current parity; advances the parity."""
        if graph.parity is not None and graph.parity != self.even:
            raise ValueError("graph was recorded from the other time-step parity")
        graph.launch(stream)
        if n_steps % 2:
            self.even = not self.even

    def lattice(self, domain) -> Lattice:
        """A Lattice view over one domain's box (shares the pool memory), for
        reading populations and macroscopic fields and for the diagnostics.
        The wall layer is marked solid, so the diagnostics skip it (the batched
        step leaves values there that belong to no node)."""
        s = V.VfxLattice()
        V.check(V.lib().vfx_batch_lattice(self.handle, int(domain), C.byref(s)), "vfx_batch_lattice")
        lat = Lattice.__new__(Lattice)
        lat.storage = self.storage
        lat.variant = "default"
        lat.device = self.device
        lat.omega = float(self.omegas[domain])
        lat.limiter = (0.01, 0.01, 0.01)
        lat.force = (0.0, 0.0, 0.0)
        lat.fp16_scale = float(s.fp16_scale)
        lat.direct = True
        nx, ny, nz = self.dims[domain]
        wall = np.ones((nz, ny, nx), dtype=bool)
        wall[1:-1, 1:-1, 1:-1] = False
        lat.solid_mask = torch.from_numpy(pack_mask(wall).view(np.int32)).to(self.device)
        lat.nb = None
        lat.fluid = None
        lat.dims = self.dims[domain]
        lat.n_nodes = int(self.sizes[domain])
        lat.n_fluid = (nx - 2) * (ny - 2) * (nz - 2)
        off = int(self.offsets[domain])
        lat.f = self.pool[off: off + 27 * lat.n_nodes]
        lat.macro = None
        lat._struct = None
        return lat
