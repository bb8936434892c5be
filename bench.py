#!/usr/bin/env python3
"""Throughput of the MI355X D3Q27 K17 stepper.

One step = one collide-and-stream pass of the K17 cumulant kernel over every
node of a 512 x 256 x 256 lattice (33.5 M nodes), in place (esoteric twist),
inputs resident in HBM.  Three storage forms run on the same lattice and the
same initial flow: FP32 arithmetic on per-lattice-scaled FP16 shifted
populations (the headline form), FP32 on FP32, and plain FP64.

Prints ONE JSON line.  ``value`` is the MLUPS of the headline form;
``roofline`` prices its kernel against MI355X's 8 TB/s with the algorithmic
54 * bytes-per-population bytes per lattice update; ``cpu_baseline`` times the
VirtualFluids reference CPU K17 path (compiled from its sources) on this
host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
DIMS = (512, 256, 256)
OMEGA = 1.9417  # Re = 200 on the mean sphere diameter of the channel-with-spheres case
U0 = 0.05


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--settle-seconds", type=float, default=0.4,
                   help="untimed stepping before the warm-up steps, so the timed steps see steady-state clocks")
    p.add_argument("--forms", default="fp16,fp32,fp32:plain,fp64",
                   help="storage[:plain]; plain = reference operation order")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-autotune", action="store_true",
                   help="time the built-in launch shape instead of the autotuned one")
    p.add_argument("--ensemble", type=int, default=512,
                   help="cavities in the batched-ensemble leg (0 skips it)")
    p.add_argument("--ensemble-steps", type=int, default=100, help="steps per captured graph of the ensemble leg")
    p.add_argument("--sparse-steps", type=int, default=50,
                   help="timed steps of the VirtualFluids sparse-layout leg (0 skips it)")
    p.add_argument("--tgv", type=int, default=256, help="Taylor-Green vortex leg edge (0 skips it)")
    p.add_argument("--tgv-steps", type=int, default=2000)
    p.add_argument("--channel-steps", type=int, default=5000,
                   help="steps of the flow-past-spheres channel leg (0 skips it)")
    p.add_argument("--channel-separate-steps", type=int, default=500,
                   help="FP16 channel steps timed per boundary-kernel setting beside the leg's own")
    p.add_argument("--cube-steps", type=int, default=10000,
                   help="coarse steps of the refined wall-mounted-cube leg (0 skips it)")
    p.add_argument("--cube-schedules", default="separate,merged,fused_outlet",
                   help="refined-cube schedules timed before the long window (the fastest runs it)")
    p.add_argument("--copy-probe", action="store_true",
                   help="only run a plain device copy of known size (reference for HBM counters and bandwidth)")
    return p.parse_args()


def copy_probe(torch, nbytes=1 << 30, reps=10):
    """Plain device-to-device copy: calibrates FETCH_SIZE/WRITE_SIZE counters
    (known bytes per dispatch) and shows the copy bandwidth this GPU reaches."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                      "ms": round(ms, 4), "gbs_read_plus_write": round(2 * nbytes / ms / 1e6, 1)}))


def initial_flow(torch, dims, device):
    """Smooth periodic shear-layer flow at |u| ~ 0.05 (lattice units), FP64."""
    nx, ny, nz = dims
    z = torch.arange(nz, device=device, dtype=torch.float64).view(nz, 1, 1)
    y = torch.arange(ny, device=device, dtype=torch.float64).view(1, ny, 1)
    x = torch.arange(nx, device=device, dtype=torch.float64).view(1, 1, nx)
    kx, ky, kz = 2 * math.pi / nx, 2 * math.pi / ny, 2 * math.pi / nz
    ux = U0 * torch.sin(ky * y) * torch.cos(kz * z) + 0.1 * U0 * torch.sin(kx * x)
    uy = 0.5 * U0 * torch.sin(kx * x) * torch.cos(kz * z) + 0 * y
    uz = 0.5 * U0 * torch.cos(kx * x) * torch.sin(ky * y) + 0 * z
    drho = 1e-3 * torch.cos(kx * x) * torch.cos(ky * y) * torch.cos(kz * z)
    shape = (nz, ny, nx)
    return [t.expand(shape).contiguous().view(-1) for t in (drho, ux, uy, uz)]


def settle(torch, st, seconds):
    """Step (untimed) until `seconds` of wall time have passed.  The first few
    dozen launches after the GPU leaves idle run up to 35% slower than the
    steady state a long simulation sees (power and clock management), so the
    timed steps start after this settling period, then the W warm-up steps."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        st.step(10)
        torch.cuda.synchronize()
        n += 10
    return n


def run_form(torch, vfx, form, fields, steps, warmup, settle_s, autotune=True):
    storage, _, variant = form.partition(":")
    lat = vfx.Lattice(DIMS, storage=storage, omega=OMEGA, device="cuda", variant=variant or "default")
    lat.init_equilibrium(*fields, even=True)
    st = vfx.Stepper(lat, "k17")
    settled = settle(torch, st, settle_s)
    launch = {"tile_nodes": 256, "xcd_run": 0, "autotuned": False}
    if autotune:
        # vfx_k17_autotune: tile width and XCD run of the structured step,
        # timed on the settled state; the timed steps run the fastest
        tile, run, ms_t = lat.autotune(st.even, steps=4)
        launch = {"tile_nodes": tile, "xcd_run": run, "autotuned": True, "autotune_ms_per_step": round(ms_t, 4)}
    st.step(warmup)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    st.step(steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    macro = lat.macroscopic(st.even)
    del lat
    return {"ms_per_step": ms / steps, "wall_s": wall, "macro": macro, "settle_steps": settled, "launch": launch}


def velocity_l2(a, b):
    _, ax, ay, az = a
    _, bx, by, bz = b
    num = ((ax - bx) ** 2 + (ay - by) ** 2 + (az - bz) ** 2).sum()
    den = (bx ** 2 + by ** 2 + bz ** 2).sum()
    return float((num / den).sqrt())


def density_drift(a, b):
    return float((a[0] - b[0]).abs().max())


def mass_error(a, b):
    return float((a[0].sum() - b[0].sum()).abs() / a[0].numel())


def cpu_model():
    """CPU model name, logical CPU count and OMP_NUM_THREADS of this host."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(target_s):
    """VirtualFluids' CPU K17 path, FP64, the way Simulation::run advances one
    level: K17CompressibleNavierStokes::calculate of every block (OpenMP
    guided over blocks, Simulation::calculateBlocks), the block exchange (ghost
    slots filled from the neighbouring blocks, as the local FullDirectConnectors
    do) and swap, on a periodic grid of 32^3 blocks, two blocks per thread.
    The collision and storage are the reference's own compiled sources; the
    exchange copy is restated.  Also reports the rate without the exchange."""
    import oracle

    if not oracle.ref_available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nblocks = max(2, 2 * threads)
    bs = 32
    secs, used, grid = oracle.ref_time_blocks(nblocks, bs, OMEGA, 2)
    per_step = secs / 2
    steps = max(2, int(target_s / max(per_step, 1e-6)))
    secs, used, grid = oracle.ref_time_blocks(nblocks, bs, OMEGA, steps)
    lups = nblocks * bs ** 3 * steps
    secs_nx, _, _ = oracle.ref_time_blocks(nblocks, bs, OMEGA, max(2, steps // 4), exchange=False)
    lups_nx = nblocks * bs ** 3 * max(2, steps // 4)
    return {"value": round(lups / secs / 1e6, 2), "unit": "MLUPS", "cores": used, "kind": "reference",
            "sample": f"{nblocks} blocks of {bs}^3 nodes on a periodic {grid[0]}x{grid[1]}x{grid[2]} block grid x "
                      f"{steps} steps, K17 FP64: calculate + block exchange + swap, {secs:.1f} s",
            "without_exchange_mlups": round(lups_nx / secs_nx / 1e6, 2),
            "cpu": cpu_model()}


def hbm_ceiling_leg(torch, nbytes, reps=20):
    """What plain streams reach on this GPU for a buffer the size of the FP16
    lattice: a copy between two buffers (torch copy_) and an in-place update of
    one buffer (torch neg_ on FP16, 16-byte vector accesses, every element read
    and written once).  The collide-stream step is in place, so the in-place
    figure is the ceiling its access pattern can approach."""
    a = torch.ones(nbytes // 2, dtype=torch.float16, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream()
    out = {"bytes": nbytes}
    for name, fn in (("copy_two_buffers", lambda: b.copy_(a)), ("in_place_update", lambda: a.neg_())):
        for _ in range(5):
            fn()
        ms = _timed_loop(torch, stream, reps, lambda i: fn()) / reps
        out[name] = {"ms": round(ms, 4), "gbs_read_plus_write": round(2 * nbytes / ms / 1e6, 1),
                     "frac_of_peak": round(2 * nbytes / ms / 1e6 / PEAK_HBM_GBS, 4)}
    del a, b
    torch.cuda.empty_cache()
    return out


def ensemble_leg(torch, vfx, n_domains, graph_steps, reps=2):
    """Lid-driven cavity ensemble (edges log-uniform in [16, 96] from seed 7,
    aspect 1, 1:1:2 or 2:1:1, Re uniform in [50, 1000], lid 0.05), FP16 shifted
    storage, FP32 arithmetic.  Times `graph_steps`-step HIP graphs of (a) the
    batched launch with the autotuned shape and (b) one launch per domain, and
    the same per-domain launches issued eagerly; aggregate MLUPS counts the
    interior (colliding) nodes."""
    dims, omegas, lids = vfx.cavity_ensemble(n_domains, seed=7)
    boxes = [(a + 2, b + 2, c + 2) for a, b, c in dims]
    ens = vfx.Ensemble(boxes, omegas, lids, storage="fp16", device="cuda")
    ens.init_rest()
    block, grid, tune_ms = ens.autotune(steps=4)
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    g_batched = ens.capture(graph_steps)
    g_domain = ens.capture(graph_steps, per_domain=True)
    ens.replay(g_batched, graph_steps)  # warm up both graphs
    ens.replay(g_domain, graph_steps)
    ms_b = timed(lambda: ens.replay(g_batched, graph_steps), reps) / (reps * graph_steps)
    ms_d = timed(lambda: ens.replay(g_domain, graph_steps), reps) / (reps * graph_steps)
    eager_steps = 10
    ms_e = timed(lambda: ens.step_per_domain(1), eager_steps) / eager_steps
    steps_run = 4 + 2 * graph_steps + 2 * reps * graph_steps + eager_steps
    fin = [bool(torch.isfinite(ens.lattice(i).f.float()).all()) for i in range(n_domains)]
    finite = sum(fin)
    lups = ens.fluid_nodes
    lups_finite = sum(int(np.prod(d)) for d, ok in zip(dims, fin) if ok)
    out = {
        "domains": n_domains, "interior_nodes": lups, "graph_steps": graph_steps,
        "launch": {"block": block, "grid": grid, "xcd_run": ens.xcd_run, "autotune_ms_per_step": round(tune_ms, 4)},
        "batched_ms_per_step": round(ms_b, 4), "batched_mlups": round(lups / ms_b / 1e3, 1),
        "per_domain_graph_ms_per_step": round(ms_d, 4), "per_domain_graph_mlups": round(lups / ms_d / 1e3, 1),
        "per_domain_eager_ms_per_step": round(ms_e, 4), "per_domain_eager_mlups": round(lups / ms_e / 1e3, 1),
        "speedup_vs_per_domain_graph": round(ms_d / ms_b, 3),
        "speedup_vs_per_domain_eager": round(ms_e / ms_b, 3),
        "achieved_gbs": round(54 * 2 * lups / ms_b / 1e6, 1),
        "domains_finite_after_run": finite, "steps_run": steps_run,
        "batched_mlups_finite_domains_only": round(lups_finite / ms_b / 1e3, 1),
        "unstable_domains": [{"dims": list(map(int, d)), "omega": round(float(o), 4)}
                             for d, o, ok in zip(dims, omegas, fin) if not ok],
        "omega_range": [round(float(omegas.min()), 4), round(float(omegas.max()), 4)],
    }
    del g_batched, g_domain, ens
    return out


def _timed_loop(torch, stream, chunks, body):
    """Run body(i) for every chunk, timing the launches with HIP events on the
    stream they go to; returns milliseconds."""
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(chunks):
        body(i)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def sparse_leg(torch, vfx, steps=50, warmup=10, forms=("fp16", "fp32", "fp64"), autotune=True):
    """The same 512 x 256 x 256 periodic flow in VirtualFluids' own sparse
    layout: neighborX/Y/Z arrays (+x/+y/+z neighbour per node, index 0 = dummy
    node) and the fluid-node list, numbered x-fastest as VirtualFluids' grid
    generator numbers a box.  This is the layout the reference's
    runCollision/K17 kernel reads (RunCollision.cuh:74-143); per node it adds the
    fluid index and three neighbour indices (16 B) to the populations, and four
    second-level neighbour reads that mostly hit in cache, except in the
    wavefronts the regular-wavefront table marks (64 consecutive nodes with one
    y and one z stride), which form their indices by arithmetic.  Each run
    tunes the step's XCD run first (vfx_k17_autotune)."""
    nx, ny, nz = DIMS
    N = nx * ny * nz
    k = np.arange(N, dtype=np.int64)
    x, y, z = k % nx, (k // nx) % ny, k // (nx * ny)

    def arrays(first):
        """Neighbour arrays and fluid list with the box's nodes numbered from
        `first` (nodes 0 .. first-1 are the dummy node and unused padding)."""
        def kof(xx, yy, zz):
            return (xx % nx) + nx * ((yy % ny) + ny * (zz % nz)) + first

        nb = [np.zeros(N + first, np.uint32) for _ in range(3)]
        nb[0][first:] = kof(x + 1, y, z)
        nb[1][first:] = kof(x, y + 1, z)
        nb[2][first:] = kof(x, y, z + 1)
        return tuple(nb), (k + first).astype(np.uint32)

    layouts = {1: arrays(1), 64: arrays(64)}
    del k, x, y, z
    base = initial_flow(torch, DIMS, "cuda")

    def fields(first):
        zero = torch.zeros(first, dtype=torch.float64, device="cuda")
        return [torch.cat([zero, f]) for f in base]

    stream = torch.cuda.current_stream()
    out = {"layout": "VirtualFluids sparse (neighbour arrays + fluid-node list), x-fastest numbering",
           "nodes": N, "steps": steps}
    for form in forms:
        s = {"fp16": 2, "fp32": 4, "fp64": 8}[form]
        # the generator's numbering (dummy node 0, nodes from 1, plane stride
        # N + 1) with the regular-wavefront table (the default); for FP16 also
        # without the table, and with the nodes numbered from 64 (stride N + 64),
        # so every row and plane starts on a 128-byte line
        runs = (((True, 1), (False, 1), (True, 64)) if form == "fp16" else ((True, 1),))
        for groups, first in runs:
            nb, fluid = layouts[first]
            lat = vfx.Lattice(storage=form, omega=OMEGA, device="cuda", neighbors=nb, fluid_nodes=fluid,
                              regular_groups=groups)
            lat.init_equilibrium(*fields(first), even=True)
            st = vfx.Stepper(lat, "k17")
            st.step(warmup)
            launch = {"xcd_run": lat.xcd_run, "autotuned": autotune}
            if autotune:
                _, run, ms_t = lat.autotune(st.even, steps=4)
                launch.update(xcd_run=run, autotune_ms_per_step=round(ms_t, 4))
                st.step(2)
            ms = _timed_loop(torch, stream, steps, lambda i: st.step(1)) / steps
            r = {"mlups": round(N / ms / 1e3, 1), "ms_per_step": round(ms, 4),
                 "population_gbs": round(54 * s * N / ms / 1e6, 1),
                 "regular_wavefront_fraction": round(lat.regular_fraction, 4), "launch": launch}
            if first != 1:
                out[form]["rows_on_line_boundaries"] = dict(r, first_node=first, plane_stride=N + first)
            elif groups:
                out[form] = r
            else:
                out[form]["without_regular_table"] = r
            del st, lat
            torch.cuda.empty_cache()
        print(f"# sparse layout {form}: {out[form]['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    return out


def tgv_leg(torch, vfx, n, steps, every=100):
    """Taylor-Green vortex n^3 (u0 = 0.04, Re = 1600), K17 and BGK each in FP64
    and FP32, `steps` steps: kinetic energy and enstrophy every `every` steps
    (FP64 GPU reductions); FP32 against FP64 and E(t)/E0 against the purely
    viscous decay exp(-6 nu k^2 t), k = 2 pi / n."""
    from virtualfluids_amd import geometry

    fields, omega, nu = geometry.taylor_green(n)
    k = 2 * math.pi / n
    stream = torch.cuda.current_stream()
    out = {"grid": f"{n}^3", "omega": round(omega, 6), "steps": steps, "sample_every": every}
    curves = {}
    for coll in ("k17", "bgk"):
        for storage in ("fp64", "fp32"):
            lat = vfx.Lattice((n, n, n), storage=storage, omega=omega)
            lat.init_equilibrium(*fields)
            st = vfx.Stepper(lat, coll)
            ke, en = [], []

            def sample():
                a = lat.kinetic_energy(st.even)
                b = lat.enstrophy(st.even)
                ke.append(a)
                en.append(b)

            sample()
            ms = 0.0
            for _ in range(steps // every):
                ms += _timed_loop(torch, stream, 1, lambda i: st.step(every))
                sample()
            kev = torch.stack(ke).cpu().numpy()
            env = torch.stack(en).cpu().numpy()
            curves[(coll, storage)] = (kev[:, 0] / kev[:, 1], env[:, 0] / env[:, 1])
            out[f"{coll}_{storage}_mlups"] = round(n ** 3 * (steps // every) * every / (ms * 1e-3) / 1e6, 1)
            del lat, st
    t = np.arange(steps // every + 1) * every
    viscous = np.exp(-6 * nu * k * k * t)
    for coll in ("k17", "bgk"):
        e64, z64 = curves[(coll, "fp64")]
        e32, z32 = curves[(coll, "fp32")]
        out[f"{coll}_fp32_ke_rel_err_vs_fp64"] = float(np.abs(e32 / e64 - 1).max())
        out[f"{coll}_fp32_enstrophy_rel_err_vs_fp64"] = float(np.abs(z32 / z64 - 1).max())
        # E(t)/E0 against the purely viscous decay exp(-6 nu k^2 t): equal at
        # early times, the vortex then transfers energy to small scales
        ratio = (e64 / e64[0]) / viscous
        out[f"{coll}_fp64_ke_over_viscous_at"] = {str(int(t[i])): round(float(ratio[i]), 6)
                                                   for i in range(0, len(t), max(1, len(t) // 5))}
    out["viscous_rate_6nuk2"] = 6 * nu * k * k
    return out


def channel_leg(torch, vfx, steps, sample_every=10, forms=("fp64", "fp32", "fp16"), separate_steps=500):
    """Flow past 64 spheres in a 512 x 256 x 256 channel (velocity inlet 0.05
    with 1% noise, pressure outlet, no-slip walls, interpolated bounce-back with
    ray-sphere q), K17 at Re = 200 on the mean sphere diameter, started from a
    uniform stream (FP16: populations stored as deviations from the inflow
    equilibrium); `steps` steps with FP64 running mean/Reynolds-stress sums
    every `sample_every` steps.  The pressure outlet runs inside the
    collide-stream kernel (Stepper(fuse=True, fuse_post=False)), the
    interpolated groups as kernels of their own; `separate_steps` FP16 steps
    with every group separate, and with every group fused, are timed beside
    it.  Reports
    whole-step MLUPS (fluid nodes, all boundary kernels and statistics
    included), the time-averaged mean velocity error of FP32/FP16 storage vs
    FP64, and the FP64 mass/momentum monitor."""
    from virtualfluids_amd import geometry

    ch = geometry.spheres_in_channel()
    bcs = geometry.boundary_conditions(ch.bcs)
    N = int(np.prod(ch.dims))
    fl = torch.from_numpy(~ch.solid).cuda()
    stream = torch.cuda.current_stream()
    zero = torch.zeros(N, dtype=torch.float64, device="cuda")
    u0 = torch.full((N,), ch.u_in, dtype=torch.float64, device="cuda")
    out = {"dims": list(ch.dims), "fluid_nodes": ch.n_fluid, "spheres": len(ch.spheres), "omega": round(ch.omega, 6),
           "steps": steps, "sample_every": sample_every,
           "bc_nodes": {str(b.kind): int(b.nodes.size) for b in ch.bcs}}
    ref_mean = None
    for storage in forms:
        # FP16 stores deviations from the inflow-velocity equilibrium
        uref = (ch.u_in, 0.0, 0.0) if storage == "fp16" else (0.0, 0.0, 0.0)
        lat = vfx.Lattice(ch.dims, storage=storage, omega=ch.omega, solid=ch.solid, device="cuda",
                          reference_velocity=uref)
        lat.init_equilibrium(zero, u0, zero, zero)
        st = vfx.Stepper(lat, "k17", bcs, fuse=True, fuse_post=False)
        plan = st.fused_plan()
        stats = vfx.Statistics(lat)
        mm0 = lat.mass_momentum(st.even).cpu().numpy()

        def chunk(i):
            st.step(sample_every)
            stats.accumulate(st.even)

        ms = _timed_loop(torch, stream, steps // sample_every, chunk)
        mm1 = lat.mass_momentum(st.even).cpu().numpy()
        mean = stats.mean()[:, fl]
        r = {"mlups": round(ch.n_fluid * (steps // sample_every) * sample_every / (ms * 1e-3) / 1e6, 1),
             "ms_per_step": round(ms / ((steps // sample_every) * sample_every), 4),
             "mass_change_per_fluid_node": float((mm1[0] - mm0[0]) / ch.n_fluid),
             "mean_ux": float(mean[0].mean()),
             "fused_entries": {str(g.bc.kind): g.bc.n for g in plan.post + ([plan.pre] if plan.pre else [])},
             "separate_kernels": len(plan.residual_post) + len(plan.residual_pre)}
        if storage == "fp16" and separate_steps > 0:
            # the leg's own policy without the statistics, the same channel
            # with one kernel per boundary group, and with every group inside
            # the collide-stream kernel
            n_sep = separate_steps // 10 * 10
            for name, kw in (("pressure_fused_without_statistics", {"fuse": True, "fuse_post": False}),
                             ("separate_bc_kernels", {}), ("all_groups_fused", {"fuse": True})):
                su = vfx.Stepper(lat, "k17", bcs, **kw)
                su.step(20)
                ms_u = _timed_loop(torch, stream, n_sep // 10, lambda i: su.step(10))
                r[name] = {"steps": n_sep, "ms_per_step": round(ms_u / n_sep, 4),
                           "mlups": round(ch.n_fluid * n_sep / (ms_u * 1e-3) / 1e6, 1)}
                del su
        if ref_mean is None:
            ref_mean = mean
        else:
            r["mean_velocity_l2_vs_fp64"] = float(torch.linalg.vector_norm(mean - ref_mean) /
                                                  torch.linalg.vector_norm(ref_mean))
        out[storage] = r
        print(f"# channel {storage}: {r['mlups']:.0f} MLUPS (whole step)", file=sys.stderr, flush=True)
        del lat, st, stats, mean
        torch.cuda.empty_cache()
    return out


def cube_leg(torch, vfx, coarse_steps, sample_every=50, forms=("fp64", "fp32", "fp16"),
             schedules=("separate", "merged", "fused_outlet")):
    """Wall-mounted cube in a boundary layer on a 256 x 128 x 128 coarse grid
    with a 2:1 nest of 192 x 128 x 96 fine nodes around the cube (power-law
    inflow with 2% noise, pressure outlet, no-slip floor, slip top, periodic
    spanwise), K17 at Re = 10^4 on the cube height.  Every coarse step (two
    fine steps, one coarse step, both levels' boundary kernels, fine-to-coarse
    and coarse-to-fine) is recorded into a HIP graph of `sample_every` coarse
    steps.  Three schedules (launch per kernel; merged: the coarse step and
    the first fine step in one launch, each phase's boundary groups of both
    levels in one launch, fine-to-coarse and coarse-to-fine in one launch;
    launch per kernel with the coarse outlet inside the coarse collide-stream
    kernel; `schedules` picks which) are timed over 500 coarse steps with
    the levels' launch shapes autotuned; the fastest runs the long window.  Time-averaged
    velocity over the second half on both levels; the
    FP32 and FP16 forms (FP16: deviations from the u_ref equilibrium) against
    FP64.  At Re = 10^4 the wake is chaotic: two forms that differ by rounding
    part ways once the rounding has grown, so the FP32 distance is the yardstick
    for the FP16 one over a window this short.  MLUPS
    counts coarse fluid nodes once and fine fluid nodes twice per coarse step."""
    from virtualfluids_amd import geometry

    case = geometry.wall_mounted_cube()
    nest = case.nest
    H = float(case.cube[3])
    stream = torch.cuda.current_stream()
    cbcs = geometry.boundary_conditions(case.coarse_bcs)
    fbcs = geometry.boundary_conditions(case.fine_bcs)

    def profile(z):
        return case.u_ref * np.clip((z - 0.5) / H, 0.0, None) ** 0.2

    nx, ny, nz = nest.coarse_dims
    zc = np.repeat(np.arange(nz, dtype=np.float64), nx * ny)
    zf = nest.fine_positions()[2]
    out = {"coarse_dims": list(nest.coarse_dims), "fine_dims": list(nest.fine_dims), "omega_coarse": case.omega,
           "coarse_steps": coarse_steps, "fine_steps": 2 * coarse_steps, "graph_coarse_steps": sample_every,
           "interface_cells": {"coarse_to_fine": int(nest.cf_coarse.size), "fine_to_coarse": int(nest.fc_coarse.size)}}
    means, early_means = {}, {}
    for storage in forms:
        uref = (case.u_ref, 0.0, 0.0) if storage == "fp16" else (0.0, 0.0, 0.0)
        def stepper(schedule):
            st = vfx.RefinedStepper(nest, case.omega, storage=storage, coarse_solid=case.coarse_solid,
                                    fine_solid=case.fine_solid, coarse_bcs=cbcs, fine_bcs=fbcs,
                                    reference_velocity=uref, merged=schedule == "merged",
                                    fuse_outlet=schedule == "fused_outlet")
            for lat, z in ((st.coarse, zc), (st.fine, zf)):
                n = z.size
                lat.init_equilibrium(np.zeros(n), profile(z), np.zeros(n), np.zeros(n))
            return st

        # both schedules with their levels' launch shapes tuned, timed over a
        # short window without statistics; the faster one runs the long window
        def timed(schedule):
            st = stepper(schedule)
            tuned = st.autotune()
            graph = st.capture(sample_every)
            st.replay(graph, sample_every)
            ms = _timed_loop(torch, stream, n_sep, lambda i: st.replay(graph, sample_every))
            launches = st.launches_per_step
            del st, graph
            torch.cuda.empty_cache()
            return ms, launches, tuned

        n_sep = 10
        sched = {name: timed(name) for name in schedules}
        best = min(sched, key=lambda k: sched[k][0])
        tuned = sched[best][2]
        st = stepper(best)
        for lat, name in ((st.coarse, "coarse"), (st.fine, "fine")):
            lat.tile_nodes, lat.xcd_run = tuned[name][0], tuned[name][1]
        graph = st.capture(sample_every)
        sc, sf = vfx.Statistics(st.coarse), vfx.Statistics(st.fine)
        sce, sfe = vfx.Statistics(st.coarse), vfx.Statistics(st.fine)
        chunks = coarse_steps // sample_every
        early = max(1, min(chunks // 2 - 1, chunks // 20))

        def chunk(i):
            st.replay(graph, sample_every)
            if i >= chunks // 2:
                sc.accumulate(st.even_c)
                sf.accumulate(st.even_f)
            elif i < early:
                sce.accumulate(st.even_c)
                sfe.accumulate(st.even_f)

        ms = _timed_loop(torch, stream, chunks, chunk)
        lups = st.lups_per_step * chunks * sample_every
        act = torch.from_numpy((nest.coarse_state == 0) & ~case.coarse_solid).cuda()
        ffl = torch.from_numpy(~case.fine_solid).cuda()
        means[storage] = (sc.mean()[:, act], sf.mean()[:, ffl])
        early_means[storage] = (sce.mean()[:, act], sfe.mean()[:, ffl])
        r = {"mlups": round(lups / (ms * 1e-3) / 1e6, 1), "ms_per_coarse_step": round(ms / (chunks * sample_every), 4),
             "schedule": best, "launches_per_coarse_step": st.launches_per_step,
             "launch": {k: {"tile_nodes": v[0], "xcd_run": v[1]} for k, v in tuned.items()},
             "schedules_without_statistics": {
                 k: {"ms_per_coarse_step": round(v[0] / (n_sep * sample_every), 4), "launches_per_coarse_step": v[1],
                     "coarse_steps": n_sep * sample_every} for k, v in sched.items()},
             "finite": bool(torch.isfinite(means[storage][1]).all())}
        out[storage] = r
        print(f"# cube {storage}: {r['mlups']:.0f} MLUPS (refined, fine counted twice)", file=sys.stderr, flush=True)
        del st, sc, sf, sce, sfe, graph
        torch.cuda.empty_cache()
    out["early_window_coarse_steps"] = [sample_every, early * sample_every]
    out["late_window_coarse_steps"] = [(chunks // 2 + 1) * sample_every, chunks * sample_every]
    if "fp64" in means:
        for s in forms:
            if s == "fp64":
                continue
            for lvl, name in ((0, "coarse"), (1, "fine")):
                # early window: the forms still follow one trajectory, so the
                # distance is the storage error; late window: dominated by the
                # chaotic divergence of two runs that differ by rounding
                for tag, mm in (("early", early_means), ("late", means)):
                    a, b = mm[s][lvl], mm["fp64"][lvl]
                    out[s][f"mean_velocity_l2_vs_fp64_{name}_{tag}"] = float(torch.linalg.vector_norm(a - b) /
                                                                             torch.linalg.vector_norm(b))
    return out


def traffic_from_profiles(form):
    """HBM bytes per launch of the headline kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "k17_fp16_512x256x256_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch[{form}]"):
            return float(line.split("=")[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import virtualfluids_amd as vfx

    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(torch)
        return
    n_nodes = DIMS[0] * DIMS[1] * DIMS[2]
    fields = initial_flow(torch, DIMS, "cuda")
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for form in forms:
        r = run_form(torch, vfx, form, fields, args.steps, args.warmup, args.settle_seconds, not args.no_autotune)
        r["mlups"] = n_nodes / (r["ms_per_step"] * 1e-3) / 1e6
        r["b_alg"] = 54 * vfx.BYTES[form.split(":")[0]]
        r["gbs"] = r["b_alg"] * n_nodes / (r["ms_per_step"] * 1e-3) / 1e9
        res[form] = r
        print(f"# {form}: {r['ms_per_step']:.3f} ms/step  {r['mlups']:.0f} MLUPS  {r['gbs']:.0f} GB/s "
              f"({100 * r['gbs'] / PEAK_HBM_GBS:.1f}% of peak)", file=sys.stderr, flush=True)
    head = forms[0]
    h = res[head]
    out = {
        "metric": "D3Q27 K17 MLUPS + %HBM peak; FP16-store/FP32 vs FP64; batched vs per-domain",
        "value": round(h["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not head.startswith("fp64") else "f64",
        "data": "synthetic smooth periodic shear flow, |u|~0.05, equilibrium start",
        "config": {"workload": f"K17 collide+stream, {DIMS[0]}x{DIMS[1]}x{DIMS[2]} D3Q27 periodic box, "
                               f"omega={OMEGA}, storage={head} (shifted f-w)",
                   "nodes": n_nodes, "storage": head, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(h["gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(h["gbs"] / PEAK_HBM_GBS, 4), "traffic": traffic_from_profiles(head),
                     "bytes_per_lup": h["b_alg"]},
        "forms": {f: {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
                      "achieved_gbs": round(r["gbs"], 1), "frac_hbm_peak": round(r["gbs"] / PEAK_HBM_GBS, 4),
                      "bytes_per_lup": r["b_alg"], "launch": r["launch"]} for f, r in res.items()},
    }
    if "fp64" in res:
        ref = res["fp64"]["macro"]
        for f, r in res.items():
            if f == "fp64":
                continue
            out["forms"][f]["speedup_vs_fp64"] = round(r["mlups"] / res["fp64"]["mlups"], 3)
            out["forms"][f]["velocity_l2_vs_fp64"] = velocity_l2(r["macro"], ref)
            out["forms"][f]["density_drift_vs_fp64"] = density_drift(r["macro"], ref)
            out["forms"][f]["mass_error_per_node_vs_fp64"] = mass_error(r["macro"], ref)
    for r in res.values():
        r.pop("macro", None)
    out["settle"] = {"seconds": args.settle_seconds, "steps_per_form": {f: r["settle_steps"] for f, r in res.items()},
                     "why": "untimed stepping before the warm-up steps: launches in the first ~30 ms after idle run "
                            "up to 35% slower (power/clock ramp)"}
    torch.cuda.empty_cache()
    ceil = hbm_ceiling_leg(torch, 54 * vfx.BYTES["fp16"] * n_nodes // 2)
    out["hbm_ceiling"] = ceil
    out["roofline"]["frac_of_in_place_ceiling"] = round(h["gbs"] / ceil["in_place_update"]["gbs_read_plus_write"], 4)
    print(f"# HBM ceilings on this GPU: two-buffer copy {ceil['copy_two_buffers']['gbs_read_plus_write']:.0f} GB/s, "
          f"in-place update {ceil['in_place_update']['gbs_read_plus_write']:.0f} GB/s", file=sys.stderr, flush=True)
    if args.ensemble > 0:
        out["ensemble"] = ensemble_leg(torch, vfx, args.ensemble, args.ensemble_steps)
        e = out["ensemble"]
        print(f"# ensemble of {e['domains']} cavities: batched {e['batched_mlups']:.0f} MLUPS, per-domain graph "
              f"{e['per_domain_graph_mlups']:.0f}, per-domain eager {e['per_domain_eager_mlups']:.0f}",
              file=sys.stderr, flush=True)
    if args.sparse_steps > 0:
        out["sparse_layout"] = sparse_leg(torch, vfx, args.sparse_steps, autotune=not args.no_autotune)
    if args.tgv > 0:
        out["taylor_green"] = tgv_leg(torch, vfx, args.tgv, args.tgv_steps)
        torch.cuda.empty_cache()
    if args.channel_steps > 0:
        out["channel"] = channel_leg(torch, vfx, args.channel_steps, separate_steps=args.channel_separate_steps)
        if "fp16" in res and "fp16" in out["channel"]:
            # the channel has the periodic box's dims: the box's step time over the channel's
            c = out["channel"]["fp16"]
            c["bulk_over_channel_step_time"] = round(res["fp16"]["ms_per_step"] / c["ms_per_step"], 4)
            if "pressure_fused_without_statistics" in c:
                # the step alone (the leg's own timing includes a statistics pass every 10 steps)
                c["bulk_over_channel_step_time_without_statistics"] = round(
                    res["fp16"]["ms_per_step"] / c["pressure_fused_without_statistics"]["ms_per_step"], 4)
    if args.cube_steps > 0:
        sch = tuple(x for x in args.cube_schedules.split(",") if x)
        if not sch or set(sch) - {"separate", "merged", "fused_outlet"}:
            sys.exit(f"--cube-schedules: pick from separate, merged, fused_outlet (got {args.cube_schedules!r})")
        out["refined_cube"] = cube_leg(torch, vfx, args.cube_steps, schedules=sch)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
