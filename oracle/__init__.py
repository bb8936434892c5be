"""ctypes access to the CPU checkers (test infrastructure only).

* ``liboracle.so``     -- this project's CPU restatement of the VirtualFluids
  D3Q27 hot path (``oracle.cpp``).
* ``_ref/libvfref.so`` -- the VirtualFluids reference CPU closure compiled from
  the reference sources with ``ref_driver.cpp`` (built only where the
  reference tree exists; the built library may travel without it).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, and only to check or time the reference behaviour.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libvfref.so")
REF_SRC = "/root/reference"

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class OrcBC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("node", C.POINTER(C.c_int64)),
                ("q", _dp), ("ux", _dp), ("uy", _dp), ("uz", _dp),
                ("cpu_float_velocity", C.c_int32), ("rho", _dp), ("nbr", C.POINTER(C.c_int64)),
                ("normal", C.c_int32), ("correction", C.c_double),
                ("wm_normal", _dp * 3), ("wm_el", _dp * 3), ("wm_wall", _dp * 3), ("wm_offset", C.POINTER(C.c_int32)),
                ("wm_z0", _dp), ("wm_ustar", _dp), ("wm_force", _dp * 3),
                ("n_pre", C.c_int32), ("pn", C.POINTER(C.c_int64) * 4), ("pw", _dp * 4), ("plast", _dp),
                ("pcur", _dp), ("tratio", C.c_double), ("vratio", C.c_double),
                ("pseries", _dp), ("n_series", C.c_int32), ("period", C.c_int32), ("t0", C.c_int32)]


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (and the reference closure when its sources exist)."""
    targets = ["all"]
    if ref is None:
        ref = os.path.isdir(REF_SRC)
    if ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, f"-j{min(8, os.cpu_count() or 1)}"] + targets, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_k17_collide.argtypes = [_dp, C.c_double, _dp, _dp, _dp]
        L.orc_bgk_collide.argtypes = [_dp, C.c_double, _dp]
        L.orc_chimera.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_macroscopic.argtypes = [_dp, _dp]
        L.orc_mean_add.argtypes = [_dp, _dp]
        L.orc_equilibrium.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_init_field.argtypes = [_dp, C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_wall_term.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_wall_term.restype = C.c_double
        L.orc_run.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _u8p, C.c_int, C.c_double, _dp, _dp,
                              C.c_int, C.POINTER(OrcBC), C.c_int, _dp]
        L.orc_run.restype = C.c_int
        L.orc_run_turb.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _u8p, C.c_double, _dp, _dp,
                                   C.c_int, C.POINTER(OrcBC), C.c_int, _dp, C.c_int, C.c_double, _dp]
        L.orc_run_turb.restype = C.c_int
        L.orc_run_body_force.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _u8p, C.c_double, _dp, _dp, C.c_double, _dp,
                                         C.c_int, C.POINTER(OrcBC), C.c_int]
        L.orc_run_body_force.restype = C.c_int
        _i32p = C.POINTER(C.c_int32)
        L.orc_pressure_bc.argtypes = [_dp, C.c_double, C.c_int, _dp, C.POINTER(C.c_int32)]
        L.orc_interp_bc.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _dp, _i32p]
        L.orc_outflow_bc.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_double, _dp, _i32p]
        _i64p = C.POINTER(C.c_int64)
        L.orc_precursor_blend.argtypes = [C.c_int, C.c_int, C.c_int] + [_i64p] * 4 + [_dp] * 6 + [C.c_double, _dp]
        L.orc_stress_bc.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp, C.c_int32, C.c_double,
                                    _dp, _dp, _dp, _i32p, _dp, _dp]
        L.orc_macroscopic_quantities.argtypes = [_dp, _dp]
        L.orc_k17_collide_turb.argtypes = [_dp, C.c_double, _dp, _dp, _dp, C.c_int, C.c_double, _dp]
        L.orc_turbulent_viscosity_amd.argtypes = [C.c_int, C.c_int, C.c_int, _u8p, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_mass_momentum.argtypes = [_dp, C.c_int64, _u8p, _dp]
        L.orc_energy.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _u8p, C.c_int, _dp]
        L.orc_source_moments.argtypes = [_dp, C.c_double, _dp]
        L.orc_interp_coefficients.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_interpolate_cf.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_interpolate_fc.argtypes = [_dp, C.c_double, C.c_double, _dp]
        _i64p = C.POINTER(C.c_int64)
        _ip = C.POINTER(C.c_int)
        for fn in (L.orc_scale_cf, L.orc_scale_fc):
            fn.argtypes = [_dp, _dp, _ip, _ip, C.c_int64, _i64p, _i64p, _dp, _dp, _dp, C.c_double, C.c_double]
        for fn in (L.orc_scale_cf_turb, L.orc_scale_fc_turb):
            fn.argtypes = [_dp, _dp, _ip, _ip, C.c_int64, _i64p, _i64p, _dp, _dp, _dp, C.c_double, C.c_double,
                           _dp, _dp]
        L.orc_interp_coefficients_omegas.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.ref_k17_collide.argtypes = [_dp, C.c_double, _dp, _dp, _dp]
        L.ref_macroscopic_quantities.argtypes = [_dp, _dp]
        L.ref_k17_collide_turb.argtypes = [_dp, C.c_double, _dp, _dp, _dp, C.c_int, C.c_double, _dp]
        L.ref_chimera.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ref_macroscopic.argtypes = [_dp, _dp]
        L.ref_run_periodic.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        L.ref_run_cavity.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int]
        L.ref_time_blocks.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ref_run_blocks.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        L.ref_run_periodic_b92.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        L.ref_time_blocks.restype = C.c_double
        L.ref_interp_cell.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      _dp, _dp]
        _ref = L
    return _ref


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


LIMITER = np.array([0.01, 0.01, 0.01])


# ------------------------------------------------------------ single node ---
def k17_collide(f27, omega, lim=LIMITER, force=(0.0, 0.0, 0.0), impl="oracle"):
    """Post-collision populations and (drho, vx, vy, vz) of one node."""
    f = np.ascontiguousarray(f27, dtype=np.float64).copy()
    lim = np.ascontiguousarray(lim, dtype=np.float64)
    fo = np.ascontiguousarray(force, dtype=np.float64)
    mv = np.zeros(4)
    L = lib() if impl == "oracle" else ref()
    fn = L.orc_k17_collide if impl == "oracle" else L.ref_k17_collide
    fn(_p(f), float(omega), _p(lim), _p(fo), _p(mv))
    return f, mv


TURBULENCE = {"none": 0, "smagorinsky": 1, "amd": 2, "qr": 3}


def k17_collide_turb(f27, omega, turb, sgs, nut, lim=LIMITER, force=(0.0, 0.0, 0.0), impl="oracle"):
    """One node with an eddy-viscosity model: (post-collision populations,
    (drho, vx, vy, vz), the node's turbulent viscosity after the collision)."""
    f = np.ascontiguousarray(f27, dtype=np.float64).copy()
    lim = np.ascontiguousarray(lim, dtype=np.float64)
    fo = np.ascontiguousarray(force, dtype=np.float64)
    mv = np.zeros(4)
    nu = np.array([float(nut)])
    L = lib() if impl == "oracle" else ref()
    fn = L.orc_k17_collide_turb if impl == "oracle" else L.ref_k17_collide_turb
    fn(_p(f), float(omega), _p(lim), _p(fo), _p(mv), TURBULENCE[turb], float(sgs), _p(nu))
    return f, mv, float(nu[0])


def macroscopic_quantities(f27, impl="oracle"):
    """(rho, vx, vy, vz, pressure) of one node as the reference's output
    kernel computes them."""
    f = np.ascontiguousarray(f27, dtype=np.float64)
    out = np.zeros(5)
    fn = lib().orc_macroscopic_quantities if impl == "oracle" else ref().ref_macroscopic_quantities
    fn(_p(f), _p(out))
    return out


def mean_add(f27, acc5):
    """calculateMeanCompressible on one fluid node: the running sums acc5 =
    (rho, vx, vy, vz, pressure) after adding populations f27."""
    f = np.ascontiguousarray(f27, dtype=np.float64)
    a = np.array(acc5, dtype=np.float64)
    lib().orc_mean_add(_p(f), _p(a))
    return a


def precursor_blend(neighbors, weights, last, current, time_ratio):
    """Plane values at the boundary nodes, last*(1 - tr) + current*tr of the
    inverse-distance weighted four precursor nodes (first weight >= 1e6: the
    0PP node alone); planes (components, n_pre) -> (components, n)."""
    nb = np.ascontiguousarray(neighbors, dtype=np.int64).reshape(4, -1)
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(4, -1)
    la = np.ascontiguousarray(last, dtype=np.float64)
    cu = np.ascontiguousarray(current, dtype=np.float64)
    comps, n_pre = la.shape
    n = nb.shape[1]
    out = np.zeros((comps, n))
    lib().orc_precursor_blend(n, n_pre, comps, *(nb[j].ctypes.data_as(C.POINTER(C.c_int64)) for j in range(4)),
                              *(_p(w[j]) for j in range(4)), _p(la), _p(cu), float(time_ratio), _p(out))
    return out


def interp_bc(kind, p27, q27, omega, u=(0.0, 0.0, 0.0)):
    """One boundary node of an interpolated kind (1, 2, 3, 5): post-collision
    populations p27, link fractions q27 -> (values, written mask) indexed by the
    direction written (the opposite of the wall link)."""
    p = np.ascontiguousarray(p27, dtype=np.float64)
    q = np.ascontiguousarray(q27, dtype=np.float64)
    out = np.zeros(27)
    wr = np.zeros(27, np.int32)
    lib().orc_interp_bc(int(kind), _p(p), _p(q), float(omega),
                        *(float(x) for x in u), _p(out), wr.ctypes.data_as(C.POINTER(C.c_int32)))
    return out, wr.astype(bool)


def pressure_bc(fn27, rho_bc, normal):
    """One node of the non-equilibrium pressure boundary (kind 4) from its
    inward neighbour's populations -> (values, written mask)."""
    f = np.ascontiguousarray(fn27, dtype=np.float64)
    out = np.zeros(27)
    wr = np.zeros(27, np.int32)
    lib().orc_pressure_bc(_p(f), float(rho_bc), int(normal), _p(out), wr.ctypes.data_as(C.POINTER(C.c_int32)))
    return out, wr.astype(bool)


def stress_bc(kind, p27, q27, omega, sample_velocity, normal, offset, z0, exchange_velocity=(0, 0, 0),
              wall_velocity=(0, 0, 0)):
    """One node of a wall-model boundary (9, 10, 11).  sample_velocity: what
    the kernel reads at the sampling node (its z from the y array).  Returns
    (values, written, new exchange velocity, new wall velocity, u_star, force)."""
    p = np.ascontiguousarray(p27, dtype=np.float64)
    q = np.ascontiguousarray(q27, dtype=np.float64)
    vs = np.ascontiguousarray(sample_velocity, dtype=np.float64)
    nrm = np.ascontiguousarray(normal, dtype=np.float64)
    el = np.array(exchange_velocity, dtype=np.float64)
    wm = np.array(wall_velocity, dtype=np.float64)
    out = np.zeros(27)
    wr = np.zeros(27, np.int32)
    us = np.zeros(1)
    fo = np.zeros(3)
    lib().orc_stress_bc(int(kind), _p(p), _p(q), float(omega), _p(vs), _p(nrm),
                        int(offset), float(z0), _p(el), _p(wm), _p(out), wr.ctypes.data_as(C.POINTER(C.c_int32)),
                        _p(us), _p(fo))
    return out, wr.astype(bool), el, wm, float(us[0]), fo


def outflow_bc(kind, f27, fn27, normal, correction=0.0):
    """One node of the non-reflecting outflow (kind 6, or 7 with pressure
    correction): pre-collision f27 of the node and fn27 of its inward
    neighbour -> (entering populations, written mask)."""
    f = np.ascontiguousarray(f27, dtype=np.float64)
    fn = np.ascontiguousarray(fn27, dtype=np.float64)
    out = np.zeros(27)
    wr = np.zeros(27, np.int32)
    lib().orc_outflow_bc(int(kind), _p(f), _p(fn), int(normal), float(correction), _p(out),
                         wr.ctypes.data_as(C.POINTER(C.c_int32)))
    return out, wr.astype(bool)


def turbulent_viscosity_amd(dims, vx, vy, vz, C_amd, solid=None):
    """AMD eddy viscosity of every non-solid node of a dense periodic box."""
    nx, ny, nz = dims
    v = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (vx, vy, vz)]
    nut = np.zeros(nx * ny * nz)
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
    lib().orc_turbulent_viscosity_amd(nx, ny, nz, sol.ctypes.data_as(_u8p) if sol is not None else None,
                                      _p(v[0]), _p(v[1]), _p(v[2]), float(C_amd), _p(nut))
    return nut


def bgk_collide(f27, omega):
    f = np.ascontiguousarray(f27, dtype=np.float64).copy()
    mv = np.zeros(4)
    lib().orc_bgk_collide(_p(f), float(omega), _p(mv))
    return f, mv


def chimera(kind, abc, vv, v2, kinv=1.0, k=1.0, impl="oracle"):
    """kind: 0 forward-with-K, 1 backward-with-K, 2 forward, 3 backward."""
    a = np.ascontiguousarray(abc, dtype=np.float64).copy()
    L = lib() if impl == "oracle" else ref()
    fn = L.orc_chimera if impl == "oracle" else L.ref_chimera
    fn(int(kind), _p(a), float(vv), float(v2), float(kinv), float(k))
    return a


def macroscopic(f27, impl="oracle"):
    f = np.ascontiguousarray(f27, dtype=np.float64)
    out = np.zeros(4)
    L = lib() if impl == "oracle" else ref()
    fn = L.orc_macroscopic if impl == "oracle" else L.ref_macroscopic
    fn(_p(f), _p(out))
    return out


def equilibrium(drho, ux, uy, uz):
    f = np.zeros(27)
    lib().orc_equilibrium(_p(f), float(drho), float(ux), float(uy), float(uz))
    return f


def init_field(drho, ux, uy, uz):
    """Node-major shifted equilibrium populations for flat field arrays."""
    drho = np.ascontiguousarray(drho, dtype=np.float64).ravel()
    ux = np.ascontiguousarray(ux, dtype=np.float64).ravel()
    uy = np.ascontiguousarray(uy, dtype=np.float64).ravel()
    uz = np.ascontiguousarray(uz, dtype=np.float64).ravel()
    F = np.zeros((drho.size, 27))
    lib().orc_init_field(_p(F), drho.size, _p(drho), _p(ux), _p(uy), _p(uz))
    return F


def wall_term(d, ux, uy, uz, cpu_float=False):
    return lib().orc_wall_term(int(d), float(ux), float(uy), float(uz), int(cpu_float))


# ------------------------------------------------------------- lattices ---
class BCGroup:
    """Boundary group in the VirtualFluids QforBoundaryConditions shape: node
    indices, q[27][n] (link d is a wall link iff 0<=q<=1), per-node wall
    velocity; kind 0 velocity bounce-back, 1 no-slip interpolated, 2 velocity
    interpolated, 3 slip, 4 non-equilibrium pressure (rho, nbr, normal),
    5 velocity interpolated with pressure, 6 non-reflecting outflow (nbr,
    normal), 7 non-reflecting outflow with pressure correction (nbr, normal,
    correction), 8 slip at the node's turbulent-viscosity rate (run_turb),
    9/10/11 wall models (interpolated, bounce-back, bounce-back with pressure:
    nbr = sampling node, wall_normal (3, n), sampling_offset, z0, optional
    initial exchange / wall velocities (3, n); the filtered velocities and the
    monitors u_star, force are updated in place), 12 precursor velocity inlet
    (q links, ux/uy/uz added to the plane velocity), 13 precursor
    distributions (no q): precursor = dict(neighbors (4, n), weights (4, n),
    last, current (components x n_pre), time_ratio, velocity_ratio), or
    instead of last/current/time_ratio a read cycle: series = planes at times
    0, period, 2 period, ..., period, and t0 = the time of the first step
    (each step at time t reads plane(t // period), the next one and
    (t - m period) / period)."""

    def __init__(self, node, q=None, ux=None, uy=None, uz=None, cpu_float_velocity=False, kind=0, rho=None,
                 nbr=None, normal=0, correction=0.0, wall_normal=None, sampling_offset=None, z0=None,
                 exchange_velocity=None, wall_velocity=None, precursor=None):
        self.node = np.ascontiguousarray(node, dtype=np.int64)
        n = self.node.size
        self.q = np.ascontiguousarray(np.full((27, n), -1.0) if q is None else q, dtype=np.float64)
        assert self.q.shape == (27, n)
        z = np.zeros(n)
        self.ux = np.ascontiguousarray(z if ux is None else np.broadcast_to(ux, (n,)), dtype=np.float64)
        self.uy = np.ascontiguousarray(z if uy is None else np.broadcast_to(uy, (n,)), dtype=np.float64)
        self.uz = np.ascontiguousarray(z if uz is None else np.broadcast_to(uz, (n,)), dtype=np.float64)
        self.rho = np.ascontiguousarray(z if rho is None else np.broadcast_to(rho, (n,)), dtype=np.float64)
        self.nbr = np.ascontiguousarray(self.node if nbr is None else nbr, dtype=np.int64)
        self.normal = int(normal)
        self.correction = float(correction)
        self.stress = 9 <= kind <= 11
        self.pre = None
        if kind in (12, 13):
            pc = precursor
            series = pc.get("series")
            if series is not None:
                series = np.ascontiguousarray(np.stack(series), dtype=np.float64)
            self.pre = {"nb": np.ascontiguousarray(pc["neighbors"], dtype=np.int64).reshape(4, n),
                        "w": np.ascontiguousarray(pc["weights"], dtype=np.float64).reshape(4, n),
                        "last": np.ascontiguousarray(pc["last"] if series is None else series[0], dtype=np.float64),
                        "cur": np.ascontiguousarray(pc["current"] if series is None else series[1], dtype=np.float64),
                        "tr": float(pc.get("time_ratio", 0.0)), "vr": float(pc.get("velocity_ratio", 1.0)),
                        "series": series, "period": int(pc.get("period", 1)), "t0": int(pc.get("t0", 1))}
            self.pre["n_pre"] = self.pre["last"].shape[-1]
        if self.stress:
            self.wall_normal = np.ascontiguousarray(np.broadcast_to(np.asarray(wall_normal, np.float64).reshape(3, -1),
                                                                    (3, n)), dtype=np.float64).copy()
            self.sampling_offset = np.ascontiguousarray(np.broadcast_to(sampling_offset, (n,)), dtype=np.int32)
            self.z0 = np.ascontiguousarray(np.broadcast_to(z0, (n,)), dtype=np.float64)
            self.exchange_velocity = np.zeros((3, n)) if exchange_velocity is None else \
                np.array(np.broadcast_to(np.asarray(exchange_velocity, np.float64).reshape(3, -1), (3, n)))
            self.wall_velocity = np.zeros((3, n)) if wall_velocity is None else \
                np.array(np.broadcast_to(np.asarray(wall_velocity, np.float64).reshape(3, -1), (3, n)))
            self.u_star = np.zeros(n)
            self.force = np.zeros((3, n))
        self.cpu_float_velocity = bool(cpu_float_velocity)
        self.kind = kind

    def c(self):
        return OrcBC(self.kind, self.node.size, self.node.ctypes.data_as(C.POINTER(C.c_int64)),
                     _p(self.q), _p(self.ux), _p(self.uy), _p(self.uz),
                     int(self.cpu_float_velocity), _p(self.rho), self.nbr.ctypes.data_as(C.POINTER(C.c_int64)),
                     self.normal, self.correction, *self._wm(), *self._pre())

    def _pre(self):
        if self.pre is None:
            return 0, (C.POINTER(C.c_int64) * 4)(), (_dp * 4)(), None, None, 0.0, 1.0, None, 0, 1, 0
        p = self.pre
        ser = p["series"]
        return (p["n_pre"], (C.POINTER(C.c_int64) * 4)(*(p["nb"][j].ctypes.data_as(C.POINTER(C.c_int64)) for j in range(4))),
                (_dp * 4)(*(_p(p["w"][j]) for j in range(4))), _p(p["last"]), _p(p["cur"]), p["tr"], p["vr"],
                None if ser is None else _p(ser), 0 if ser is None else ser.shape[0], p["period"], p["t0"])

    def _wm(self):
        z3 = (_dp * 3)(None, None, None)
        if not self.stress:
            return z3, z3, z3, None, None, None, z3
        rows = lambda a: (_dp * 3)(*(_p(a[j]) for j in range(3)))  # noqa: E731
        return (rows(self.wall_normal), rows(self.exchange_velocity), rows(self.wall_velocity),
                self.sampling_offset.ctypes.data_as(C.POINTER(C.c_int32)), _p(self.z0), _p(self.u_star),
                rows(self.force))


def run(F, dims, nsteps, omega, collision="k17", solid=None, bcs=(), lim=LIMITER, force=(0.0, 0.0, 0.0),
        want_macro=False):
    """Advance node-major populations F (N x 27, x fastest) by nsteps."""
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    assert F.shape == (nx * ny * nz, 27)
    sol = None
    if solid is not None:
        sol = np.ascontiguousarray(solid, dtype=np.uint8).ravel()
        assert sol.size == nx * ny * nz
    bcl = list(bcs)
    arr = (OrcBC * max(1, len(bcl)))(*[b.c() for b in bcl]) if bcl else None
    macro = np.zeros((nx * ny * nz, 4)) if want_macro else None
    lim = np.ascontiguousarray(lim, dtype=np.float64)
    fo = np.ascontiguousarray(force, dtype=np.float64)
    rc = lib().orc_run(_p(F), nx, ny, nz, sol.ctypes.data_as(_u8p) if sol is not None else None,
                       0 if collision == "k17" else 1, float(omega), _p(lim), _p(fo), len(bcl), arr,
                       int(nsteps), _p(macro) if macro is not None else None)
    if rc == -2:
        raise RuntimeError("orc_run failed: a precursor plane series ends before the last step")
    if rc != 0:
        raise RuntimeError(f"orc_run failed ({rc}): a fluid->solid link has no boundary condition")
    return (F, macro) if want_macro else F


def run_turb(F, dims, nsteps, omega, turb, sgs, nut=None, solid=None, bcs=(), lim=LIMITER,
             force=(0.0, 0.0, 0.0), want_macro=False):
    """`run` with the K17 collision and an eddy-viscosity model; returns
    (F, nut[, macro]) with the turbulent viscosity of every node."""
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    assert F.shape == (nx * ny * nz, 27)
    nut = np.zeros(nx * ny * nz) if nut is None else np.ascontiguousarray(nut, dtype=np.float64).copy()
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
    bcl = list(bcs)
    arr = (OrcBC * max(1, len(bcl)))(*[b.c() for b in bcl]) if bcl else None
    macro = np.zeros((nx * ny * nz, 4)) if want_macro else None
    lim = np.ascontiguousarray(lim, dtype=np.float64)
    fo = np.ascontiguousarray(force, dtype=np.float64)
    rc = lib().orc_run_turb(_p(F), nx, ny, nz, sol.ctypes.data_as(_u8p) if sol is not None else None, float(omega),
                            _p(lim), _p(fo), len(bcl), arr, int(nsteps), _p(macro) if macro is not None else None,
                            TURBULENCE[turb], float(sgs), _p(nut))
    if rc != 0:
        raise RuntimeError(f"orc_run_turb failed ({rc}): a fluid->solid link has no boundary condition")
    return (F, nut, macro) if want_macro else (F, nut)


def run_body_force(F, dims, nsteps, omega, forces, force_factor, body_force, solid=None, bcs=(), lim=LIMITER):
    """`run` with the ApplyBodyForce template: returns (F, body_force after
    the run -- zero where a node collided)."""
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    bf = np.ascontiguousarray(body_force, dtype=np.float64).reshape(3, -1).copy()
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
    bcl = list(bcs)
    arr = (OrcBC * max(1, len(bcl)))(*[b.c() for b in bcl]) if bcl else None
    lim = np.ascontiguousarray(lim, dtype=np.float64)
    g = np.ascontiguousarray(forces, dtype=np.float64)
    rc = lib().orc_run_body_force(_p(F), nx, ny, nz, sol.ctypes.data_as(_u8p) if sol is not None else None, float(omega),
                                  _p(lim), _p(g), float(force_factor), _p(bf), len(bcl), arr, int(nsteps))
    if rc != 0:
        raise RuntimeError("orc_run_body_force failed")
    return F, bf


def mass_momentum(F, solid=None):
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.zeros(4)
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
    lib().orc_mass_momentum(_p(F), F.shape[0], sol.ctypes.data_as(_u8p) if sol is not None else None, _p(out))
    return out


def energy(F, dims, enstrophy=False, solid=None):
    """(sum, count) of the kinetic-energy or enstrophy analyzer."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.zeros(2)
    sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
    lib().orc_energy(_p(F), dims[0], dims[1], dims[2], sol.ctypes.data_as(_u8p) if sol is not None else None,
                     int(bool(enstrophy)), _p(out))
    return out


def ref_run_periodic(F, dims, nsteps, omega):
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    ref().ref_run_periodic(_p(F), nx, ny, nz, float(omega), int(nsteps))
    return F


def ref_run_periodic_b92(F, dims, nsteps, omega):
    """The reference's incompressible BGK CPU kernel (B92IncompressibleNavierStokes)
    on a periodic block."""
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    ref().ref_run_periodic_b92(_p(F), nx, ny, nz, float(omega), int(nsteps))
    return F


def ref_run_cavity(F, dims, nsteps, omega, lid=(0.05, 0.0, 0.0)):
    nx, ny, nz = dims
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    ref().ref_run_cavity(_p(F), nx, ny, nz, float(omega), float(lid[0]), float(lid[1]), float(lid[2]),
                         int(nsteps))
    return F


def ref_time_blocks(nblocks, bs, omega, nsteps, exchange=True):
    """Wall seconds of nsteps reference K17 steps on nblocks bs^3 blocks with
    (or without) the block exchange; returns (seconds, OpenMP threads, block grid)."""
    t = C.c_int(0)
    g = (C.c_int * 3)()
    secs = ref().ref_time_blocks(int(nblocks), int(bs), float(omega), int(nsteps), int(bool(exchange)), C.byref(t), g)
    return secs, t.value, tuple(g)


def ref_run_blocks(F, grid, bs, nsteps, omega):
    """The CPU baseline's block grid (grid = blocks per axis, bs^3 nodes each,
    periodic) on one node-major field of grid*bs nodes per axis."""
    F = np.ascontiguousarray(F, dtype=np.float64).copy()
    ref().ref_run_blocks(_p(F), int(grid[0]), int(grid[1]), int(grid[2]), int(bs), float(omega), int(nsteps))
    return F


# ---------------------------------------------------------------- refinement
def source_moments(f27, omega):
    """(drho, vx, vy, vz, kxy, kyz, kxz, kxxMyy, kxxMzz) of one source node."""
    f = np.ascontiguousarray(f27, dtype=np.float64)
    out = np.zeros(9)
    lib().orc_source_moments(_p(f), float(omega), _p(out))
    return out


def interp_coefficients(f8x27, omega, offset=(0.0, 0.0, 0.0)):
    """42 interpolation coefficients of one cell (8 corners, bit 0 = +x)."""
    f = np.ascontiguousarray(f8x27, dtype=np.float64).reshape(8, 27)
    out = np.zeros(42)
    lib().orc_interp_coefficients(_p(f), float(omega), *(float(o) for o in offset), _p(out))
    return out


def interpolate_cf(coeff42, omega_fine, x, y, z, eps=0.5):
    k = np.ascontiguousarray(coeff42, dtype=np.float64)
    out = np.zeros(27)
    lib().orc_interpolate_cf(_p(k), float(omega_fine), float(eps), float(x), float(y), float(z), _p(out))
    return out


def interpolate_fc(coeff42, omega_coarse, eps=2.0):
    k = np.ascontiguousarray(coeff42, dtype=np.float64)
    out = np.zeros(27)
    lib().orc_interpolate_fc(_p(k), float(omega_coarse), float(eps), _p(out))
    return out


def interp_cell(kind, f8x27, omega_src, omega_dst, offset=(0.0, 0.0, 0.0), impl="oracle"):
    """One interface cell end to end: kind "cf" -> (coeff42, 8 x 27 fine nodes),
    kind "fc" -> (coeff42, 27 coarse populations)."""
    f = np.ascontiguousarray(f8x27, dtype=np.float64).reshape(8, 27)
    if impl == "ref":
        k = np.zeros(42)
        out = np.zeros((8, 27) if kind == "cf" else 27)
        ref().ref_interp_cell(0 if kind == "cf" else 1, _p(f), float(omega_src), float(omega_dst),
                              *(float(o) for o in offset), _p(k), _p(out))
        return k, out
    k = interp_coefficients(f, omega_src, offset)
    if kind == "cf":
        out = np.stack([interpolate_cf(k, omega_dst, (0.25 if c & 1 else -0.25), (0.25 if c & 2 else -0.25),
                                       (0.25 if c & 4 else -0.25)) for c in range(8)])
    else:
        out = interpolate_fc(k, omega_dst)
    return k, out


def interp_coefficients_omegas(f8x27, omega8, offset=(0.0, 0.0, 0.0)):
    """interp_coefficients with one omega per corner (corner bit order)."""
    f = np.ascontiguousarray(f8x27, dtype=np.float64)
    om = np.ascontiguousarray(omega8, dtype=np.float64)
    out = np.zeros(42)
    lib().orc_interp_coefficients_omegas(_p(f), _p(om), *(float(o) for o in offset), _p(out))
    return out


def _scale(fn, Fc, Ff, dims_c, dims_f, coarse, fine, offsets, omega_c, omega_f, nut=None):
    n = len(coarse)
    co = np.ascontiguousarray(coarse, dtype=np.int64)
    fi = np.ascontiguousarray(fine, dtype=np.int64)
    dc = (C.c_int * 3)(*dims_c)
    df = (C.c_int * 3)(*dims_f)
    offs = [None, None, None]
    if offsets is not None:
        offs = [np.ascontiguousarray(o, dtype=np.float64) for o in offsets]
    i64 = C.POINTER(C.c_int64)
    extra = ()
    if nut is not None:
        nc, nf = (np.ascontiguousarray(a, dtype=np.float64) for a in nut)
        extra = (_p(nc), _p(nf))
    fn(_p(Fc), _p(Ff), dc, df, n, co.ctypes.data_as(i64), fi.ctypes.data_as(i64),
       *(_p(o) if o is not None else None for o in offs), float(omega_c), float(omega_f), *extra)


def scale_cf(Fc, Ff, dims_c, dims_f, coarse_cells, fine_cells, omega_c, omega_f, offsets=None, nut=None):
    """Coarse-to-fine over cells; returns the updated fine populations.
    nut = (coarse, fine) turbulent viscosity arrays: per-node omegas."""
    Fc = np.ascontiguousarray(Fc, dtype=np.float64)
    Ff = np.ascontiguousarray(Ff, dtype=np.float64).copy()
    fn = lib().orc_scale_cf if nut is None else lib().orc_scale_cf_turb
    _scale(fn, Fc, Ff, dims_c, dims_f, coarse_cells, fine_cells, offsets, omega_c, omega_f, nut)
    return Ff


def scale_fc(Fc, Ff, dims_c, dims_f, coarse_nodes, fine_cells, omega_c, omega_f, offsets=None, nut=None):
    """Fine-to-coarse over cells; returns the updated coarse populations.
    nut = (coarse, fine) turbulent viscosity arrays: per-node omegas."""
    Fc = np.ascontiguousarray(Fc, dtype=np.float64).copy()
    Ff = np.ascontiguousarray(Ff, dtype=np.float64)
    fn = lib().orc_scale_fc if nut is None else lib().orc_scale_fc_turb
    _scale(fn, Fc, Ff, dims_c, dims_f, coarse_nodes, fine_cells, offsets, omega_c, omega_f, nut)
    return Fc
