"""Precursor inflows in the oracle (PrecursorNonReflectiveCompressible.cu,
PrecursorDistributions.cu) and the plane time cycling of PrecursorPlanes
(BoundaryConditionKernelManager.cpp:346-382), without a GPU.

Pins: the plane stencil reproduces a constant plane and averages four equal
weights, takes the 0PP node alone from a first weight of 1e6 on, and blends
linearly in time with the end points last / current; the velocity inlet equals
VelocityWithPressureInterpolatedCompressible (kind 5, pinned against the
compiled reference) at the blended velocity on every link but d0PM / d0MP, and
there with f_0PP in place of f_0MP; in a run, the distributions kind sets the
nine positive-x populations of the next state to the plane values and the
velocity inlet's value enters two steps after it was computed (the first two
states read the initial populations).  The cycling matches the closed form of
the reference's schedule: at step t, last = plane(mT), current =
plane((m+1)T), ratio (t - mT)/T with m = floor(t/T).
"""
import numpy as np
import pytest
import torch

import oracle
from tests import cases
from virtualfluids_amd.lattice import BoundaryCondition, PrecursorPlanes

W = cases.W
PDIRS = [1, 7, 9, 11, 13, 19, 21, 23, 25]  # dP00 dPP0 dPM0 dP0P dP0M dPPP dPMP dPPM dPMM


def _stencil(n, n_pre, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n_pre, (4, n)), rng.uniform(0.1, 2.0, (4, n))


def test_plane_stencil_closed_forms():
    nb, w = _stencil(7, 11, 1)
    const = np.full((3, 11), 0.0375)
    assert np.allclose(oracle.precursor_blend(nb, w, const, const, 0.3), 0.0375, rtol=1e-15, atol=0)
    rng = np.random.default_rng(2)
    last, cur = rng.normal(size=(9, 11)), rng.normal(size=(9, 11))
    ones = np.ones((4, 7))
    mean = oracle.precursor_blend(nb, ones, last, last, 0.0)
    assert np.allclose(mean, last[:, nb].mean(axis=1), rtol=1e-14, atol=1e-15)
    w1 = w.copy()
    w1[0, :3] = 1e6
    lone = oracle.precursor_blend(nb, w1, last, cur, 0.0)
    assert np.array_equal(lone[:, :3], last[:, nb[0, :3]])
    a0 = oracle.precursor_blend(nb, w, last, cur, 0.0)
    a1 = oracle.precursor_blend(nb, w, last, cur, 1.0)
    assert np.allclose(oracle.precursor_blend(nb, w, last, last, 0.0), oracle.precursor_blend(nb, w, cur, last, 1.0))
    for tr in (0.25, 0.5, 0.8):
        assert np.allclose(oracle.precursor_blend(nb, w, last, cur, tr), (1 - tr) * a0 + tr * a1, rtol=1e-13, atol=1e-14)


def test_velocity_inlet_is_kind5_but_for_two_links():
    rng = np.random.default_rng(3)
    p = oracle.equilibrium(0.01, 0.03, -0.01, 0.02) + rng.normal(0, 1e-3, 27) * W
    u = (0.031, -0.004, 0.012)
    q = np.where(rng.uniform(size=27) < 0.6, rng.uniform(0, 1, 27), -1.0).astype(np.float32)
    q[[17, 18]] = -1.0
    a, wa = oracle.interp_bc(12, p, q, 1.7, u)
    b, wb = oracle.interp_bc(5, p, q, 1.7, u)
    assert np.array_equal(wa, wb) and wa.sum() > 5 and np.array_equal(a[wa], b[wb])
    # d0PM / d0MP: the value is affine in (f, fi) at fixed moments, so
    # substituting f_0PP for f_0MP shifts it by the coefficient of the slot
    # it enters times (f_0PP - f_0MP)
    om = 1.7
    for d, slot in ((17, "fi"), (18, "f")):  # d0PM, d0MP
        qd = np.full(27, -1.0, np.float32)
        qd[d] = 0.4
        e = cases.INV[d]
        a = oracle.interp_bc(12, p, qd, om, u)[0][e]
        b = oracle.interp_bc(5, p, qd, om, u)[0][e]
        qq = 0.4
        sgn = -1.0 if slot == "fi" else 1.0
        coef = (1 - qq) / (1 + qq) * (sgn + 1 / (1 - om)) / 2 + qq / (1 + qq)
        assert np.isclose(a - b, coef * (p[15] - p[18]), rtol=1e-6, atol=1e-15) and a != b
        p_eq = p.copy()
        p_eq[18] = p_eq[15]
        assert oracle.interp_bc(12, p_eq, qd, om, u)[0][e] == oracle.interp_bc(5, p_eq, qd, om, u)[0][e]


def _channel(dims=(9, 6, 6)):
    solid, groups, onodes, onbrs = cases.inflow_channel(*dims)
    inlet, iq = groups["inlet"]
    F0 = oracle.init_field(*cases.random_flow(dims, seed=5, u_amp=0.01))
    outlet = oracle.BCGroup(onodes, kind=6, nbr=onbrs, normal=1)
    return solid, groups, inlet, iq, F0, outlet


def test_distributions_set_the_next_state():
    dims = (9, 6, 6)
    solid, groups, inlet, iq, F0, outlet = _channel(dims)
    n, n_pre = inlet.size, 30
    nb, w = _stencil(n, n_pre, 7)
    rng = np.random.default_rng(8)
    last, cur = rng.normal(0, 1e-3, (9, n_pre)), rng.normal(0, 1e-3, (9, n_pre))
    pre = dict(neighbors=nb, weights=w, last=last, current=cur, time_ratio=0.35)
    bcs = [outlet, oracle.BCGroup(*groups["wall"], kind=1), oracle.BCGroup(inlet, kind=13, precursor=pre)]
    F1 = oracle.run(F0, dims, 1, 1.8, solid=solid, bcs=bcs)
    want = oracle.precursor_blend(nb, w, last, cur, 0.35)
    assert np.array_equal(F1[inlet][:, PDIRS], want.T)


def test_velocity_inlet_two_step_lag():
    dims = (9, 6, 6)
    solid, groups, inlet, iq, F0, outlet = _channel(dims)
    n, n_pre = inlet.size, 30
    nb, w = _stencil(n, n_pre, 9)
    rng = np.random.default_rng(10)
    pre = dict(neighbors=nb, weights=w, last=rng.normal(0, 1e-3, (3, n_pre)), current=rng.normal(0, 1e-3, (3, n_pre)),
               time_ratio=0.6, velocity_ratio=1.0)
    ux = 0.02 + rng.uniform(0, 1e-3, n)

    def bcs():
        return [outlet, oracle.BCGroup(*groups["wall"], kind=1), oracle.BCGroup(inlet, iq, ux, kind=12, precursor=pre)]

    nx, ny, _ = dims
    links = [(i, d) for i in range(n) for d in range(1, 27) if 0 <= iq[d, i] <= 1]
    F1 = oracle.run(F0, dims, 1, 1.8, solid=solid, bcs=bcs())
    for i, d in links:  # the first state after the start reads the initial element
        k = inlet[i]
        ghost = k + cases.CX[d] + nx * (cases.CY[d] + ny * cases.CZ[d])
        assert F1[k, cases.INV[d]] == F0[ghost, d]
    F2 = oracle.run(F0, dims, 2, 1.8, solid=solid, bcs=bcs())
    F3 = oracle.run(F0, dims, 3, 1.8, solid=solid, bcs=bcs())
    vel = oracle.precursor_blend(nb, w, pre["last"], pre["current"], 0.6)
    for i in range(n):  # state 3 holds what the inlet computed from the collision of state 1
        k = inlet[i]
        post = oracle.k17_collide(F1[k], 1.8)[0]
        out, wr = oracle.interp_bc(12, post, iq[:, i], 1.8, (ux[i] + vel[0, i], vel[1, i], vel[2, i]))
        assert np.allclose(F3[k][wr], out[wr], rtol=1e-12, atol=1e-15)
    assert not np.array_equal(F2[inlet], F3[inlet])


def test_plane_cycling_closed_form():
    T = 3
    loads = []

    def reader(t):
        loads.append(t)
        return np.full((9, 4), t)

    nb = np.zeros((4, 2), np.uint32)
    pc = PrecursorPlanes(nb, np.ones((4, 2)), 4, 9, reader=reader, time_steps_between_reads=T, physical_time_ratio=0.5,
                         device="cpu")
    for t in range(1, 14):
        tr = pc.advance(t)
        m = t // T
        assert tr == (t - m * T) / T
        assert float(pc.last[0, 0]) == m * T * 0.5 and float(pc.current[0, 0]) == (m + 1) * T * 0.5
    assert loads == [k * T * 0.5 for k in range(len(loads))]
    b = BoundaryCondition(13, np.arange(2), precursor=pc, device="cpu")
    q = b.tables(torch.float32)["q"]
    assert q.shape == (27, 2) and q.dtype == torch.float32 and not ((q >= 0) & (q <= 1)).any()
    # the planes rotate through three arrays allocated once: a new read lands in
    # the array that held `last`
    ptrs = {pc.last.data_ptr(), pc.current.data_ptr(), pc.next.data_ptr()}
    pc.advance(14)
    pc.advance(15)
    assert {pc.last.data_ptr(), pc.current.data_ptr(), pc.next.data_ptr()} == ptrs
    try:
        BoundaryCondition(12, np.arange(2), np.zeros((27, 2)), precursor=pc, device="cpu")
    except ValueError:
        pass
    else:
        raise AssertionError("a velocity inlet must refuse nine-component planes")
    assert torch.equal(pc.neighbors, torch.zeros((4, 2), dtype=torch.int32))


def test_plane_series_is_the_stepwise_closed_form_cycle():
    """The oracle's plane series (one run through the read cycle) equals
    stepping one state at a time with last = plane(mT), current = plane((m+1)T)
    and ratio (t - mT)/T for the distributions kind (no state carried between
    steps).  For the velocity kind one run is needed: its two-step lag carries
    each step's value into the state after next."""
    dims = (10, 6, 6)
    solid, groups, onodes, onbrs = cases.inflow_channel(*dims)
    inlet, iq = groups["inlet"]
    n = inlet.size
    rng = np.random.default_rng(2)
    nb = rng.integers(0, 12, (4, n))
    w = rng.uniform(0.1, 2.0, (4, n))
    base = oracle.init_field(np.zeros(1), np.full(1, 0.03), np.zeros(1), np.zeros(1))[0][PDIRS]
    T, steps = 2, 7
    series = [base[:, None] + 1e-4 * rng.normal(size=(9, 12)) for _ in range(steps // T + 2)]
    F0 = oracle.init_field(*cases.random_flow(dims, seed=1, u_amp=0.01))
    F0[solid] = 0.0
    walls = [oracle.BCGroup(onodes, kind=6, nbr=onbrs, normal=1), oracle.BCGroup(*groups["wall"], kind=1)]
    pre = dict(neighbors=nb, weights=w, series=series, period=T, t0=1)
    one = oracle.run(F0, dims, steps, 1.9, solid=solid, bcs=walls + [oracle.BCGroup(inlet, kind=13, precursor=pre)])
    F = F0
    for t in range(1, steps + 1):
        m = t // T
        p = dict(neighbors=nb, weights=w, last=series[m], current=series[m + 1], time_ratio=(t - m * T) / T)
        F = oracle.run(F, dims, 1, 1.9, solid=solid, bcs=walls + [oracle.BCGroup(inlet, kind=13, precursor=p)])
    assert np.array_equal(one, F)
    # a series too short for the steps is refused
    with pytest.raises(RuntimeError):
        oracle.run(F0, dims, steps + 4, 1.9, solid=solid, bcs=walls + [oracle.BCGroup(inlet, kind=13, precursor=pre)])
