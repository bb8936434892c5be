"""The oracle's calculateMeanCompressible (oracle.mean_add), without a GPU:

* from zero sums it gives the node's density, velocity and pressure, as the
  output kernel's macroscopic quantities (pinned by the compiled reference
  functions) -- equal up to the different density summation order;
* density and velocity sums are sums of the per-call values;
* the pressure sum uses the new density and velocity sums in place of the
  node's own values, as the reference kernel reads rhoD/vxD back: for two
  identical calls on a fluid at rest P2 = 2 P1 + drho / 2, and on a moving
  fluid of zero density deviation P2 = 2 P1 - 3/2 |u|^2; a restatement that
  used the node's own values would give P2 = 2 P1 in both cases.
"""
import numpy as np

import oracle


def _state(drho, u, noise, seed):
    rng = np.random.default_rng(seed)
    return oracle.equilibrium(drho, *u) + rng.normal(0.0, noise, 27)


def test_first_call_gives_the_macroscopic_quantities():
    for seed in range(20):
        f = _state(0.02 * np.sin(seed), (0.03, -0.02 * seed / 20, 0.01), 1e-4, seed)
        got = oracle.mean_add(f, np.zeros(5))
        want = oracle.macroscopic_quantities(f)
        assert np.allclose(got, want, rtol=1e-14, atol=1e-17)


def test_density_and_velocity_sums_add_up():
    f1 = _state(0.01, (0.02, 0.01, -0.03), 1e-4, 1)
    f2 = _state(-0.02, (-0.01, 0.04, 0.0), 1e-4, 2)
    a = oracle.mean_add(f2, oracle.mean_add(f1, np.zeros(5)))
    m1, m2 = oracle.macroscopic_quantities(f1), oracle.macroscopic_quantities(f2)
    assert np.allclose(a[:4], m1[:4] + m2[:4], rtol=1e-14, atol=1e-17)


def test_pressure_sum_reads_the_new_sums_back():
    d = 0.013
    rest = oracle.equilibrium(d, 0.0, 0.0, 0.0)
    p1 = oracle.mean_add(rest, np.zeros(5))[4]
    p2 = oracle.mean_add(rest, oracle.mean_add(rest, np.zeros(5)))[4]
    assert abs(p2 - (2 * p1 + d / 2)) < 1e-15
    assert abs(p2 - 2 * p1) > 1e-3
    u = np.array([0.04, -0.03, 0.02])
    moving = oracle.equilibrium(0.0, *u)
    q1 = oracle.mean_add(moving, np.zeros(5))
    q2 = oracle.mean_add(moving, q1)
    uu = float(q1[1:4] @ q1[1:4])  # |u| of the first call
    assert abs(q2[4] - (2 * q1[4] - 1.5 * uu)) < 1e-15
    assert abs(q2[4] - 2 * q1[4]) > 1e-4

