"""Writes the golden vectors of the K17 hot path from the VirtualFluids reference
CPU closure (oracle/_ref/libvfref.so, built by `make -C oracle ref` from the
reference sources).  Run once in a container that has the reference tree:

    python tests/golden/make_golden.py

Fixtures (inputs and reference outputs only; parameters in golden.json):
  g1_k17_nodes.npz  single-node vf::lbm::runK17CompressibleNavierStokes<None>
                    (K17CompressibleNavierStokes.h:87-591) on random shifted states
  g2_periodic.npz   8^3 periodic block, K17CompressibleNavierStokes::calculate +
                    EsoSplit (K17CompressibleNavierStokes.cpp:77-162), 10 and 100 steps
  g3_cavity.npz     10^3 lid-driven cavity, VelocityBounceBack strategy
                    (VelocityBounceBack.cpp:59-79) on every wall link, 50 steps
  g4_interp.npz     refinement interpolation of 128 random 8-node cells with the
                    header-only MomentsOnSourceNodeSet / interpolateCF / interpolateFC
                    (InterpolationCoefficients.h:74-432, InterpolationCF.h:57-306,
                    InterpolationFC.h:56-263)
  g6_k17_turbulence.npz  single-node runK17CompressibleNavierStokes<Smagorinsky / AMD / QR>
                    (K17CompressibleNavierStokes.h:245, 355-368, TurbulentViscosity.h:63-89)
                    with random turbulent viscosity in; populations, macroscopic
                    values and turbulent viscosity out

`python tests/golden/make_golden.py g4` writes only the named fixtures.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests import cases  # noqa: E402

W = cases.W


def g1(seed=20240601, n=384):
    rng = np.random.default_rng(seed)
    omegas = np.array([0.9, 1.2, 1.6778523, 1.9417, 1.998])
    f_in = np.zeros((n, 27))
    om = np.zeros(n)
    force = np.zeros((n, 3))
    lim = np.zeros((n, 3))
    for i in range(n):
        scale = [1e-4, 1e-2, 1e-1][i % 3]
        f_in[i] = rng.uniform(-1, 1, 27) * scale * W
        om[i] = omegas[i % len(omegas)]
        lim[i] = (0.01, 0.01, 0.01) if i % 4 else rng.uniform(1e-3, 1.0, 3)
        if i % 7 == 0:
            force[i] = rng.uniform(-1e-5, 1e-5, 3)
    f_out = np.zeros_like(f_in)
    macro = np.zeros((n, 4))
    for i in range(n):
        f_out[i], macro[i] = oracle.k17_collide(f_in[i], om[i], lim=lim[i], force=force[i], impl="ref")
    np.savez(os.path.join(HERE, "g1_k17_nodes.npz"), f_in=f_in, omega=om, limiter=lim, force=force,
             f_out=f_out, macro=macro)
    return {"seed": seed, "n": n, "omegas": omegas.tolist(), "note": "limiter 0.01 except every 4th node"}


def g2(seed=8, n=8, omega=1.6):
    dims = (n, n, n)
    fields = cases.random_flow(dims, seed=seed)
    F0 = oracle.init_field(*fields)
    F10 = oracle.ref_run_periodic(F0, dims, 10, omega)
    F100 = oracle.ref_run_periodic(F0, dims, 100, omega)
    np.savez(os.path.join(HERE, "g2_periodic.npz"), F0=F0, F10=F10, F100=F100)
    return {"seed": seed, "dims": list(dims), "omega": omega, "steps": [10, 100]}


def g3(n=10, omega=1.6778523, lid=(0.05, 0.0, 0.0), steps=50):
    F0 = np.zeros((n ** 3, 27))
    F = oracle.ref_run_cavity(F0, (n, n, n), steps, omega, lid=lid)
    np.savez(os.path.join(HERE, "g3_cavity.npz"), F=F)
    return {"n_fluid": n, "omega": omega, "lid": list(lid), "steps": steps,
            "layout": "fluid nodes only, x fastest; walls are the surrounding layer"}


def g4(seed=4242, n=128):
    """Cells of near-equilibrium corner states with random drho, velocity and
    non-equilibrium parts; every other cell carries a random offset (the
    extrapolation used next to walls)."""
    rng = np.random.default_rng(seed)
    f8 = np.zeros((n, 8, 27))
    offsets = np.zeros((n, 3))
    omega = np.zeros((n, 2))  # (coarse, fine)
    for i in range(n):
        u0 = rng.uniform(-0.05, 0.05, 3)
        grad = rng.uniform(-0.02, 0.02, (3, 3))
        for c in range(8):
            s = np.array([(1 if c & 1 else -1), (1 if c & 2 else -1), (1 if c & 4 else -1)]) * 0.5
            u = u0 + grad @ s
            f8[i, c] = oracle.equilibrium(rng.uniform(-0.01, 0.01), *u) + rng.normal(0, 1e-3, 27) * W
        if i % 2:
            offsets[i] = rng.uniform(-0.5, 0.5, 3)
        wc = rng.uniform(1.6, 1.99)
        omega[i] = (wc, 1.0 / (2.0 / wc - 0.5))  # fine omega for nu_f = 2 nu_c
    coeff_cf = np.zeros((n, 42))
    coeff_fc = np.zeros((n, 42))
    cf = np.zeros((n, 8, 27))
    fc = np.zeros((n, 27))
    for i in range(n):
        coeff_cf[i], cf[i] = oracle.interp_cell("cf", f8[i], omega[i, 0], omega[i, 1], offsets[i], impl="ref")
        coeff_fc[i], fc[i] = oracle.interp_cell("fc", f8[i], omega[i, 1], omega[i, 0], offsets[i], impl="ref")
    np.savez(os.path.join(HERE, "g4_interp.npz"), f8=f8, offsets=offsets, omega=omega, coeff_cf=coeff_cf,
             coeff_fc=coeff_fc, cf=cf, fc=fc)
    return {"seed": seed, "n": n, "omega": "coarse U[1.6,1.99], fine 1/(2/omega_c - 1/2)",
            "cf": "coarse moments (omega_c) -> 8 fine nodes at +-1/4, eps 1/2, omega_f",
            "fc": "fine moments (omega_f) -> cell centre, eps 2, omega_c"}


def g6(seed=6606, n=192):
    """Per model (smagorinsky, amd, qr): random shifted states at three
    amplitudes, omega in [0.6, 1.999], turbulent viscosity in [0, 0.05) (0 for
    every 8th node), SGS constant in [0.05, 0.5]."""
    rng = np.random.default_rng(seed)
    out = {}
    for model in ("smagorinsky", "amd", "qr"):
        f_in = np.zeros((n, 27))
        om = np.zeros(n)
        nut_in = np.zeros(n)
        sgs = np.zeros(n)
        for i in range(n):
            f_in[i] = rng.uniform(-1, 1, 27) * [1e-4, 1e-2, 1e-1][i % 3] * W
            om[i] = rng.uniform(0.6, 1.999)
            nut_in[i] = 0.0 if i % 8 == 0 else rng.uniform(0.0, 0.05)
            sgs[i] = rng.uniform(0.05, 0.5)
        f_out = np.zeros_like(f_in)
        macro = np.zeros((n, 4))
        nut_out = np.zeros(n)
        for i in range(n):
            f_out[i], macro[i], nut_out[i] = oracle.k17_collide_turb(f_in[i], om[i], model, sgs[i], nut_in[i],
                                                                     impl="ref")
        out.update({f"{model}_f_in": f_in, f"{model}_omega": om, f"{model}_nut_in": nut_in,
                    f"{model}_sgs": sgs, f"{model}_f_out": f_out, f"{model}_macro": macro,
                    f"{model}_nut_out": nut_out})
    np.savez(os.path.join(HERE, "g6_k17_turbulence.npz"), **out)
    return {"seed": seed, "n_per_model": n, "models": ["smagorinsky", "amd", "qr"], "limiter": 0.01,
            "note": "nut_in 0 for every 8th node, else U[0, 0.05); SGS constant U[0.05, 0.5]"}


GENERATORS = {"g1": g1, "g2": g2, "g3": g3, "g4": g4, "g6": g6}


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    path = os.path.join(HERE, "golden.json")
    meta = json.load(open(path)) if os.path.exists(path) else {}
    meta.update({"generator": "tests/golden/make_golden.py", "reference": "oracle/_ref/libvfref.so"})
    for name in sys.argv[1:] or list(GENERATORS):
        meta[name] = GENERATORS[name]()
    with open(path, "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
