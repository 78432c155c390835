"""Write the golden fixtures of tests/golden/ from the reference's own gravity
code, compiled in this container by `make -C oracle ref` into
oracle/_ref/libvipp_ref.so.

Every fixture is data: inputs and the reference's outputs for them.  Run from
the repository root:  python tests/golden/make_golden.py
(needs the reference checkout; the fixtures it writes are committed, so the
tests never need the reference itself).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import Oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def tree_arrays(prefix, t, lists=None):
    d = {f"{prefix}parent": t.parent, f"{prefix}level": t.level, f"{prefix}bmin": t.bmin, f"{prefix}bmax": t.bmax,
         f"{prefix}child_base": t.child_base, f"{prefix}pt_off": t.pt_off, f"{prefix}pts": t.pts}
    if lists is not None:
        for k in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
            d[f"{prefix}{k}"] = getattr(lists, k)
    return d


def uniform(n, seed):
    return np.random.default_rng(seed).random((n, 3))


def main():
    ref = Oracle("ref")
    # -- trees and interaction lists (octree.hpp, solver_ref.hpp:115-162) ----
    cases = {}
    x0, _, _ = ref.problem_initialize(16384, 0, 1e-3)      # Plummer-like Problem::initialize, N=16384
    cases["plummer16k_s64"] = (x0, 64)
    cases["plummer4k_s5"] = (ref.problem_initialize(4096, 3, 1e-3)[0], 5)
    cases["uniform2k_s1"] = (uniform(2000, 1), 1)
    cases["uniform20k_s64"] = (uniform(20000, 2), 64)
    xs = uniform(3000, 4)
    xs[:700] = 0.25 + 1e-3 * uniform(700, 5)                # a dense clump: deep, ragged tree
    cases["clump3k_s2"] = (xs, 2)
    for name, (x, s) in cases.items():
        t, lists = ref.tree_and_lists(x, s)
        np.savez_compressed(os.path.join(OUT, f"tree_{name}.npz"), x=x, s=s, **tree_arrays("", t, lists))
        print(name, "cells", t.ncells)

    # -- FMM expansions and accelerations (solver_ref.hpp:164-232) ----------
    xf, _, _ = ref.problem_initialize(2000, 11, 1e-3)
    d = {"x": xf, "s": 10, "eps": 1e-3, "G": 1.0}
    for p in (2, 4, 8):
        with ref.build(xf, 10) as h:
            out = h.fmm(p, m=1.0 / len(xf), G=1.0, eps=1e-3, diagnostics=True)
        d[f"a_p{p}"] = out["a"]
        if p <= 4:
            d[f"M_p{p}"] = out["M"]
            d[f"L_p{p}"] = out["L"]
    d["a_direct"] = ref.direct(xf, eps=1e-3)
    np.savez_compressed(os.path.join(OUT, "fmm_plummer2k.npz"), **d)

    # -- one RK4 step, FMM and direct (solver_ref.hpp:36-104, solver_ref_naive.hpp:33-100)
    xr, vr, _ = ref.problem_initialize(1000, 5, 1e-3)
    dt = 1e-3
    x1, v1 = ref.rk4_fmm(xr, vr, dt, 10, 4, eps=1e-3)
    x2, v2 = ref.rk4_direct(xr, vr, dt, eps=1e-3)
    np.savez_compressed(os.path.join(OUT, "rk4_1k.npz"), x=xr, v=vr, dt=dt, s=10, p=4, eps=1e-3,
                        x_fmm=x1, v_fmm=v1, x_direct=x2, v_direct=v2)

    # -- special functions (fmm.hpp:20-221) -------------------------------
    angles = np.array([[0.3, 0.7], [1.1, -2.0], [2.9, 3.0], [np.pi / 2, 0.0], [0.0, 0.0], [1e-3, 1.0]])
    sp = {"angles": angles}
    for k, (th, ph) in enumerate(angles):
        sp[f"P{k}"] = ref.P(12, np.cos(th))
        sp[f"DP{k}"] = ref.DP(12, th)
        sp[f"Y{k}"] = ref.Y(24, th, ph)
    sp["A"] = np.array([[ref.A(n, m) if abs(m) <= n else 0.0 for m in range(-24, 25)] for n in range(25)])
    np.savez_compressed(os.path.join(OUT, "special.npz"), **sp)

    # -- Problem::initialize stream and delta_t (problem.hpp:188-243) --------
    pi = {}
    for n, seed, eps in ((1000, 0, None), (257, 7, 1e-3)):
        x, v, e = ref.problem_initialize(n, seed, eps)
        pi[f"x_{n}_{seed}"] = x
        pi[f"v_{n}_{seed}"] = v
        pi[f"eps_{n}_{seed}"] = e
        pi[f"dt_{n}_{seed}"] = ref.delta_t(e)
    np.savez_compressed(os.path.join(OUT, "problem_init.npz"), **pi)

    # -- E(T) curve fit (benchmark.hpp:198-241) -----------------------------
    t = np.array([1.0, 2, 4, 8, 16])
    rng = np.random.default_rng(9)
    fits = []
    for _ in range(6):
        e = 1e-7 * t ** rng.uniform(-5, 1) * (1 + 0.1 * rng.standard_normal(5))
        fits.append(np.concatenate([e, ref.curve_fit(e, t)]))
    np.savez_compressed(os.path.join(OUT, "curve_fit.npz"), t=t, rows=np.array(fits))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
