"""Synthetic initial conditions of the benchmark workloads (numpy, seeded).

Plummer spheres follow the reference's recipe (problem.hpp:188-211,250-292):
radius from the inverse mass profile r = a / sqrt(u^(-2/3) - 1), isotropic
direction by calc_xyz (x = (1 - 2 X3) r, the rest on a circle of angle 2 pi X4),
speed q sqrt(2) (1 + r^2)^(-1/4) with q drawn by Aarseth's rejection on
q^2 (1 - q^2)^(7/2) under P_MAX = sqrt(7) 686 / 19683.  Unlike the reference,
u is drawn on (0, u_max] with u_max = (1 + (a / r_max)^2)^(-3/2), i.e. the
sphere is truncated at r_max = 20 a, and the stream is numpy's PCG64.
"""
from __future__ import annotations

import math

import numpy as np

P_MAX = math.sqrt(7.0) * 686 / 19683


def _calc_xyz(norm, r, X):
    x = (1 - 2 * r) * norm
    yz = np.sqrt(np.maximum(norm * norm - x * x, 0.0))
    return np.stack([x, yz * np.cos(2 * np.pi * X), yz * np.sin(2 * np.pi * X)], 1)


def _aarseth_q(rng, n):
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x1 = rng.random(todo.size)
        x2 = rng.random(todo.size)
        ok = P_MAX * x1 < x2 * x2 * (1 - x2 * x2) ** 3.5
        q[todo[ok]] = x2[ok]
        todo = todo[~ok]
    return q


def plummer(n, seed=0, a=1.0, rmax_over_a=20.0, G=1.0, M=1.0, rng=None):
    """Positions and velocities of an n-body Plummer sphere of scale a and mass M."""
    rng = np.random.default_rng(seed) if rng is None else rng
    umax = (1 + rmax_over_a ** -2) ** -1.5
    u = rng.uniform(0, umax, n)
    u = np.where(u <= 0, umax, u)
    r = 1.0 / np.sqrt(u ** (-2.0 / 3) - 1)
    x = _calc_xyz(r, rng.random(n), rng.random(n)) * a
    q = _aarseth_q(rng, n)
    vnorm = q * math.sqrt(2.0) * (1 + r * r) ** -0.25 * math.sqrt(G * M / a)
    v = _calc_xyz(vnorm, rng.random(n), rng.random(n))
    return x, v


def uniform_cube(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    return x, np.zeros_like(x), m


def plummer_mixture(n, k=32, seed=0, amin=0.02, amax=0.2):
    """k equal-mass Plummer spheres, scales log-uniform in [amin, amax], centres uniform in the unit cube."""
    rng = np.random.default_rng(seed)
    per = n // k
    xs, vs = [], []
    for i in range(k):
        cnt = per if i < k - 1 else n - per * (k - 1)
        a = math.exp(rng.uniform(math.log(amin), math.log(amax)))
        c = rng.random(3)
        x, v = plummer(cnt, a=a, M=1.0 / k, rng=rng)
        xs.append(x + c)
        vs.append(v)
    return np.concatenate(xs), np.concatenate(vs)


def two_plummer(n, seed=0, sep=5.0, vbulk=0.3):
    rng = np.random.default_rng(seed)
    h = n // 2
    x1, v1 = plummer(h, M=0.5, rng=rng)
    x2, v2 = plummer(n - h, M=0.5, rng=rng)
    x1[:, 0] -= sep
    v1[:, 0] += vbulk
    x2[:, 0] += sep
    v2[:, 0] -= vbulk
    return np.concatenate([x1, x2]), np.concatenate([v1, v2])


# The workloads: N, s, p, softening, dt, steps of the configuration and the
# arithmetic it names (f64, f32 = plain FP32, split = BF16 hi/lo on MFMA)
WORKLOADS = {
    "plummer16k_p4": dict(n=16384, s=64, p=4, eps=1e-3, dt=1e-3, steps=10, mode="f64"),
    "uniform1m_p8": dict(n=1 << 20, s=64, p=8, eps=1e-4, dt=1e-4, steps=10, mode="f32"),
    "plummer4m_p12": dict(n=1 << 22, s=64, p=12, eps=1e-3, dt=5e-4, steps=20, mode="split"),
    # the mixture's softening is left open by its configuration: 1e-4 (the
    # uniform workload's) keeps the softened-direct-sum floor of the FP64
    # p = 12 FMM at 2e-6 median; at 1e-3 the deepest cells (level 11, width
    # 7e-4) are smaller than the softening and that floor is 1.8e-4
    "mixture4m_p4": dict(n=1 << 22, s=64, p=4, eps=1e-4, dt=5e-4, steps=20, mode="split"),
    "collision16m": dict(n=1 << 24, s=64, p=12, eps=1e-3, dt=1e-3, steps=100, mode="split"),
    # the reference's own initial conditions (Problem::initialize, problem.hpp:188-211,
    # mt19937_64 seed 0, the mass fraction truncated at 1 - eps) at the headline's size
    "refinit4m_p12": dict(n=1 << 22, s=64, p=12, eps=1e-3, dt=5e-4, steps=3, mode="split",
                          source="the reference's Problem::initialize stream (vipp_grav_problem_initialize)"),
}


def make(name, seed=0):
    """(x, v, masses or None) of a named workload."""
    w = WORKLOADS[name]
    n = w["n"]
    if name.startswith("plummer"):
        x, v = plummer(n, seed)
        return x, v, None
    if name.startswith("uniform"):
        return uniform_cube(n, seed)
    if name.startswith("mixture"):
        x, v = plummer_mixture(n, seed=seed)
        return x, v, None
    if name.startswith("collision"):
        x, v = two_plummer(n, seed)
        return x, v, None
    if name.startswith("refinit"):
        import vipp_gravity
        x, v, _ = vipp_gravity.problem_initialize(n, seed, w["eps"])
        return x, v, None
    raise KeyError(name)
