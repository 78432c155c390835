#!/usr/bin/env python3
"""RK4 steps per second of the MI355X gravity FMM on one GPU.

Workload (default): N = 2^22 Plummer sphere (a = 1, truncated at 20 a), leaf
parameter s = 64 (split above 181 points), expansion order p = 12, softening
1e-3, dt = 5e-4; RK4 state in FP64, P2P in FP32, M2L by mode (plain FP32 on
VALU or BF16 hi/lo split on MFMA with FP32 accumulation).

One step is one classic RK4 step: four force evaluations, each rebuilding the
octree and the interaction lists as the reference does (solver_ref.hpp:36-104,
164-177).  The particles are uploaded once; the timed region holds K steps on
device-resident data, bracketed by device synchronisation.

Prints one JSON line.  Besides value (steps/s) it carries:
  roofline      the dominant kernel's algorithmic FLOP/s from its HIP-event time
  cpu_baseline  the reference's own FMM (oracle/_ref, single thread) on a
                bounded sample of one force evaluation of the same input
  force_err     relative force error against an FP64 direct sum on a sample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vipp-nbody-bench_amd"))

METRIC = "FMM RK4 steps/sec at N=4M; force err vs FP64 direct; learned vs plain GPU"
PEAK_FP32_VECTOR_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plummer4m_p12")
    ap.add_argument("--mode", default="split", choices=["plain", "split", "f64"])
    ap.add_argument("--p", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--err-sample", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)

    import vipp_gravity as vgm
    import workloads

    w = dict(workloads.WORKLOADS[args.workload])
    if args.p:
        w["p"] = args.p
    x, v, masses = workloads.make(args.workload)
    if args.n:
        x, v = x[:args.n], v[:args.n]
        masses = None if masses is None else masses[:args.n]
    n = len(x)
    mode = {"plain": "f32", "split": "split", "f64": "f64"}[args.mode]
    g = vgm.Gravity(n, w["s"], w["p"], mode, eps=w["eps"], G=1.0, mass=1.0 / n)
    g.upload(x, v, masses)

    # accuracy on the initial state (untimed)
    a = g.accelerations()
    cnt = g.counts()
    err = None
    if args.err_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        idx = np.random.default_rng(1).choice(n, size=min(n, args.err_sample), replace=False)
        d = Oracle("orc").direct_subset(x, idx, m=1.0 / n, G=1.0, eps=w["eps"], masses=masses)
        e = np.linalg.norm(a[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
        err = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max()), "sample": int(len(idx))}
    g.upload(x, v, masses)

    for _ in range(args.warmup):
        g.step(w["dt"], 1)
    g.set_timing(True)
    g.phase_times(reset=True)
    t0 = time.perf_counter()
    g.step(w["dt"], args.steps)  # returns after the device finished
    t1 = time.perf_counter()
    ph = g.phase_times()
    g.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 1e3 / ms_per_step
    nev = 4 * args.steps
    per_eval = {k: v / nev for k, v in ph.items()}

    p = w["p"]
    m2l_flops = 2.0 * (p + 1) ** 4 * cnt["m2l_live_pairs"]
    p2p_flops = 20.0 * cnt["p2p_pairs"]
    m2l_peak = PEAK_BF16_MFMA_TFLOPS if mode == "split" else PEAK_FP32_VECTOR_TFLOPS
    cand = {
        "m2l": (m2l_flops, per_eval["m2l"], m2l_peak),
        "p2p": (p2p_flops, per_eval["p2p"], PEAK_FP32_VECTOR_TFLOPS),
    }
    dom = max(cand, key=lambda k: cand[k][1])
    fl, ms, peak = cand[dom]
    achieved = fl / (ms * 1e-3) / 1e12
    roofline = {"kernel": dom, "bound": "mfma" if (dom == "m2l" and mode == "split") else "valu", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "work_per_launch": fl, "ms_per_launch": ms}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, w, n)

    line = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if mode == "f64" else "f32", "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "s": w["s"], "p": p, "eps": w["eps"], "dt": w["dt"],
                   "m2l": {"f32": "fp32 VALU", "split": "bf16 hi/lo split on MFMA, fp32 accumulate", "f64": "fp64"}[mode],
                   "p2p": "fp64" if mode == "f64" else "fp32", "state": "fp64", "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phase_ms_per_eval": per_eval,
        "p2p_pairs_per_s": cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3) if per_eval["p2p"] > 0 else None,
        "m2l_pairs_per_s": cnt["m2l_live_pairs"] / (per_eval["m2l"] * 1e-3) if per_eval["m2l"] > 0 else None,
        "counts": cnt,
        "force_err": err,
    }
    print(json.dumps(line))


def cpu_baseline(x, w, n):
    """The reference's FMM compiled from its own headers (oracle/_ref), one thread,
    on a bounded sample of one force evaluation of the same input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.available("ref"):
        return None
    ref = oracle.Oracle("ref")
    per = 16
    secs, frac = ref.time_sample(x, w["s"], w["p"], 1.0 / n, 1.0, w["eps"], per)
    t_eval = float(secs.sum())
    return {"value": 1.0 / (4 * t_eval), "unit": "steps/s", "cores": 1, "kind": "reference",
            "sample": (f"one force evaluation of the same {n}-particle input: Octree build and calc_link in full; P2M, M2M, M2L, "
                       f"P2P, L2L/L2P on {per} evenly strided work items each (sampled work fractions "
                       f"{', '.join(f'{f:.1e}' for f in frac[2:])}), scaled by exact work counts; 4 evaluations per RK4 step"),
            "seconds_per_eval": {k: float(s) for k, s in zip(["tree", "lists", "p2m", "m2m", "m2l", "p2p", "l2l_l2p"], secs)}}


if __name__ == "__main__":
    main()
