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
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6


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
    ap.add_argument("--err-sample", type=int, default=65536)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size, the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to HBM bytes")
    ap.add_argument("--learned-workload", default="mixture4m_p4",
                    help="workload of the learned-closure comparison (p=4 + closure against plain FP32 p=12); 'none' skips it")
    ap.add_argument("--fallback-fraction", type=float, default=0.1,
                    help="fraction of leaves (on the training pairs) whose closure estimate sends them to the order-12 far field")
    ap.add_argument("--learned-err-sample", type=int, default=16384)
    ap.add_argument("--drift-workload", default="collision16m",
                    help="workload of the energy-drift comparison (split p=12 against plain FP32 p=12); 'none' skips it")
    ap.add_argument("--drift-steps", type=int, default=20)
    ap.add_argument("--no-compare-plain", dest="compare_plain", action="store_false",
                    help="skip the plain FP32 p=12 GPU FMM run reported beside the main mode")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)

    import vipp_gravity as vgm
    import workloads

    if args.calib_copy_mib > 0:
        calibration_copy(args.calib_copy_mib << 20)

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

    # accuracy on the initial state (untimed): FP64 direct sum on the device for
    # a random sample, cross-checked against the CPU oracle on its first 32
    a = g.accelerations()
    cnt = g.counts()
    err = None
    err_ref = None
    if args.err_sample > 0:
        idx = np.sort(np.random.default_rng(1).choice(n, size=min(n, args.err_sample), replace=False))
        d = g.direct(idx)
        err_ref = (idx, d)
        e = np.linalg.norm(a[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
        err = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max()), "sample": int(len(idx)),
               "reference": "fp64 direct sum on the device (vipp_grav_direct)"}
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        k = min(32, len(idx))
        dc = Oracle("orc").direct_subset(x, idx[:k], m=1.0 / n, G=1.0, eps=w["eps"], masses=masses)
        err["direct_vs_cpu_oracle_max_rel"] = float(np.max(np.linalg.norm(d[:k] - dc, axis=1) / np.linalg.norm(dc, axis=1)))
    g.upload(x, v, masses)

    for _ in range(args.warmup):
        g.step(w["dt"], 1)
    # the timed steps run without per-phase events (those synchronise after
    # every evaluation); the same number of steps follows with them, for the
    # phase breakdown and the M2L kernel time
    t0 = time.perf_counter()
    g.step(w["dt"], args.steps)  # returns after the device finished
    t1 = time.perf_counter()
    g.set_timing(True)
    g.phase_times(reset=True)
    g.step(w["dt"], args.steps)
    ph = g.phase_times()
    g.set_timing(False)
    cnt = g.counts()  # work of the last evaluation
    # host-buffer round trip as the reference's preprocess/solve/postprocess
    # sees it: upload x, v; one step; download x, v (not the headline value)
    t2 = time.perf_counter()
    g.upload(x, v, masses)
    g.step(w["dt"], 1)
    g.download()
    pcie_steps = 1.0 / (time.perf_counter() - t2)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 1e3 / ms_per_step
    nev = 4 * args.steps
    per_eval = {k: v / nev for k, v in ph.items()}
    g.close()

    # the dominant single kernel: M2L is one launch per evaluation
    p = w["p"]
    m2l_flops = 2.0 * (p + 1) ** 4 * cnt["m2l_live_pairs"]
    ms = per_eval["m2l"]
    split = mode == "split"
    peak = {"split": PEAK_BF16_MFMA_TFLOPS, "f32": PEAK_FP32_TFLOPS, "f64": PEAK_FP64_TFLOPS}[mode]
    achieved = m2l_flops / (ms * 1e-3) / 1e12
    roofline = {"kernel": "k_m2l_split" if split else "k_m2l_plain", "bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": m2l_traffic(args.workload, mode),
                "flop_per_launch": m2l_flops, "ms_per_launch": ms,
                "work": f"2(p+1)^4 = {2 * (p + 1) ** 4} FLOP per M2L pair whose source and target hold particles"}

    plain = None
    if args.compare_plain and mode != "f32":
        plain = plain_reference(vgm, x, v, masses, w, n, args.steps, value, err_ref)

    learned = None
    if args.learned_workload != "none":
        learned = learned_comparison(vgm, workloads, args)

    drift = None
    if args.drift_workload != "none" and args.drift_steps > 0:
        drift = energy_drift(vgm, workloads, args)

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
        "vs_plain_gpu": plain,
        "learned": learned,
        "energy_drift": drift,
        "host_roundtrip_steps_per_s": pcie_steps,
        "phase_ms_per_eval": per_eval,
        "p2p_pairs_per_s": cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3) if per_eval["p2p"] > 0 else None,
        "p2p_tflops": 20.0 * cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3) / 1e12 if per_eval["p2p"] > 0 else None,
        "m2l_pairs_per_s": cnt["m2l_live_pairs"] / (per_eval["m2l"] * 1e-3) if per_eval["m2l"] > 0 else None,
        "counts": cnt,
        "force_err": err,
    }
    print(json.dumps(line))


def plain_reference(vgm, x, v, masses, w, n, steps, value, err_ref=None):
    """The plain GPU FMM (FP32 P2P, FP32 M2L on the vector units, same p) on the
    same input and step count: the comparison the headline metric asks for.
    err_ref = (idx, direct accelerations): its force error on the same sample."""
    g = vgm.Gravity(n, w["s"], w["p"], "f32", eps=w["eps"], G=1.0, mass=1.0 / n)
    g.upload(x, v, masses)
    err = None
    if err_ref is not None:
        idx, d = err_ref
        a = g.accelerations()
        e = np.linalg.norm(a[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
        err = {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "sample": int(len(idx))}
    g.step(w["dt"], 1)
    g.upload(x, v, masses)
    t0 = time.perf_counter()
    g.step(w["dt"], steps)
    t1 = time.perf_counter()
    g.close()
    plain = steps / (t1 - t0)
    return {"mode": "plain fp32 p2p + fp32 m2l", "steps_per_s": plain, "speedup": value / plain, "force_err": err}


def energy_drift(vgm, workloads, args):
    """Relative total-energy drift |E(T) - E(0)| / |E(0)| over K RK4 steps of the
    two-Plummer collision, split p = 12 beside plain FP32 p = 12 (FP64 state
    and energy reductions in both)."""
    wl = args.drift_workload
    wdef = workloads.WORKLOADS[wl]
    x, v, masses = workloads.make(wl, seed=0)
    n = len(x)
    out = {"workload": wl, "n": n, "steps": args.drift_steps, "dt": wdef["dt"], "p": wdef["p"]}
    for name, mode in (("split_p12", "split"), ("plain_fp32_p12", "f32")):
        with vgm.Gravity(n, wdef["s"], wdef["p"], mode, eps=wdef["eps"], G=1.0, mass=1.0 / n) as g:
            g.upload(x, v, masses)
            e0 = sum(g.energy())
            t0 = time.perf_counter()
            g.step(wdef["dt"], args.drift_steps)
            t1 = time.perf_counter()
            e1 = sum(g.energy())
        out[name] = {"relative_drift": abs(e1 - e0) / abs(e0), "E0": e0, "E1": e1, "steps_per_s": args.drift_steps / (t1 - t0)}
    return out


def learned_comparison(vgm, workloads, args):
    """p = 4 FMM + learned closure (+ order-12 fallback) against the plain FP32
    p = 12 GPU FMM on the same input: steps/s and force errors.  The network is
    trained on the device on pairs the FMM generates from three other seeds
    of the workload (never the evaluated one)."""
    import closure
    wl = args.learned_workload
    wdef = workloads.WORKLOADS[wl]
    s_, eps, dt = wdef["s"], wdef["eps"], wdef["dt"]
    t0 = time.perf_counter()
    X, T = [], []
    for seed in (1, 2, 3):
        xs, _, ms = workloads.make(wl, seed=seed)
        a, b = closure.training_pairs(xs, s_, eps, ms)
        X.append(a)
        T.append(b)
    X, T = np.concatenate(X), np.concatenate(T)
    t1 = time.perf_counter()
    params, io, hist = closure.train(X, T, epochs=40, batch=512, lr=2e-3, seed=0)
    t2 = time.perf_counter()
    bound = closure.bound_for_fraction(params, io, X, args.fallback_fraction)
    x, v, masses = workloads.make(wl, seed=0)
    n = len(x)
    idx = np.sort(np.random.default_rng(1).choice(n, size=min(n, args.learned_err_sample), replace=False))

    def run(p, mode, with_closure, b=float("inf")):
        g = vgm.Gravity(n, s_, p, mode, eps=eps, G=1.0, mass=1.0 / n)
        g.upload(x, v, masses)
        if with_closure:
            g.set_closure(params, io, b)
        a = g.accelerations()
        cnt = g.counts()
        d = g.direct(idx) if run.d is None else run.d
        run.d = d
        g.step(dt, 1)  # warm-up
        g.upload(x, v, masses)
        g.set_timing(True)
        g.phase_times(reset=True)
        ta = time.perf_counter()
        g.step(dt, args.steps)
        tb = time.perf_counter()
        ph = g.phase_times()
        g.close()
        e = np.linalg.norm(a[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
        return {"steps_per_s": args.steps / (tb - ta), "err_median": float(np.median(e)), "err_p99": float(np.percentile(e, 99)),
                "phase_ms_per_eval": {k: v_ / (4 * args.steps) for k, v_ in ph.items()}, "fallback_leaves": cnt["fallback_leaves"],
                "leaves": cnt["leaves"]}, a
    run.d = None
    plain, a12 = run(12, "f32", False)
    split12, _ = run(12, "split", False)
    low, a4 = run(4, "split", False)
    learned, ac = run(4, "split", True, bound)
    learned_nofb, acn = run(4, "split", True)
    diff_to_p12 = lambda a: float(np.median(np.linalg.norm(a[idx] - a12[idx], axis=1) / np.linalg.norm(a12[idx], axis=1)))
    return {
        "workload": wl, "n": n, "train_pairs": int(len(X)), "train_seeds": [1, 2, 3], "eval_seed": 0,
        "pairs_s": t1 - t0, "train_s": t2 - t1, "train_loss_first_last": [float(hist[0]), float(hist[-1])],
        "fallback_fraction_target": args.fallback_fraction, "bound": bound,
        "plain_fp32_p12": plain, "split_p12": split12, "p4_no_closure": low, "p4_closure": learned_nofb, "p4_closure_fallback": learned,
        "speedup_vs_plain": learned["steps_per_s"] / plain["steps_per_s"],
        "median_diff_to_p12": {"p4": diff_to_p12(a4), "p4_closure": diff_to_p12(acn), "p4_closure_fallback": diff_to_p12(ac)},
        "err_sample": int(len(idx)), "err_reference": "fp64 direct sum on the device",
    }


def calibration_copy(nbytes):
    """One hipMemcpy device-to-device of nbytes (reads and writes nbytes once)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(C.byref(buf), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0 or hip.hipMemset(b, 0, C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # 3: device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def m2l_traffic(workload, mode):
    """HBM bytes per M2L launch from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (scaled by the copy-kernel calibration), if present."""
    name = {"split": "k_m2l_split", "f32": "k_m2l_plain"}.get(mode)
    path = os.path.join(ROOT, "profiles", f"{name}_{workload}_traffic.txt")
    if name is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(x, w, n):
    """The reference's FMM compiled from its own headers (oracle/_ref), one thread,
    on a bounded sample of one force evaluation of the same input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.available("ref"):
        return None
    ref = oracle.Oracle("ref")
    per = 32
    secs, frac = ref.time_sample(x, w["s"], w["p"], 1.0 / n, 1.0, w["eps"], per)
    t_eval = float(secs.sum())
    return {"value": 1.0 / (4 * t_eval), "unit": "steps/s", "cores": 1, "kind": "reference",
            "sample": (f"one force evaluation of the same {n}-particle input: Octree build and calc_link in full; P2M, M2M, M2L, "
                       f"P2P, L2L/L2P on {per} evenly strided work items each (sampled work fractions "
                       f"{', '.join(f'{f:.1e}' for f in frac[2:])}), scaled by exact work counts; 4 evaluations per RK4 step"),
            "seconds_per_eval": {k: float(s) for k, s in zip(["tree", "lists", "p2m", "m2m", "m2l", "p2p", "l2l_l2p"], secs)}}


if __name__ == "__main__":
    main()
