#!/usr/bin/env python3
"""RK4 steps per second of the MI355X gravity FMM on one GPU.

Workload (default): N = 2^22 Plummer sphere (a = 1, truncated at 20 a), leaf
parameter s = 64 (split above 181 points), expansion order p = 12, softening
1e-3, dt = 5e-4; RK4 state in FP64, P2P in FP32, M2L, M2M and L2L on the
matrix cores with BF16 hi/lo split operands and FP32 accumulation.

One step is one classic RK4 step: four force evaluations, each rebuilding the
octree and the interaction lists as the reference does (solver_ref.hpp:36-104,
164-177).  The particles are uploaded once; the timed region holds K steps on
device-resident data, bracketed by device synchronisation.

Prints one JSON line.  Besides value (steps/s) it carries:
  roofline      the dominant kernel (split M2L) at its algorithmic FLOP/s
                from its HIP-event time, against dense BF16 MFMA peak
  phases        every phase's algorithmic work, time and fraction of the
                ceiling that bounds it, and the step composite
  cpu_baseline  the reference's own FMM (oracle/_ref) on the host's cores, on
                a bounded sample of one force evaluation of the same input
  force_err     relative force error against an FP64 direct sum on a sample
  vs_plain_gpu  the plain FP32 p = 12 GPU FMM on the same input
  order_sweep   the same input at other orders (split p = 8 and 10 by
                default): steps/s and force error, and the speed-up over
                the plain FP32 p = 12 run
  configs       the other workloads: 16k Plummer p = 4 FP64, 2^20 uniform
                p = 8 FP32 with per-particle masses, and the 4M Plummer from
                the reference's own initial-condition stream
  learned       the 4M mixture of 32 Plummer spheres: p = 4 + learned closure
                + order-12 fallback against plain FP32 p = 12 and split p = 12
  energy_drift  the 16M two-Plummer collision over 100 steps
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
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3          # vector FP32 with packed (two-wide) FMA
PEAK_FP32_SCALAR_TFLOPS = 78.6    # the same without packed instructions (this build emits none)
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6
TREE_BYTES_PER_PARTICLE = 136     # AABB 24 + codes 24+8 + single-pass sort 24 + permute 48 + counts 8
LIST_BYTES_PER_ENTRY = 8          # read the parent-neighbour id, write the entry
RK4_BYTES_PER_PARTICLE = 228      # per stage: FP64 x0 v0 v xo vo read, FP32 a read, x v xo vo written
P2P_FLOP_PER_PAIR = 20


def m2l_flop_per_pair(p):
    return 2.0 * (p + 1) ** 4


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
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="host threads of the CPU baseline (default: OMP_NUM_THREADS, else every core this process may use)")
    ap.add_argument("--err-sample", type=int, default=65536)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size, the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to HBM bytes")
    ap.add_argument("--configs", default="plummer16k_p4,uniform1m_p8,refinit4m_p12",
                    help="other workloads reported beside the headline (comma list; 'none' skips them)")
    ap.add_argument("--learned-workload", default="mixture4m_p4",
                    help="workload of the learned-closure comparison (p=4 + closure against plain FP32 p=12); 'none' skips it")
    ap.add_argument("--learned-err-sample", type=int, default=16384)
    ap.add_argument("--drift-workload", default="collision16m",
                    help="workload of the energy-drift comparison; 'none' skips it")
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--no-compare-plain", dest="compare_plain", action="store_false",
                    help="skip the plain FP32 p=12 GPU FMM run reported beside the main mode")
    ap.add_argument("--order-sweep", default="split:8,10",
                    help="the headline input at other expansion orders, as mode:p,p;mode:p (modes plain, split, f64; 'none' skips): "
                         "steps/s and force error, with the speed-up over the plain run at the headline order")
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
    err = None
    err_ref = None
    if args.err_sample > 0:
        idx = np.sort(np.random.default_rng(1).choice(n, size=min(n, args.err_sample), replace=False))
        d = g.direct(idx)
        err_ref = (idx, d)
        err = rel_err(a[idx], d)
        err["reference"] = "fp64 direct sum on the device (vipp_grav_direct)"
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
    cnt = g.counts()  # work of an evaluation at the stepped positions
    # host-buffer round trip as the reference's preprocess/solve/postprocess
    # sees it: upload x, v; one step; download x, v (not the headline value)
    t2 = time.perf_counter()
    g.upload(x, v, masses)
    g.step(w["dt"], 1)
    g.download()
    pcie_steps = 1.0 / (time.perf_counter() - t2)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 1e3 / ms_per_step
    per_eval = {k: v_ / (4 * args.steps) for k, v_ in ph.items()}
    g.close()

    p = w["p"]
    peak_m2l = {"split": PEAK_BF16_MFMA_TFLOPS, "f32": PEAK_FP32_TFLOPS, "f64": PEAK_FP64_TFLOPS}[mode]
    phases = phase_rooflines(per_eval, cnt, n, p, mode, ms_per_step)
    m2l_flops = m2l_flop_per_pair(p) * cnt["m2l_live_pairs"]
    ms = per_eval["m2l"]
    achieved = m2l_flops / (ms * 1e-3) / 1e12
    roofline = {"kernel": "k_m2l_split" if mode == "split" else "k_m2l_plain", "bound": "mfma" if mode == "split" else "valu",
                "achieved": achieved, "peak": peak_m2l, "unit": "TFLOP/s", "frac": achieved / peak_m2l,
                "traffic": m2l_traffic(args.workload, mode), "flop_per_launch": m2l_flops, "ms_per_launch": ms,
                "work": f"2(p+1)^4 = {int(m2l_flop_per_pair(p))} FLOP per M2L pair whose source and target hold particles"}

    plain = None
    if args.compare_plain and mode != "f32":
        plain = plain_reference(vgm, x, v, masses, w, n, args.steps, value, err_ref)

    sweep = None
    if args.order_sweep != "none":
        sweep = order_sweep(vgm, x, v, masses, w, parse_sweep(args.order_sweep), args.steps, err_ref, plain)

    configs = None
    if args.configs != "none":
        configs = {name: config_line(vgm, workloads, name) for name in args.configs.split(",") if name}

    learned = None
    if args.learned_workload != "none":
        learned = learned_comparison(vgm, workloads, args)

    drift = None
    if args.drift_workload != "none" and args.drift_steps > 0:
        drift = energy_drift(vgm, workloads, args)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, w, n, args.cpu_threads)

    line = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"split": "bf16-split+f32+f64", "f32": "f32+f64", "f64": "f64"}[mode], "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "s": w["s"], "p": p, "eps": w["eps"], "dt": w["dt"],
                   "m2l": {"f32": "fp32 VALU", "split": "bf16 hi/lo split on MFMA (3 products), fp32 accumulate", "f64": "fp64"}[mode],
                   "m2m_l2l": {"f32": "fp32 VALU", "split": "bf16 three-way split on MFMA (6 products), fp32 accumulate", "f64": "fp64"}[mode],
                   "p2p": "fp64" if mode == "f64" else "fp32", "p2m_l2p": "fp64" if mode == "f64" else "fp32",
                   "state": "fp64", "parallelism": "1 GPU"},
        "roofline": roofline,
        "phases": phases,
        "cpu_baseline": cpu,
        "vs_plain_gpu": plain,
        "order_sweep": sweep,
        "configs": configs,
        "learned": learned,
        "energy_drift": drift,
        "host_roundtrip_steps_per_s": pcie_steps,
        "phase_ms_per_eval": per_eval,
        "tree_GBps": phases["tree"]["achieved"],
        "lists_GBps": phases["lists"]["achieved"],
        "p2p_pairs_per_s": cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3) if per_eval["p2p"] > 0 else None,
        "p2p_tflops": P2P_FLOP_PER_PAIR * cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3) / 1e12 if per_eval["p2p"] > 0 else None,
        "m2l_pairs_per_s": cnt["m2l_live_pairs"] / (per_eval["m2l"] * 1e-3) if per_eval["m2l"] > 0 else None,
        "counts": cnt,
        "force_err": err,
    }
    print(json.dumps(line))


def rel_err(a, d):
    e = np.linalg.norm(a - d, axis=1) / np.linalg.norm(d, axis=1)
    return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)), "max": float(e.max()), "sample": int(len(e))}


def phase_rooflines(per_eval, cnt, n, p, mode, ms_per_step):
    """Each phase's algorithmic work per force evaluation, its time alone (HIP
    events, one stream), the ceiling that bounds it and the fraction reached;
    and the step composite sum_phase (work / peak) / measured step time, where
    the step holds 4 evaluations and 4 RK4 stages."""
    split = mode == "split"
    mfma_or_valu = PEAK_BF16_MFMA_TFLOPS if split else (PEAK_FP64_TFLOPS if mode == "f64" else PEAK_FP32_TFLOPS)
    p2p_peak = PEAK_FP64_TFLOPS if mode == "f64" else PEAK_FP32_TFLOPS
    inner_cells = (cnt["cells"] - 1) / 8.0  # every internal cell has 8 children
    rows = {
        "tree": (TREE_BYTES_PER_PARTICLE * n / 1e9, "GB", PEAK_HBM_GBS, "GB/s", "hbm", f"{TREE_BYTES_PER_PARTICLE} B per particle"),
        "lists": (LIST_BYTES_PER_ENTRY * cnt["list_entries"] / 1e9, "GB", PEAK_HBM_GBS, "GB/s", "hbm", f"{LIST_BYTES_PER_ENTRY} B per list entry"),
        "p2p": (P2P_FLOP_PER_PAIR * cnt["p2p_pairs"] / 1e12, "TFLOP", p2p_peak, "TFLOP/s", "valu", "20 FLOP per pair"),
        "m2l": (m2l_flop_per_pair(p) * cnt["m2l_live_pairs"] / 1e12, "TFLOP", mfma_or_valu, "TFLOP/s", "mfma" if split else "valu",
                "2(p+1)^4 FLOP per live M2L pair"),
        # M2M and L2L: 2(p+1)^4 per child-parent pair; the phase time also holds P2M / L2P (no work unit: time share only)
        "upward": (m2l_flop_per_pair(p) * 8 * inner_cells / 1e12, "TFLOP", mfma_or_valu, "TFLOP/s", "mfma" if split else "valu",
                   "M2M 2(p+1)^4 per child-parent pair (P2M time share only)"),
        "downward": (m2l_flop_per_pair(p) * 8 * inner_cells / 1e12, "TFLOP", mfma_or_valu, "TFLOP/s", "mfma" if split else "valu",
                     "L2L 2(p+1)^4 per child-parent pair (L2P time share only)"),
        "rk4": (RK4_BYTES_PER_PARTICLE * n / 1e9, "GB", PEAK_HBM_GBS, "GB/s", "hbm", f"{RK4_BYTES_PER_PARTICLE} B per particle per stage"),
        "operators": (0.0, "-", None, "-", "-", "operator tables for the current root box: time share only"),
    }
    out = {}
    ideal_ms = 0.0
    for name, (work, wunit, peak, unit, bound, note) in rows.items():
        t = per_eval.get(name, 0.0)
        ach = work / (t * 1e-3) if t > 0 and work > 0 else None
        out[name] = {"ms_per_eval": t, "work": work, "work_unit": wunit, "achieved": ach, "peak": peak, "unit": unit, "bound": bound,
                     "frac": (ach / peak) if ach is not None and peak else None, "per_unit": note}
        if peak and work > 0:
            ideal_ms += 4 * work / peak * 1e3  # 4 evaluations (and 4 RK4 stages) per step
    out["composite"] = {"ideal_ms_per_step": ideal_ms, "measured_ms_per_step": ms_per_step, "frac": ideal_ms / ms_per_step,
                        "definition": "sum over phases of work / peak, over the measured step time (4 evaluations and stages per step)"}
    ach = out["p2p"]["achieved"]
    out["p2p"]["frac_of_scalar_fp32"] = ach / PEAK_FP32_SCALAR_TFLOPS if ach and mode != "f64" else None
    return out


def timed_run(vgm, x, v, masses, w, mode, p, steps, warmup=1, closure=None, err_idx=None, d=None):
    """steps/s of `steps` RK4 steps (after `warmup`), the phase times of as many
    more, the work counts, and the force error on err_idx against d."""
    n = len(x)
    with vgm.Gravity(n, w["s"], p, mode, eps=w["eps"], G=1.0, mass=1.0 / n) as g:
        g.upload(x, v, masses)
        if closure is not None:
            g.set_closure(*closure)
        out = {}
        if err_idx is not None:
            a = g.accelerations()
            if d is None:
                d = g.direct(err_idx)
            out["force_err"] = rel_err(a[err_idx], d)
        cnt = g.counts()
        for _ in range(warmup):
            g.step(w["dt"], 1)
        g.upload(x, v, masses)
        t0 = time.perf_counter()
        g.step(w["dt"], steps)
        t1 = time.perf_counter()
        g.set_timing(True)
        g.phase_times(reset=True)
        g.step(w["dt"], steps)
        ph = g.phase_times()
    per_eval = {k: v_ / (4 * steps) for k, v_ in ph.items()}
    out.update({"steps_per_s": steps / (t1 - t0), "phase_ms_per_eval": per_eval, "counts": cnt})
    if per_eval["p2p"] > 0:
        out["p2p_pairs_per_s"] = cnt["p2p_pairs"] / (per_eval["p2p"] * 1e-3)
    if per_eval["m2l"] > 0:
        out["m2l_pairs_per_s"] = cnt["m2l_live_pairs"] / (per_eval["m2l"] * 1e-3)
    return out, d


def parse_sweep(spec):
    """'split:8,10;plain:10' -> [("split", 8), ("split", 10), ("plain", 10)]."""
    out = []
    for part in spec.split(";"):
        if not part.strip():
            continue
        m, ps = part.split(":")
        if m not in ("plain", "split", "f64"):
            raise ValueError(f"unknown mode {m!r} in --order-sweep")
        out += [(m, int(q)) for q in ps.split(",") if q.strip()]
    return out


def order_sweep(vgm, x, v, masses, w, runs, steps, err_ref, plain):
    """The headline input at other expansion orders: steps/s and force error
    against the same FP64 direct-sum sample, and the speed-up over the plain
    FP32 run at the headline order (the error each order buys, side by side)."""
    idx, d = err_ref if err_ref is not None else (None, None)
    out = []
    for m, p in runs:
        mode = {"plain": "f32", "split": "split", "f64": "f64"}[m]
        r, _ = timed_run(vgm, x, v, masses, w, mode, p, steps, err_idx=idx, d=d)
        line = {"mode": m, "p": p, "steps_per_s": r["steps_per_s"], "force_err": r.get("force_err"),
                "phase_ms_per_eval": r["phase_ms_per_eval"]}
        if plain is not None:
            line["speedup_vs_plain_at_headline_p"] = r["steps_per_s"] / plain["steps_per_s"]
        out.append(line)
    return out


def config_line(vgm, workloads, name):
    """steps/s, interactions/s and force error of one of the other workloads,
    in the arithmetic its configuration names."""
    w = dict(workloads.WORKLOADS[name])
    mode = w.get("mode", "split")
    x, v, masses = workloads.make(name)
    n = len(x)
    idx = np.sort(np.random.default_rng(1).choice(n, size=min(n, 65536), replace=False))
    out, _ = timed_run(vgm, x, v, masses, w, mode, w["p"], w["steps"], err_idx=idx)
    out.update({"n": n, "s": w["s"], "p": w["p"], "eps": w["eps"], "dt": w["dt"], "steps": w["steps"], "mode": mode,
                "masses": "per particle" if masses is not None else "1/N", "source": w.get("source", "numpy generators (workloads.py)")})
    out["force_err"]["reference"] = "fp64 direct sum on the device"
    return out


def plain_reference(vgm, x, v, masses, w, n, steps, value, err_ref=None):
    """The plain GPU FMM (FP32 P2P, FP32 M2L on the vector units, same p) on the
    same input and step count: the comparison the headline metric asks for."""
    idx, d = err_ref if err_ref is not None else (None, None)
    out, _ = timed_run(vgm, x, v, masses, w, "f32", w["p"], steps, err_idx=idx, d=d)
    ms = out["phase_ms_per_eval"]["m2l"]
    flops = m2l_flop_per_pair(w["p"]) * out["counts"]["m2l_live_pairs"]
    tf = flops / (ms * 1e-3) / 1e12 if ms > 0 else None
    out.update({"mode": "plain fp32 p2p + fp32 m2l", "speedup": value / out["steps_per_s"],
                "m2l_tflops": tf, "m2l_frac_of_fp32": tf / PEAK_FP32_TFLOPS if tf else None,
                "m2l_frac_of_scalar_fp32": tf / PEAK_FP32_SCALAR_TFLOPS if tf else None})
    return out


def train_closure(vgm, closure, workloads, wl, seeds, n=None):
    X, T = [], []
    wdef = workloads.WORKLOADS[wl]
    for seed in seeds:
        xs, _, ms = workloads.make(wl, seed=seed)
        if n:
            xs, ms = xs[:n], (None if ms is None else ms[:n])
        a, b = closure.training_pairs(xs, wdef["s"], wdef["eps"], ms)
        X.append(a)
        T.append(b)
    X, T = np.concatenate(X), np.concatenate(T)
    params, io, hist = closure.train(X, T, epochs=40, batch=512, lr=2e-3, seed=0)
    return X, params, io, hist


def learned_comparison(vgm, workloads, args):
    """p = 4 FMM + learned closure (+ order-12 fallback) against the plain FP32
    p = 12 GPU FMM and the split p = 12 FMM on the same input.  The network is
    trained on the device on pairs the FMM generates from three other seeds of
    the workload (never the evaluated one).  The fallback bound is swept over
    the fraction of leaves sent to order 12; errors are stated against the
    FP64 direct sum and against the FP64 order-12 FMM (the reference of the
    closure), and the speed-up is stated where the closure's error reaches
    the plain FMM's."""
    import closure
    wl = args.learned_workload
    wdef = dict(workloads.WORKLOADS[wl])
    steps = max(2, args.steps)
    t0 = time.perf_counter()
    X, params, io, hist = train_closure(vgm, closure, workloads, wl, (1, 2, 3))
    t1 = time.perf_counter()
    x, v, masses = workloads.make(wl, seed=0)
    n = len(x)
    idx = np.sort(np.random.default_rng(1).choice(n, size=min(n, args.learned_err_sample), replace=False))
    with vgm.Gravity(n, wdef["s"], 12, "f64", eps=wdef["eps"], G=1.0, mass=1.0 / n) as g:
        g.upload(x, v, masses)
        ref12 = g.accelerations()[idx]
        d = g.direct(idx)

    def run(p, mode, cl=None):
        out, _ = timed_run(vgm, x, v, masses, wdef, mode, p, steps, closure=cl, err_idx=idx, d=d)
        with vgm.Gravity(n, wdef["s"], p, mode, eps=wdef["eps"], G=1.0, mass=1.0 / n) as g:
            g.upload(x, v, masses)
            if cl is not None:
                g.set_closure(*cl)
            a = g.accelerations()[idx]
            out["fallback_leaves"] = g.counts()["fallback_leaves"]
        out["err_vs_fp64_p12"] = rel_err(a, ref12)
        return out

    plain = run(12, "f32")
    split12 = run(12, "split")
    low = run(4, "split")
    nofb = run(4, "split", (params, io, float("inf")))
    sweep = []
    for frac in (0.1, 0.25, 0.5, 0.75, 1.0):
        b = closure.bound_for_fraction(params, io, X, frac) if frac < 1.0 else 0.0
        r = run(4, "split", (params, io, b))
        r.update({"fallback_fraction_target": frac, "bound": b})
        sweep.append(r)
    # equal error: the smallest swept fallback whose median error (against the
    # FP64 direct sum, and against the FP64 p = 12 FMM) is at or below plain's
    def at_equal(key):
        target = plain[key]["median"]
        for r in sweep:
            if r[key]["median"] <= target:
                return {"fallback_fraction_target": r["fallback_fraction_target"], "steps_per_s": r["steps_per_s"],
                        "median": r[key]["median"], "plain_median": target, "speedup_vs_plain": r["steps_per_s"] / plain["steps_per_s"]}
        return {"reached": False, "plain_median": target, "best_median": min(r[key]["median"] for r in sweep)}
    leaves = low["counts"]["leaves"]
    return {
        "workload": wl, "n": n, "eps": wdef["eps"], "train_pairs": int(len(X)), "train_seeds": [1, 2, 3], "eval_seed": 0,
        "train_s": t1 - t0, "train_loss_first_last": [float(hist[0]), float(hist[-1])], "leaves": leaves,
        "plain_fp32_p12": plain, "split_p12": split12, "p4_no_closure": low, "p4_closure": nofb, "p4_closure_fallback_sweep": sweep,
        "closure_gain_vs_p4": {"median_vs_fp64_p12": [low["err_vs_fp64_p12"]["median"], nofb["err_vs_fp64_p12"]["median"]],
                               "median_vs_direct": [low["force_err"]["median"], nofb["force_err"]["median"]]},
        "equal_error_vs_direct": at_equal("force_err"), "equal_error_vs_fp64_p12": at_equal("err_vs_fp64_p12"),
        "split_p12_speedup_vs_plain": split12["steps_per_s"] / plain["steps_per_s"],
        "err_sample": int(len(idx)),
    }


def energy_drift(vgm, workloads, args):
    """Relative total-energy drift |E(T) - E(0)| / |E(0)| over K RK4 steps of the
    two-Plummer collision: split p = 12 M2L with the learned closure (p = 4,
    trained on another seed of the collision, order-12 fallback for a quarter
    of the leaves), split p = 12 alone, and the plain FP32 p = 12 GPU FMM (FP64
    state and FP64 energy reductions in all), with each one's force error on
    the initial state against an FP64 direct sum."""
    import closure
    wl = args.drift_workload
    wdef = workloads.WORKLOADS[wl]
    x, v, masses = workloads.make(wl, seed=0)
    n = len(x)
    idx = np.sort(np.random.default_rng(1).choice(n, size=16384, replace=False))
    X, params, io, _ = train_closure(vgm, closure, workloads, wl, (1,))
    bound = closure.bound_for_fraction(params, io, X, 0.25)
    out = {"workload": wl, "n": n, "steps": args.drift_steps, "dt": wdef["dt"], "eps": wdef["eps"], "train_pairs": int(len(X))}
    d = None
    for name, mode, p, cl in (("split_closure_p4_fb12", "split", 4, (params, io, bound)), ("split_p12", "split", 12, None),
                              ("plain_fp32_p12", "f32", 12, None)):
        with vgm.Gravity(n, wdef["s"], p, mode, eps=wdef["eps"], G=1.0, mass=1.0 / n) as g:
            g.upload(x, v, masses)
            if cl is not None:
                g.set_closure(*cl)
            a = g.accelerations()[idx]
            if d is None:
                d = g.direct(idx)
            e0 = sum(g.energy())
            t0 = time.perf_counter()
            g.step(wdef["dt"], args.drift_steps)
            t1 = time.perf_counter()
            e1 = sum(g.energy())
            fb = g.counts()["fallback_leaves"]
        out[name] = {"relative_drift": abs(e1 - e0) / abs(e0), "E0": e0, "E1": e1, "steps_per_s": args.drift_steps / (t1 - t0),
                     "force_err": rel_err(a, d), "fallback_leaves": fb}
        print(f"# drift {name}: {out[name]['relative_drift']:.3e} at {out[name]['steps_per_s']:.2f} steps/s", file=sys.stderr, flush=True)
    return out


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


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(x, w, n, threads=None):
    """The reference's FMM compiled from its own headers (oracle/_ref) on the
    host's cores, on a bounded sample of one force evaluation of the same
    input.  Tree and calc_link run in full with the reference's own threading
    (the octree insertion is serial); P2M, M2M, P2P and L2L/L2P run sampled
    items on all threads; M2L runs on one thread (its loop in the reference is
    not race-free) and is counted at ideal scaling over the threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if not oracle.available("ref"):
        return None
    host = host_info()
    if threads is None:
        env = os.environ.get("OMP_NUM_THREADS")
        threads = min(int(env), host["nproc"]) if env and env.isdigit() and int(env) > 0 else host["nproc"]
    ref = oracle.Oracle("ref")
    per = 32
    secs, frac = ref.time_sample(x, w["s"], w["p"], 1.0 / n, 1.0, w["eps"], per, threads)
    names = ["tree", "lists", "p2m", "m2m", "m2l_one_thread", "p2p", "l2l_l2p"]
    t_eval = float(secs.sum() - secs[4] + secs[4] / threads)
    return {"value": 1.0 / (4 * t_eval), "unit": "steps/s", "cores": int(threads), "kind": "reference",
            "host": host,
            "sample": (f"one force evaluation of the same {n}-particle input: Octree build (serial, as the reference) and calc_link "
                       f"in full; P2M, M2M, P2P, L2L/L2P on {per * threads} evenly strided work items on {threads} threads; M2L on "
                       f"{per} items on one thread, divided by {threads} (ideal scaling: the reference's M2L loop races, "
                       f"fmm.hpp:279-281); sampled work fractions {', '.join(f'{f:.1e}' for f in frac[2:])}, scaled by exact work "
                       f"counts; 4 evaluations per RK4 step; extrapolated"),
            "seconds_per_eval": {k: float(s) for k, s in zip(names, secs)}}


if __name__ == "__main__":
    main()
