"""bench.py's contract pieces that need no GPU: it refuses more than one GPU
before touching one, and its per-phase roofline arithmetic (work per unit,
fractions of peak, the step composite) follows the stated definitions."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("p", [1, 4, 8, 12])
def test_m2l_flop_per_pair(p):
    # a K x K real operator times a K vector, K = (p + 1)^2: 2 K^2 FLOP
    assert bench.m2l_flop_per_pair(p) == 2 * ((p + 1) ** 2) ** 2


def test_phase_rooflines_arithmetic():
    cnt = {"cells": 8 * 1000 + 1, "list_entries": 2_000_000, "p2p_pairs": 1e10, "m2l_live_pairs": 1e7}
    per_eval = {"tree": 0.5, "lists": 0.1, "p2p": 5.0, "m2l": 3.0, "upward": 0.4, "downward": 0.5, "rk4": 0.15, "operators": 0.2}
    n, p, ms_step = 1 << 22, 12, 40.0
    out = bench.phase_rooflines(per_eval, cnt, n, p, "split", ms_step)
    # tree: 136 B per particle against HBM
    assert out["tree"]["work"] == pytest.approx(136 * n / 1e9)
    assert out["tree"]["achieved"] == pytest.approx(136 * n / 1e9 / 0.5e-3)
    assert out["tree"]["frac"] == pytest.approx(out["tree"]["achieved"] / 8000.0)
    # P2P: 20 FLOP per pair against the FP32 vector peak, also stated against the non-packed rate
    assert out["p2p"]["work"] == pytest.approx(20 * 1e10 / 1e12)
    assert out["p2p"]["frac"] == pytest.approx(0.2 / 5e-3 / 157.3)
    assert out["p2p"]["frac_of_scalar_fp32"] == pytest.approx(0.2 / 5e-3 / 78.6)
    # M2L in split mode: 2 (p+1)^4 per live pair against dense BF16
    assert out["m2l"]["bound"] == "mfma" and out["m2l"]["peak"] == 2500.0
    assert out["m2l"]["work"] == pytest.approx(2 * 169 ** 2 * 1e7 / 1e12)
    # M2M / L2L: one operator product per child-parent pair, 8 per internal cell
    assert out["upward"]["work"] == pytest.approx(2 * 169 ** 2 * 8 * 1000 / 1e12)
    # composite: 4 evaluations per step, each phase's work at its peak, over the measured step
    ideal = 4 * sum(r["work"] / r["peak"] * 1e3 for k, r in out.items() if k != "composite" and r["peak"] and r["work"] > 0)
    assert out["composite"]["ideal_ms_per_step"] == pytest.approx(ideal)
    assert out["composite"]["frac"] == pytest.approx(ideal / ms_step)
    # a phase with no work unit reports its time only
    assert out["operators"]["frac"] is None and out["operators"]["ms_per_eval"] == 0.2


def test_phase_rooflines_plain_and_f64_peaks():
    cnt = {"cells": 9, "list_entries": 10, "p2p_pairs": 1e6, "m2l_live_pairs": 1e3}
    per_eval = {"p2p": 1.0, "m2l": 1.0}
    f32 = bench.phase_rooflines(per_eval, cnt, 1000, 4, "f32", 10.0)
    f64 = bench.phase_rooflines(per_eval, cnt, 1000, 4, "f64", 10.0)
    assert f32["m2l"]["bound"] == "valu" and f32["m2l"]["peak"] == 157.3
    assert f64["m2l"]["peak"] == 78.6 and f64["p2p"]["peak"] == 78.6 and f64["p2p"]["frac_of_scalar_fp32"] is None


def test_order_sweep_spec():
    assert bench.parse_sweep("split:8,10") == [("split", 8), ("split", 10)]
    assert bench.parse_sweep("split:8;plain:10,12") == [("split", 8), ("plain", 10), ("plain", 12)]
    with pytest.raises(ValueError):
        bench.parse_sweep("bf16:8")
