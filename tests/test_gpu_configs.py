"""The reference configurations' own checks, at their sizes, on the device.

config 0: N = 16,384 Plummer, s = 64, p = 4, eps = 1e-3, FP64: ten RK4 steps
  follow the oracle's FP64 FMM trajectory (the reference's algorithm) to 1e-9
  of the displacement, and two steps stay within 1e-3 of the displacement of
  an FP64 direct-sum trajectory (the p = 4 truncation error).
config 1: N = 2^20 uniform, masses U[0.5, 1.5]/N, v = 0, eps = 1e-4, p = 8,
  FP32: after an RK4 step of dt = 1e-4 the tree and lists built from the new
  positions are bit-exact to the oracle's.
config 3: N = 2^22 mixture of 32 Plummer spheres (depth 11), s = 64: codes,
  tree and lists bit-exact to the oracle's on the initial positions and again
  after a split-mode p = 4 RK4 step, in one context.
config 4: the two-Plummer collision.  At its full N = 2^24, tree and lists
  after a split-mode p = 12 step are bit-exact to the oracle's; at 2^16 of the
  same geometry the split-mode p = 12 forces agree with the oracle's FP64
  p = 12 FMM (median 1e-5, the BF16 hi/lo M2L) and the FP64 direct sum
  (median 1e-5), and the total energy drifts less than 1e-5 over 10 steps.
Energy: total energy of a 16k Plummer sphere under 20 split-mode RK4 steps
  drifts less than 1e-5 (relative).
Scale: one split-mode p = 8 evaluation of 2^26 uniform points (16 times the
  headline size) against the FP64 direct sum on a sample."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vg():
    import vipp_gravity
    return vipp_gravity


@pytest.fixture(scope="module")
def workloads():
    import workloads
    return workloads


def test_config0_fmm_trajectory_f64(vg, orc, workloads):
    x, v, _ = workloads.make("plummer16k_p4")
    x0 = x.copy()
    with vg.Gravity(len(x), 64, 4, "f64", eps=1e-3) as g:
        g.upload(x, v)
        g.step(1e-3, 10)
        xg, vg_ = g.download()
    xr, vr = x.copy(), v.copy()
    for _ in range(10):
        xr, vr = orc.rk4_fmm(xr, vr, 1e-3, 64, 4, eps=1e-3)
    disp = np.abs(xr - x0).max()
    assert np.abs(xg - xr).max() <= 1e-9 * disp, (np.abs(xg - xr).max(), disp)
    assert np.abs(vg_ - vr).max() <= 1e-9 * np.abs(vr - v).max()


def test_config0_against_direct_sum(vg, orc, workloads):
    x, v, _ = workloads.make("plummer16k_p4")
    with vg.Gravity(len(x), 64, 4, "f64", eps=1e-3) as g:
        g.upload(x, v)
        g.step(1e-3, 2)
        xg, _ = g.download()
    xd, vd = x.copy(), v.copy()
    for _ in range(2):
        xd, vd = orc.rk4_direct(xd, vd, 1e-3, eps=1e-3)
    # the FMM's force error moves positions by ~|da| dt^2; against the full displacement
    # that is far below the 1e-3 bound, which an indexing or sign mistake would break
    disp = np.linalg.norm(xd - x, axis=1)
    err = np.linalg.norm(xg - xd, axis=1)
    assert np.median(err / disp) < 1e-3, np.median(err / disp)


def test_config1_trees_after_a_step(vg, orc, workloads):
    x, v, m = workloads.make("uniform1m_p8")
    with vg.Gravity(len(x), 64, 8, "f32", eps=1e-4) as g:
        g.upload(x, v, m)
        g.step(1e-4, 1)
        x1, _ = g.download()
    assert not np.array_equal(x1, x)
    with vg.Gravity(len(x1), 64, 8, "f32", eps=1e-4) as g:
        g.upload(x1, None, m)
        t, l = g.tree(), g.lists()
    rt, rl = orc.tree_and_lists(x1, 64)
    for f in ("parent", "level", "child_base", "pt_off", "pts"):
        assert np.array_equal(getattr(t, f), getattr(rt, f)), f
    assert np.array_equal(t.bmin, rt.bmin) and np.array_equal(t.bmax, rt.bmax)
    for f in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
        assert np.array_equal(getattr(l, f), getattr(rl, f)), f


def test_energy_conservation_split(vg, workloads):
    x, v, _ = workloads.make("plummer16k_p4")
    with vg.Gravity(len(x), 64, 8, "split", eps=1e-3) as g:
        g.upload(x, v)
        e0 = sum(g.energy())
        g.step(1e-3, 20)
        e1 = sum(g.energy())
    drift = abs(e1 - e0) / abs(e0)
    print(f"relative energy drift over 20 steps: {drift:.3e}")
    assert drift < 1e-5, drift


def _assert_same_tree(t, l, rt, rl):
    for f in ("parent", "level", "child_base", "pt_off", "pts"):
        assert np.array_equal(getattr(t, f), getattr(rt, f)), f
    assert np.array_equal(t.bmin, rt.bmin) and np.array_equal(t.bmax, rt.bmax)
    for f in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
        assert np.array_equal(getattr(l, f), getattr(rl, f)), f


def test_config3_mixture_trees_bit_exact(vg, orc, workloads):
    x, v, _ = workloads.make("mixture4m_p4")
    with vg.Gravity(len(x), 64, 4, "split", eps=1e-3) as g:
        g.upload(x, v)
        t, l = g.tree(), g.lists()
        assert t.level.max() >= 10
        rt, rl = orc.tree_and_lists(x, 64)
        _assert_same_tree(t, l, rt, rl)
        g.step(5e-4, 1)
        x1, _ = g.download()
        assert not np.array_equal(x1, x)
        t, l = g.tree(), g.lists()  # rebuilt at the stepped positions
    rt, rl = orc.tree_and_lists(x1, 64)
    _assert_same_tree(t, l, rt, rl)


def test_config4_collision_trees_after_split_step(vg, orc, workloads):
    x, v, _ = workloads.make("collision16m")
    with vg.Gravity(len(x), 64, 12, "split", eps=1e-3) as g:
        g.upload(x, v)
        g.step(1e-3, 1)
        x1, _ = g.download()
        t, l = g.tree(), g.lists()
    del x, v
    rt, rl = orc.tree_and_lists(x1, 64)
    _assert_same_tree(t, l, rt, rl)


def test_config4_collision_split_vs_oracle(vg, orc, workloads):
    x, v = workloads.two_plummer(1 << 16, seed=0)
    ref = orc.fmm(x, 64, 12, eps=1e-3)
    sample = np.random.default_rng(1).choice(len(x), 4096, replace=False)
    with vg.Gravity(len(x), 64, 12, "split", eps=1e-3) as g:
        g.upload(x, v)
        a = g.accelerations()
        d = g.direct(sample)
        e0 = sum(g.energy())
        g.step(1e-3, 10)
        e1 = sum(g.energy())
    rel = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-4, (np.median(rel), np.percentile(rel, 99))
    rel_d = np.linalg.norm(a[sample] - d, axis=1) / np.linalg.norm(d, axis=1)
    assert np.median(rel_d) < 1e-5, np.median(rel_d)
    drift = abs(e1 - e0) / abs(e0)
    print(f"collision 2^16: median vs FP64 FMM {np.median(rel):.2e}, vs direct {np.median(rel_d):.2e}, drift {drift:.2e}")
    assert drift < 1e-5, drift


def test_large_n_uniform_split(vg):
    """2^26 points, 16 times the headline size (index arithmetic of the sort,
    the tree, the lists, P2P work items and the particle arrays at that scale):
    one split-mode evaluation against the FP64 direct sum on a sample."""
    n = 1 << 26
    x = np.random.default_rng(21).random((n, 3))
    with vg.Gravity(n, 64, 8, "split", eps=1e-4, G=1.0, mass=1.0 / n) as g:
        g.upload(x, np.zeros_like(x))
        a = g.accelerations()
        idx = np.sort(np.random.default_rng(22).choice(n, 2048, replace=False))
        d = g.direct(idx)
        info = g.counts()
    rel = np.linalg.norm(a[idx] - d, axis=1) / np.linalg.norm(d, axis=1)
    print(f"2^26 uniform: cells {info['cells']}, leaves {info['leaves']}, depth {info['depth']}, median rel {np.median(rel):.2e}")
    assert np.isfinite(a).all()
    assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-3
