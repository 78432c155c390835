"""The reference configurations' own checks, at their sizes, on the device.

config 0: N = 16,384 Plummer, s = 64, p = 4, eps = 1e-3, FP64: ten RK4 steps
  follow the oracle's FP64 FMM trajectory (the reference's algorithm) to 1e-9
  of the displacement, and two steps stay within 1e-3 of the displacement of
  an FP64 direct-sum trajectory (the p = 4 truncation error).
config 1: N = 2^20 uniform, masses U[0.5, 1.5]/N, v = 0, eps = 1e-4, p = 8,
  FP32: after an RK4 step of dt = 1e-4 the tree and lists built from the new
  positions are bit-exact to the oracle's.
Energy: total energy of a 16k Plummer sphere under 20 split-mode RK4 steps
  drifts less than 1e-5 (relative)."""
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
