"""Parity of the HIP path (libvipp_gravity.so through its C ABI) with the
reference: golden vectors written by the reference's own code, and the CPU
oracle on the same seeded inputs.

Tolerances: trees, cell numbering, leaf lists and interaction lists are
compared bit for bit.  FP64 accelerations and expansions are compared with a
relative tolerance of 1e-11 of the largest value (summation order and libm
differ from the reference).  The FP32 path is compared with the FP64 oracle at
the FP32 accuracy it can reach (stated per test)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TREES = ["plummer16k_s64", "plummer4k_s5", "uniform2k_s1", "uniform20k_s64", "clump3k_s2"]


@pytest.fixture(scope="module")
def vg():
    import vipp_gravity
    return vipp_gravity


def gpu_tree(vg, x, s, p=4, precision="f64", eps=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    with vg.Gravity(len(x), s, p, precision, eps=eps) as g:
        g.upload(x)
        return g.tree(), g.lists()


def assert_tree_equal(t, ref):
    for f in ("parent", "level", "child_base", "pt_off", "pts"):
        assert np.array_equal(np.asarray(getattr(t, f)), np.asarray(getattr(ref, f) if hasattr(ref, f) else ref[f])), f
    bmin = ref.bmin if hasattr(ref, "bmin") else ref["bmin"]
    bmax = ref.bmax if hasattr(ref, "bmax") else ref["bmax"]
    assert np.array_equal(t.bmin, bmin) and np.array_equal(t.bmax, bmax)


def assert_lists_equal(l, ref):
    for f in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
        assert np.array_equal(getattr(l, f), getattr(ref, f) if hasattr(ref, f) else ref[f]), f


# ------------------------------------------------------------- tree / lists --
@pytest.mark.parametrize("name", TREES)
def test_tree_lists_golden(vg, golden, name):
    g = golden("tree_" + name)
    t, l = gpu_tree(vg, g["x"], int(g["s"]))
    assert_tree_equal(t, g)
    assert_lists_equal(l, g)


@pytest.mark.parametrize("x", [
    [[1, 1, 1], [-1, -1, -1], [1, -1, 1]],            # test_octree.cpp:113-128
    [[4, 4, 4], [-4, -4, -4], [2, 2, 2], [3, 3, 3]],  # test_octree.cpp:130-154
    [[0, 0, 0]],                                      # test_octree.cpp:105-111
])
def test_tree_kats(vg, orc, x):
    x = np.array(x, float)
    t, l = gpu_tree(vg, x, 1)
    rt, rl = orc.tree_and_lists(x, 1)
    assert_tree_equal(t, rt)
    assert_lists_equal(l, rl)


def test_tree_ties_on_split_planes(vg, orc):
    """Points exactly on the midpoints of every level go to the high side (octree.hpp:136,153-158)."""
    g = np.arange(-8, 9, dtype=float)
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = np.unique(np.concatenate([x, x[::7] * 0.5, x[::5] * 0.25]), axis=0)  # distinct points
    for s in (1, 3, 10):
        t, l = gpu_tree(vg, x, s)
        rt, rl = orc.tree_and_lists(x, s)
        assert_tree_equal(t, rt)
        assert_lists_equal(l, rl)


def test_tree_uniform_1m_s64(vg, orc):
    """The 2^20-particle uniform workload: codes, tree and lists bit-exact."""
    x = np.random.default_rng(1).random((1 << 20, 3))
    t, l = gpu_tree(vg, x, 64, precision="f32")
    rt, rl = orc.tree_and_lists(x, 64)
    assert_tree_equal(t, rt)
    assert_lists_equal(l, rl)


def test_tree_deep_cluster(vg, orc):
    """A cluster 3e-6 wide inside the unit cube needs 19 levels: the
    build that sorts the first 16 levels' digits must redo the sort in full."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.random((1000, 3)), 0.3 + 3e-6 * rng.random((400, 3))])
    t, l = gpu_tree(vg, x, 64)
    rt, rl = orc.tree_and_lists(x, 64)
    assert t.level.max() > 16
    assert_tree_equal(t, rt)
    assert_lists_equal(l, rl)


def test_tree_deeper_than_previous(vg, orc):
    """A context sorts the code digits of two levels beyond its previous tree.
    A shallow tree, then a deep one, then the shallow one again in the same
    context: each must equal the oracle's tree bit for bit (the deep one
    through the full re-sort)."""
    rng = np.random.default_rng(5)
    shallow = rng.random((1400, 3))
    deep = np.concatenate([rng.random((1000, 3)), 0.3 + 3e-6 * rng.random((400, 3))])
    with vg.Gravity(1400, 64, 4, "f64") as g:
        for x in (shallow, deep, shallow):
            g.upload(x)
            rt, rl = orc.tree_and_lists(x, 64)
            assert_tree_equal(g.tree(), rt)
            assert_lists_equal(g.lists(), rl)


def test_tree_device_levels(vg, orc):
    """From its second tree on, a context keeps the level sizes on the device,
    in a cell capacity of twice its previous tree's cells (at least 1024).
    A uniform tree (585 cells), then one with two dense regions (1721 cells,
    over that capacity: built again level by level), the same again and a
    moved copy (on the device), then the first again: each must equal the
    oracle's tree and lists bit for bit, and the forces of the last one the
    forces of a fresh context."""
    n = 1 << 16
    rng = np.random.default_rng(12)
    a = rng.random((n, 3))
    b = np.concatenate([rng.random((n // 2, 3)), 0.01 + 0.23 * rng.random((n // 4, 3)), 0.51 + 0.23 * rng.random((n // 4, 3))])
    c = b + 1e-4 * rng.standard_normal(b.shape)
    with vg.Gravity(n, 64, 4, "f64", eps=1e-3) as g:
        for x in (a, b, b, c, a):
            g.upload(x)
            rt, rl = orc.tree_and_lists(x, 64)
            assert_tree_equal(g.tree(), rt)
            assert_lists_equal(g.lists(), rl)
        acc = g.accelerations()
    with vg.Gravity(n, 64, 4, "f64", eps=1e-3) as g:
        g.upload(a)
        assert np.array_equal(g.accelerations(), acc)


def test_tree_plummer_16k(vg, orc):
    x, _, _ = orc.problem_initialize(16384, 0, 1e-3)
    t, l = gpu_tree(vg, x, 64)
    rt, rl = orc.tree_and_lists(x, 64)
    assert_tree_equal(t, rt)
    assert_lists_equal(l, rl)


# ----------------------------------------------------------------- FP64 FMM --
@pytest.mark.parametrize("p", [2, 4, 8])
def test_fmm_f64_golden(vg, golden, p):
    g = golden("fmm_plummer2k")
    x = g["x"]
    with vg.Gravity(len(x), int(g["s"]), p, "f64", eps=float(g["eps"])) as gr:
        gr.upload(x)
        a = gr.accelerations()
        ref = g[f"a_p{p}"]
        assert np.abs(a - ref).max() <= 1e-11 * np.abs(ref).max()
        if f"M_p{p}" in g:
            M, L = gr.expansions()
            t = gr.tree()
            live = np.array([(t.level[c] >= 2) and (t.child_base[c] >= 0 or t.pt_off[c + 1] > t.pt_off[c]) for c in range(t.ncells)])
            Mr, Lr = g[f"M_p{p}"], g[f"L_p{p}"]
            assert np.abs(M[live] - Mr[live]).max() <= 1e-12 * np.abs(Mr).max()
            assert np.abs(L[live] - Lr[live]).max() <= 1e-11 * np.abs(Lr[live]).max()


def test_fmm_f64_plummer_p12_vs_direct(vg, orc):
    x, _, _ = orc.problem_initialize(16384, 2, 1e-3)
    d = orc.direct(x, eps=1e-3)
    with vg.Gravity(len(x), 64, 12, "f64", eps=1e-3) as gr:
        gr.upload(x)
        a = gr.accelerations()
    err = np.linalg.norm(a - d, axis=1) / np.linalg.norm(d, axis=1)
    assert np.median(err) < 2e-6 and np.percentile(err, 99) < 1e-4, (np.median(err), np.percentile(err, 99))


def test_fmm_f64_vs_oracle_masses_and_G(vg, orc):
    rng = np.random.default_rng(5)
    x = rng.random((6000, 3))
    m = rng.uniform(0.5, 1.5, len(x)) / len(x)
    for far_g in (False, True):
        with vg.Gravity(len(x), 10, 6, "f64", eps=1e-4, G=2.5, ref_far_g=not far_g) as gr:
            gr.upload(x, masses=m)
            a = gr.accelerations()
        ref = orc.fmm(x, 10, 6, G=2.5, eps=1e-4, masses=m, far_g=far_g)
        assert np.abs(a - ref).max() <= 1e-11 * np.abs(ref).max()


def test_masses_and_G_closed_forms(vg):
    """The device's mass and G handling against what the mathematics fixes,
    independently of the oracle: the two-body direct sum and potential with
    distinct masses and G != 1; Newton's third law of the direct sum; the FMM
    exactly linear in G with G on the far field, affine with the reference's
    far field (G only in P2P, fmm.hpp:539), and equal (to rounding) on a
    uniform mass array and the scalar-mass path."""
    G, eps, m0, m1 = 2.5, 0.1, 0.3, 1.7
    x = np.array([[0.1, -0.2, 0.3], [0.45, 0.05, -0.15]])
    d = x[1] - x[0]
    r2 = d @ d + eps * eps
    with vg.Gravity(2, 1, 4, "f64", eps=eps, G=G) as gr:
        gr.upload(x, masses=np.array([m0, m1]))
        a, pot = gr.direct([0, 1], potential=True)
    assert np.allclose(a, [G * m1 * d / r2 ** 1.5, -G * m0 * d / r2 ** 1.5], rtol=1e-14, atol=0)
    assert np.allclose(pot, [m1 / np.sqrt(r2), m0 / np.sqrt(r2)], rtol=1e-14, atol=0)
    rng = np.random.default_rng(11)
    x = rng.random((4000, 3))
    m = rng.uniform(0.1, 3.0, len(x)) / len(x)
    with vg.Gravity(len(x), 10, 8, "f64", eps=1e-3, G=1.9) as gr:
        gr.upload(x, masses=m)
        a = gr.direct(np.arange(len(x)))
    assert np.abs((m[:, None] * a).sum(0)).max() <= 1e-12 * (m[:, None] * np.abs(a)).sum()

    def fmm(G, ref_far_g, masses=m, mass=None):
        with vg.Gravity(len(x), 10, 8, "f64", eps=1e-3, G=G, mass=mass, ref_far_g=ref_far_g) as gr:
            gr.upload(x, masses=masses)
            return gr.accelerations()
    one = fmm(1.0, False)
    assert np.abs(fmm(2.5, False) - 2.5 * one).max() <= 1e-14 * 2.5 * np.abs(one).max()
    far = fmm(0.0, True)  # the reference's far field alone
    assert np.abs(fmm(2.5, True) - (far + 2.5 * (one - far))).max() <= 1e-13 * np.abs(one).max()
    # the scalar-mass path factors the mass out of the pair sums: equal to rounding
    u = fmm(1.0, False, masses=np.full(len(x), 1.0 / len(x)))
    assert np.abs(u - fmm(1.0, False, masses=None, mass=1.0 / len(x))).max() <= 1e-13 * np.abs(u).max()


def test_rk4_step_f64_golden(vg, golden):
    g = golden("rk4_1k")
    x, v = g["x"], g["v"]
    with vg.Gravity(len(x), int(g["s"]), int(g["p"]), "f64", eps=float(g["eps"])) as gr:
        gr.upload(x, v)
        gr.step(float(g["dt"]), 1)
        x1, v1 = gr.download()
    assert np.abs(x1 - g["x_fmm"]).max() <= 1e-11 * np.abs(g["x_fmm"] - x).max()
    assert np.abs(v1 - g["v_fmm"]).max() <= 1e-11 * np.abs(g["v_fmm"] - v).max()


def test_rk4_f64_two_body_closed_form(vg):
    """test_solver_ref_naive.cpp:47-79 through the FMM path (two points: one leaf, pure P2P)."""
    x = np.array([[0, 0, 0], [10, 0, 0]], float)
    eps = 1e-4
    with vg.Gravity(2, 1, 4, "f64", eps=eps, mass=1.0) as gr:
        gr.upload(x)
        gr.step(1.0, 1)
        x1, v1 = gr.download()
    d = 10.0
    k1 = d / (d * d + eps * eps) ** 1.5
    l2 = k1 / 2
    k3 = (d - l2) / ((d - l2) ** 2 + eps * eps) ** 1.5
    k4 = (d - 2 * l2) / ((d - 2 * l2) ** 2 + eps * eps) ** 1.5
    xd = (2 * l2 + 2 * l2 + k3) / 6
    vd = (k1 + 2 * k1 + 2 * k3 + k4) / 6
    np.testing.assert_allclose(x1, [[xd, 0, 0], [d - xd, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(v1, [[vd, 0, 0], [-vd, 0, 0]], atol=1e-15)


# ----------------------------------------------------------------- FP32 FMM --
@pytest.mark.parametrize("p", [4, 8])
def test_fmm_f32_vs_oracle(vg, orc, p):
    """FP32 P2P/M2L against the FP64 oracle of the same algorithm: median 1e-6, 99th percentile 1e-4."""
    x, _, _ = orc.problem_initialize(20000, 4, 1e-3)
    ref = orc.fmm(x, 64, p, eps=1e-3)
    with vg.Gravity(len(x), 64, p, "f32", eps=1e-3) as gr:
        gr.upload(x)
        a = gr.accelerations()
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) < 1e-6 and np.percentile(err, 99) < 1e-4, (np.median(err), np.percentile(err, 99))


@pytest.mark.parametrize("p", [4, 8, 12])
def test_fmm_split_vs_oracle(vg, orc, p):
    """BF16 hi/lo split M2L on MFMA (FP32 elsewhere) against the FP64 oracle: median 2e-6, 99th percentile 1e-4."""
    x, _, _ = orc.problem_initialize(20000, 4, 1e-3)
    ref = orc.fmm(x, 64, p, eps=1e-3)
    with vg.Gravity(len(x), 64, p, "split", eps=1e-3) as gr:
        gr.upload(x)
        a = gr.accelerations()
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(err) < 2e-6 and np.percentile(err, 99) < 1e-4, (np.median(err), np.percentile(err, 99))


def test_large_leaves(vg, orc):
    """s = 256 makes leaves of up to floor(sqrt(8) 256) = 724 points, so the
    P2P takes its batched path (256 targets at a time) for the largest ones:
    FP64 against the oracle (1e-11 of max |a|), FP32 against FP64 (median
    1e-6)."""
    x, _, _ = orc.problem_initialize(30000, 11, 1e-3)
    ref = orc.fmm(x, 256, 4, eps=1e-3)
    out = {}
    for mode in ("f64", "f32"):
        with vg.Gravity(len(x), 256, 4, mode, eps=1e-3) as gr:
            gr.upload(x)
            out[mode] = gr.accelerations()
            if mode == "f64":
                t = gr.tree()
    cnt = np.diff(np.asarray(t.pt_off))
    par = np.asarray(t.parent)
    leaf = np.ones(len(cnt), bool)
    leaf[par[par >= 0]] = False
    assert cnt[leaf].max() > 384, cnt[leaf].max()  # some leaf takes the batched path
    assert np.abs(out["f64"] - ref).max() <= 1e-11 * np.abs(ref).max()
    rel = np.linalg.norm(out["f32"] - out["f64"], axis=1) / np.linalg.norm(out["f64"], axis=1)
    assert np.median(rel) < 1e-6, np.median(rel)


@pytest.mark.parametrize("mode", ["f32", "split"])
def test_uniform_mass_path(vg, orc, mode):
    """With no per-particle masses the P2P kernels leave the mass out of the
    pair terms and scale the sums once; uploading the same mass for every
    particle takes the general path.  The two agree to FP32 rounding: median
    1e-6 and at most 1e-4 of each |a| (a missing or doubled mass would be off
    by orders of magnitude)."""
    x, _, _ = orc.problem_initialize(20000, 9, 1e-3)
    out = []
    for masses in (None, np.full(len(x), 1.0 / len(x))):
        with vg.Gravity(len(x), 64, 8, mode, eps=1e-3) as gr:
            gr.upload(x, masses=masses)
            out.append(gr.accelerations())
    rel = np.linalg.norm(out[0] - out[1], axis=1) / np.linalg.norm(out[1], axis=1)
    assert rel.max() < 1e-4 and np.median(rel) < 1e-6, (rel.max(), np.median(rel))


def test_split_m2l_locals_match_plain(vg, orc):
    """The local expansions from the MFMA path agree with the plain FP32 path to the split's 2^-16 precision."""
    x, _, _ = orc.problem_initialize(30000, 5, 1e-3)
    Ls = {}
    for mode in ("f32", "split"):
        with vg.Gravity(len(x), 64, 8, mode, eps=1e-3) as gr:
            gr.upload(x)
            gr.accelerations(copy=False)
            Ls[mode] = gr.expansions()[1]
    a, b = Ls["f32"], Ls["split"]
    scale = np.abs(a).max(axis=(1, 2), keepdims=True) + 1e-300
    assert (np.abs(a - b) / scale).max() < 1e-4


@pytest.mark.parametrize("size,p", [("small", 3), ("small", 5), ("small", 8), ("small", 9), ("small", 11), ("small", 12),
                                    ("large", 3), ("large", 8), ("large", 12)])
def test_plain_m2l_locals_match_f64(vg, orc, size, p):
    """The FP32 VALU M2L (1 to 6 row groups of 32; the even ones on two thread
    groups) gives the FP64 path's local expansions to FP32 accuracy, cell by
    cell.  "small": a clustered input whose blocks hold empty octants, few
    parents, so blocks take half the targets; "large": 2^21 uniform points at
    s = 16, 37,448 parents, so blocks take the full target count."""
    if size == "small":
        x, _, _ = orc.problem_initialize(40000, 9, 1e-3)
        x[:20000] *= 0.05  # a dense core inside a sparse halo
        s = 32
    else:
        x = np.random.default_rng(p).random((1 << 21, 3))
        s = 16
    Ls = {}
    for mode in ("f64", "f32"):
        with vg.Gravity(len(x), s, p, mode, eps=1e-3) as gr:
            gr.upload(x)
            gr.accelerations(copy=False)
            Ls[mode] = gr.expansions()[1]
            if size == "large":
                assert gr.counts()["cells"] > 8 * 37000
    a, b = Ls["f64"], Ls["f32"]
    scale = np.abs(a).max(axis=(1, 2), keepdims=True) + 1e-300
    rel = np.abs(a - b) / scale
    # FP32 sums of 189 K products per coefficient: about 1e-7 typically; the
    # maximum over the 2.4e7 coefficients of the large input reaches 2e-5
    assert np.median(rel[np.abs(a) > 1e-6 * scale]) < 1e-6
    assert rel.max() < (2e-5 if size == "small" else 5e-5)


def test_energy_f64(vg, orc):
    x, v, _ = orc.problem_initialize(5000, 6, 1e-3)
    _, pot = orc.direct(x, eps=1e-3, potential=True)
    m = 1.0 / len(x)
    with vg.Gravity(len(x), 10, 10, "f64", eps=1e-3) as gr:
        gr.upload(x, v)
        ke, pe = gr.energy()
    assert abs(ke - 0.5 * m * (v ** 2).sum()) <= 1e-14 * ke
    pe_ref = -0.5 * m * pot.sum()
    assert abs(pe - pe_ref) <= 1e-6 * abs(pe_ref), (pe, pe_ref)


@pytest.mark.parametrize("eps,with_masses", [(1e-3, False), (0.0, True)])
def test_direct_sum_vs_oracle(vg, orc, eps, with_masses):
    """Device FP64 direct sum (solver_ref_naive.hpp:103-119) against the CPU
    oracle: relative 1e-13 of |a| (sqrt/division rounding and summation order
    differ); the sample includes duplicates-free indices at both ends."""
    rng = np.random.default_rng(7)
    n = 20000
    x = rng.standard_normal((n, 3))
    masses = rng.uniform(0.5, 1.5, n) / n if with_masses else None
    idx = np.concatenate([[0, n - 1], rng.choice(np.arange(1, n - 1), 300, replace=False)])
    a_ref, pot_ref = orc.direct_subset(x, idx, m=1.0 / n, G=0.7, eps=eps, masses=masses, potential=True)
    with vg.Gravity(n, 64, 4, "f32", eps=eps, G=0.7) as gr:
        gr.upload(x, None, masses)
        a, pot = gr.direct(idx, potential=True)
        a2 = gr.direct(idx[:5])
    rel = np.linalg.norm(a - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
    assert rel.max() < 1e-13, rel.max()
    assert np.abs(pot - pot_ref).max() < 1e-13 * np.abs(pot_ref).max()
    assert np.array_equal(a2, a[:5])  # same source slicing for both calls: same bits


def test_direct_sum_rejects_bad_index(vg):
    with vg.Gravity(100, 4, 4, "f32") as gr:
        gr.upload(np.random.default_rng(0).random((100, 3)))
        with pytest.raises(vg.GravityError) as e:
            gr.direct([0, 100])
        assert e.value.code == vg.E_INVALID


# -------------------------------------------------------------- error paths --
def test_non_finite_input_rejected(vg):
    x = np.random.default_rng(0).random((100, 3))
    x[17, 1] = np.nan
    with vg.Gravity(len(x), 1, 4, "f32") as gr:
        gr.upload(x)
        with pytest.raises(vg.GravityError) as e:
            gr.accelerations()
        assert e.value.code == vg.E_INPUT


def test_coincident_points_depth_error(vg):
    x = np.concatenate([np.zeros((10, 3)), [[1.0, 1.0, 1.0]]])
    with vg.Gravity(len(x), 1, 4, "f32") as gr:  # threshold 2 < 10 coincident points
        gr.upload(x)
        with pytest.raises(vg.GravityError) as e:
            gr.accelerations()
        assert e.value.code == vg.E_DEPTH


def test_depth_error_then_recovery(vg, orc):
    """A context whose previous tree exists builds the next from the sorted
    digits of two more levels; coincident points then reach the 21-level
    limit through the full re-sort and fail with E_DEPTH.  The same context
    must then build a good input's tree and lists exactly as the oracle."""
    rng = np.random.default_rng(21)
    good = rng.random((11, 3))
    bad = np.concatenate([np.zeros((10, 3)), [[1.0, 1.0, 1.0]]])
    with vg.Gravity(11, 1, 4, "f32") as gr:  # threshold 2
        for x, ok in ((good, True), (bad, False), (good, True)):
            gr.upload(x)
            if ok:
                rt, rl = orc.tree_and_lists(x, 1)
                assert_tree_equal(gr.tree(), rt)
                assert_lists_equal(gr.lists(), rl)
            else:
                with pytest.raises(vg.GravityError) as e:
                    gr.accelerations()
                assert e.value.code == vg.E_DEPTH


def test_nonfinite_then_recovery(vg, orc):
    """Split mode, second tree on: a non-finite coordinate is caught by the
    device-driven build (E_INPUT) before any force work is queued on the
    second stream; the same context then gives the forces of a fresh one,
    bit for bit."""
    x, _, _ = orc.problem_initialize(20000, 3, 1e-3)
    bad = x.copy()
    bad[123, 1] = np.nan
    with vg.Gravity(len(x), 64, 8, "split", eps=1e-3) as gr:
        gr.upload(x)
        gr.accelerations()
        gr.upload(bad)
        with pytest.raises(vg.GravityError) as e:
            gr.accelerations()
        assert e.value.code == vg.E_INPUT
        gr.upload(x)
        a = gr.accelerations()
    with vg.Gravity(len(x), 64, 8, "split", eps=1e-3) as gr:
        gr.upload(x)
        assert np.array_equal(gr.accelerations(), a)


@pytest.mark.parametrize("mode", ["f64", "split"])
def test_failed_step_restores_state(vg, orc, mode):
    """A stage of an RK4 step whose Euler state overflows (x0 + v dt/2 = inf)
    fails with E_INPUT after the first stage's tree, lists and forces were
    built.  x and v must be put back to the state the failed step started
    from, and every query after it must describe that state: tree, lists and
    work counts equal to a fresh context holding it, never the stale tree of
    the failed stage."""
    x, v, _ = orc.problem_initialize(3000, 5, 1e-3)
    with vg.Gravity(len(x), 8, 4, mode, eps=1e-3) as gr:
        gr.upload(x, v)
        gr.step(1e-3, 1)  # one good step: the context holds a tree
        x1, v1 = gr.download()
        vbad = v1.copy()
        vbad[7] = [1e300, 0.0, 0.0]
        gr.upload(x1, vbad)
        with pytest.raises(vg.GravityError) as e:
            gr.step(1e10, 1)  # stage 1 position 1e300 * 5e9 overflows
        assert e.value.code == vg.E_INPUT
        x2, v2 = gr.download()
        assert np.array_equal(x2, x1) and np.array_equal(v2, vbad)
        t, l = gr.tree(), gr.lists()
        counts = gr.counts()
    rt, rl = orc.tree_and_lists(x1, 8)
    assert_tree_equal(t, rt)
    assert_lists_equal(l, rl)
    with vg.Gravity(len(x), 8, 4, mode, eps=1e-3) as gr:
        gr.upload(x1, vbad)
        assert gr.counts() == counts


def test_upload_checks_sizes(vg):
    with vg.Gravity(100, 4, 4, "f32") as gr:
        x = np.random.default_rng(0).random((100, 3))
        with pytest.raises(ValueError):
            gr.upload(x, masses=np.ones(99))
        with pytest.raises(ValueError):
            gr.upload(x[:99])


def test_two_streams_equal_one(vg, orc):
    """Split mode runs the lists and P2P on a second stream unless per-phase
    timing is on; both schedules must give the same forces bit for bit, over
    several evaluations and RK4 steps of one context."""
    x, v, _ = orc.problem_initialize(50000, 4, 1e-3)
    runs = []
    for timing in (False, True):
        with vg.Gravity(len(x), 64, 12, "split", eps=1e-3) as gr:
            gr.set_timing(timing)
            gr.upload(x, v)
            a0 = gr.accelerations()
            gr.step(1e-3, 2)
            runs.append((a0, gr.accelerations(), gr.download()))
    (a0, a1, (x1, v1)), (b0, b1, (y1, w1)) = runs
    assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    assert np.array_equal(x1, y1) and np.array_equal(v1, w1)


def test_deterministic(vg, orc):
    x, _, _ = orc.problem_initialize(30000, 8, 1e-3)
    outs = []
    for _ in range(2):
        with vg.Gravity(len(x), 64, 8, "f32", eps=1e-3) as gr:
            gr.upload(x)
            outs.append(gr.accelerations())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("mode", ["f64", "f32", "split"])
@pytest.mark.parametrize("n", [1, 2, 5, 100])
def test_tiny_inputs(vg, orc, mode, n):
    """Fewer points than a leaf holds: the root is the only leaf, the far field
    is empty and the forces are the direct sum.  One point feels nothing.  A
    step and the energy work at these sizes too."""
    x = np.random.default_rng(n).random((n, 3))
    v = np.zeros_like(x)
    with vg.Gravity(n, 64, 4, mode, eps=1e-2) as gr:
        gr.upload(x, v)
        a = gr.accelerations()
        d = gr.direct(np.arange(n))
        gr.step(1e-3, 2)
        x1, v1 = gr.download()
        e = gr.energy()
        assert gr.counts()["cells"] == 1
    ref = orc.direct(x, eps=1e-2)
    tol = 1e-12 if mode == "f64" else 2e-6
    assert np.allclose(d, ref, rtol=1e-13, atol=0)
    assert np.abs(a - ref).max() <= tol * max(np.abs(ref).max(), 1e-300)
    if n == 1:
        assert not a.any() and np.array_equal(x1, x)
    assert np.isfinite(x1).all() and np.isfinite(v1).all() and np.isfinite(e).all()


@pytest.mark.parametrize("shape", ["plane", "line"])
def test_degenerate_geometry(vg, orc, shape):
    """Points in a plane (z = 0) or on a line: the root box is 2e-5 thick in
    the flat axes (the reference's padding), so the cells are extremely flat
    and the offsets of the operators nearly degenerate.  Tree and lists stay
    bit-exact, FP64 forces match the oracle's FMM to 1e-11, and split mode
    matches the FP64 direct sum."""
    rng = np.random.default_rng(31)
    n = 6000
    x = np.zeros((n, 3))
    x[:, 0] = rng.normal(size=n)
    if shape == "plane":
        x[:, 1] = rng.normal(size=n)
    with vg.Gravity(n, 16, 6, "f64", eps=1e-3) as gr:
        gr.upload(x)
        a64 = gr.accelerations()
        t, l = gr.tree(), gr.lists()
    rt, rl = orc.tree_and_lists(x, 16)
    for f in ("parent", "level", "child_base", "pt_off", "pts"):
        assert np.array_equal(getattr(t, f), getattr(rt, f)), f
    for f in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
        assert np.array_equal(getattr(l, f), getattr(rl, f)), f
    ref = orc.fmm(x, 16, 6, eps=1e-3)
    assert np.abs(a64 - ref).max() <= 1e-11 * np.abs(ref).max()
    with vg.Gravity(n, 16, 6, "split", eps=1e-3) as gr:
        gr.upload(x)
        a = gr.accelerations()
        d = gr.direct(np.arange(0, n, 7))
    rel = np.linalg.norm(a[::7] - d, axis=1) / np.linalg.norm(d, axis=1)
    assert np.median(rel) < 1e-3, np.median(rel)


@pytest.mark.parametrize("mode", ["f64", "split"])
def test_stage_boxes_equal_fresh_boxes(vg, orc, mode):
    """Inside one step call each tree takes its root box from the partial
    boxes the previous RK4 stage left; a call starts from a box computed from
    x.  Minimum and maximum are exact, so one call of two steps and two calls
    of one step give the same state bit for bit."""
    x, v, _ = orc.problem_initialize(20000, 12, 1e-3)
    outs = []
    for calls in ((2,), (1, 1)):
        with vg.Gravity(len(x), 64, 6, mode, eps=1e-3) as gr:
            gr.upload(x, v)
            for k in calls:
                gr.step(1e-3, k)
            outs.append(gr.download())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


# ------------------------------------------------------------- every order --
@pytest.mark.parametrize("p", list(range(1, 13)))
def test_every_order(vg, orc, p):
    """Each order p is its own instantiation of L2P and of the split M2L: FP64
    against the oracle (1e-11 of max |a|), FP32 and split against the FP64
    result of the same tree (median 1e-5: FP32 rounding, BF16 hi/lo split)."""
    x, _, _ = orc.problem_initialize(2000, 7, 1e-3)
    ref = orc.fmm(x, 16, p, eps=1e-3)
    out = {}
    for mode in ("f64", "f32", "split"):
        with vg.Gravity(len(x), 16, p, mode, eps=1e-3) as gr:
            gr.upload(x)
            out[mode] = gr.accelerations()
    assert np.abs(out["f64"] - ref).max() <= 1e-11 * np.abs(ref).max()
    for mode in ("f32", "split"):
        rel = np.linalg.norm(out[mode] - out["f64"], axis=1) / np.linalg.norm(out["f64"], axis=1)
        assert np.median(rel) < 1e-5, (mode, np.median(rel))
