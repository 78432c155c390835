"""Pin the CPU oracle against the golden vectors written by the reference's
own code (tests/golden/*.npz, made by tests/golden/make_golden.py) and against
invariants that the mathematics fixes."""
import numpy as np
import pytest

TREES = ["plummer16k_s64", "plummer4k_s5", "uniform2k_s1", "uniform20k_s64", "clump3k_s2"]


@pytest.mark.parametrize("name", TREES)
def test_tree_and_lists_bit_exact(orc, golden, name):
    g = golden("tree_" + name)
    t, lists = orc.tree_and_lists(g["x"], int(g["s"]))
    for f in ("parent", "level", "child_base", "pt_off", "pts"):
        assert np.array_equal(getattr(t, f), g[f]), f
    assert np.array_equal(t.bmin, g["bmin"]) and np.array_equal(t.bmax, g["bmax"])
    for f in ("ap_off", "ap", "ni_off", "ni", "nb_off", "nb"):
        assert np.array_equal(getattr(lists, f), g[f]), f


@pytest.mark.parametrize("p", [2, 4, 8])
def test_fmm_accelerations(orc, golden, p):
    g = golden("fmm_plummer2k")
    out = orc.fmm(g["x"], int(g["s"]), p, eps=float(g["eps"]), diagnostics=True)
    ref = g[f"a_p{p}"]
    scale = np.abs(ref).max()
    assert np.abs(out["a"] - ref).max() <= 1e-13 * scale
    if f"M_p{p}" in g:
        assert np.abs(out["M"] - g[f"M_p{p}"]).max() <= 1e-14 * np.abs(g[f"M_p{p}"]).max()
        assert np.abs(out["L"] - g[f"L_p{p}"]).max() <= 1e-13 * np.abs(g[f"L_p{p}"]).max()


def test_direct_sum_bitwise(orc, golden):
    g = golden("fmm_plummer2k")
    a = orc.direct(g["x"], eps=float(g["eps"]))
    assert np.array_equal(a, g["a_direct"])


def test_rk4_steps(orc, golden):
    g = golden("rk4_1k")
    x1, v1 = orc.rk4_direct(g["x"], g["v"], float(g["dt"]), eps=float(g["eps"]))
    assert np.array_equal(x1, g["x_direct"]) and np.array_equal(v1, g["v_direct"])
    x2, v2 = orc.rk4_fmm(g["x"], g["v"], float(g["dt"]), int(g["s"]), int(g["p"]), eps=float(g["eps"]))
    dx = np.abs(g["x_fmm"] - g["x"]).max()
    assert np.abs(x2 - g["x_fmm"]).max() <= 1e-12 * dx
    assert np.abs(v2 - g["v_fmm"]).max() <= 1e-12 * np.abs(g["v_fmm"] - g["v"]).max()


def test_special_functions(orc, golden):
    g = golden("special")
    for k, (th, ph) in enumerate(g["angles"]):
        assert np.allclose(orc.P(12, np.cos(th)), g[f"P{k}"], rtol=1e-14, atol=1e-14)
        assert np.allclose(orc.DP(12, th), g[f"DP{k}"], rtol=1e-14, atol=1e-14)
        assert np.allclose(orc.Y(24, th, ph), g[f"Y{k}"], rtol=1e-13, atol=1e-15)
    A = np.array([[orc.A(n, m) if abs(m) <= n else 0.0 for m in range(-24, 25)] for n in range(25)])
    assert np.array_equal(A, g["A"])


def test_problem_initialize_stream(orc, golden):
    g = golden("problem_init")
    for n, seed, eps in ((1000, 0, None), (257, 7, 1e-3)):
        x, v, e = orc.problem_initialize(n, seed, eps)
        assert np.array_equal(x, g[f"x_{n}_{seed}"]) and np.array_equal(v, g[f"v_{n}_{seed}"])
        assert e == g[f"eps_{n}_{seed}"] and orc.delta_t(e) == g[f"dt_{n}_{seed}"]


def test_curve_fit_rows(orc, golden):
    g = golden("curve_fit")
    for row in g["rows"]:
        a, c = orc.curve_fit(row[:5], g["t"])
        assert a == row[5] and c == row[6]


# ------------------------------------------------------------ invariants ----
def _subtree_points(t, c, out):
    if t.child_base[c] < 0:
        out.extend(t.leaf_points(c))
    else:
        for q in range(8):
            _subtree_points(t, t.child_base[c] + q, out)
    return out


@pytest.mark.parametrize("name", ["uniform2k_s1", "clump3k_s2", "plummer4k_s5"])
def test_every_pair_covered_once(golden, orc, name):
    """Each ordered pair (i, j != i) is summed exactly once: by P2P from the NI
    list of i's leaf, or by M2L from exactly one AP entry of i's leaf or one of
    its ancestors at level >= 2 (the far field reaches a point through L2L/L2P)."""
    g = golden("tree_" + name)
    t, lists = orc.tree_and_lists(g["x"], int(g["s"]))
    n = len(g["x"])
    sub = {}

    def pts(c):
        if c not in sub:
            sub[c] = np.array(_subtree_points(t, c, []), dtype=np.int64)
        return sub[c]

    rng = np.random.default_rng(0)
    leaf_of = np.empty(n, np.int64)
    for c in range(t.ncells):
        if t.child_base[c] < 0:
            leaf_of[t.leaf_points(c)] = c
    for i in rng.choice(n, size=min(n, 150), replace=False):
        cover = np.zeros(n, np.int64)
        leaf = leaf_of[i]
        for c in lists.get("ni", leaf):
            cover[pts(c)] += 1
        a = leaf
        while a > 0:
            if t.level[a] >= 2:
                for c in lists.get("ap", a):
                    cover[pts(c)] += 1
            a = t.parent[a]
        cover[i] -= 1
        assert (cover[np.arange(n) != i] == 1).all() and cover[i] == 0


def test_fmm_converges_to_direct(orc):
    x, _, _ = orc.problem_initialize(3000, 3, 1e-3)
    d = orc.direct(x, eps=1e-3)
    errs = []
    for p in (2, 4, 6, 8):
        a = orc.fmm(x, 10, p, eps=1e-3)
        errs.append(np.median(np.linalg.norm(a - d, axis=1) / np.linalg.norm(d, axis=1)))
    assert all(b < a for a, b in zip(errs, errs[1:])) and errs[-1] < 1e-4, errs


def test_m2m_equals_direct_p2m(orc):
    """M2M(P2M(children)) equals P2M of all points about the parent centre."""
    x, _, _ = orc.problem_initialize(2500, 4, 1e-3)
    p = 6
    with orc.build(x, 5) as h:
        t = h.export()
        M = h.upward(p, 1.0 / len(x), min_level=2)
    # re-expand every level-2 internal cell from its points directly: build a one-cell tree about its centre
    checked = 0
    for c in np.where((t.level == 2) & (t.child_base >= 0))[0][:6]:
        idx = np.array(_subtree_points(t, c, []), dtype=np.int64)
        cen = (t.bmin[c] + t.bmax[c]) / 2
        # a single-leaf tree whose root centre is `cen`: two zero-mass points at cen +- R
        R = np.abs(x[idx] - cen).max(0) + 1.0
        xs = np.concatenate([x[idx], [cen + R, cen - R]])
        masses = np.concatenate([np.full(len(idx), 1.0 / len(x)), [0.0, 0.0]])
        lo, hi = xs.min(0) - 1e-5, xs.max(0) + 1e-5
        assert np.allclose((lo + hi) / 2, cen, rtol=0, atol=1e-12)
        with orc.build(xs, 10 ** 6) as h2:
            M2 = h2.upward(p, 0.0, min_level=0, masses=masses)
        ref = M2[0]
        assert np.abs(M[c] - ref).max() <= 1e-11 * np.abs(ref).max()
        checked += 1
    assert checked > 0
