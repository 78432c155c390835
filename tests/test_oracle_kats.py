"""The reference's own known-answer tests, restated against the CPU oracle.

Each test cites the gtest case it restates (gravity/test/*.cpp of the
reference).  The expected values are the closed forms those tests assert.
"""
import math

import numpy as np
import pytest

W = 1e-5  # Problem::OCTREE_WIDTH_PADDING (problem.hpp:167)
TOL = 1e-15


def box(lo, hi):
    return np.array(lo, float), np.array(hi, float)


# ---------------------------------------------------------------- octree ----
def test_empty_tree(orc):  # test_octree.cpp:97-103
    t = orc.tree(np.zeros((0, 3)), 1)
    assert t.ncells == 1 and t.child_base[0] == -1
    assert (t.bmin[0] == 0).all() and (t.bmax[0] == 0).all() and len(t.leaf_points(0)) == 0


def test_one_point_tree(orc):  # test_octree.cpp:105-111
    t = orc.tree(np.zeros((1, 3)), 1)
    assert t.ncells == 1
    np.testing.assert_allclose(t.bmin[0], -W, atol=TOL)
    np.testing.assert_allclose(t.bmax[0], W, atol=TOL)
    assert list(t.leaf_points(0)) == [0]


def _check_children(t, base, parent, lo, hi, mid, pts):
    """Children at base..base+7: offset bit 0/1/2 == 0 selects the high x/y/z half (octree.hpp:164-176)."""
    for off in range(8):
        c = base + off
        assert t.parent[c] == parent
        exp_lo = [mid[a] if ((off >> a) & 1) == 0 else lo[a] for a in range(3)]
        exp_hi = [hi[a] if ((off >> a) & 1) == 0 else mid[a] for a in range(3)]
        np.testing.assert_allclose(t.bmin[c], exp_lo, atol=TOL)
        np.testing.assert_allclose(t.bmax[c], exp_hi, atol=TOL)
        if pts.get(off) is not None:
            assert list(t.leaf_points(c)) == pts[off], (off, list(t.leaf_points(c)))


def test_three_point_tree(orc):  # test_octree.cpp:113-128
    x = np.array([[1, 1, 1], [-1, -1, -1], [1, -1, 1]], float)
    t = orc.tree(x, 1)
    assert t.child_base[0] == 1 and t.ncells == 9
    lo, hi = [-W - 1] * 3, [W + 1] * 3
    _check_children(t, 1, 0, lo, hi, [0, 0, 0], {0: [0], 1: [], 2: [2], 3: [], 4: [], 5: [], 6: [], 7: [1]})


def test_four_point_tree(orc):  # test_octree.cpp:130-154
    x = np.array([[4, 4, 4], [-4, -4, -4], [2, 2, 2], [3, 3, 3]], float)
    t = orc.tree(x, 1)
    assert t.child_base[0] == 1 and t.child_base[1] == 9 and t.ncells == 17
    _check_children(t, 1, 0, [-W - 4] * 3, [W + 4] * 3, [0, 0, 0], {1: [], 2: [], 7: [1]})
    h = (W + 4) / 2
    _check_children(t, 9, 1, [0, 0, 0], [W + 4] * 3, [h, h, h], {0: [0, 3], 1: [], 6: [], 7: [2]})


# ------------------------------------------------------ special functions ---
def test_legendre_at_minus_one(orc):  # test_fmm.cpp:131-159
    P = orc.P(5, -1.0)
    for n in range(6):
        assert abs(P[n, 0] - (-1) ** n) <= TOL
        for m in range(1, n + 1):
            assert abs(P[n, m]) <= TOL


def test_legendre_at_zero(orc):  # test_fmm.cpp:161-189
    P = orc.P(5, 0.0)
    exp = {(0, 0): 1, (1, 0): 0, (2, 0): -0.5, (3, 0): 0, (4, 0): 3 / 8, (5, 0): 0, (1, 1): -1, (2, 1): 0, (3, 1): 1.5,
           (4, 1): 0, (5, 1): -15 / 8, (2, 2): 3, (3, 2): 0, (4, 2): -7.5, (5, 2): 0, (3, 3): -15, (4, 3): 0,
           (5, 3): 52.5, (4, 4): 105, (5, 4): 0, (5, 5): -945}
    for (n, m), v in exp.items():
        assert abs(P[n, m] - v) <= TOL, (n, m, P[n, m], v)


def test_legendre_at_one(orc):  # test_fmm.cpp:191-219
    P = orc.P(5, 1.0)
    for n in range(6):
        assert abs(P[n, 0] - 1) <= TOL
        for m in range(1, n + 1):
            assert abs(P[n, m]) <= TOL


def test_y_at_pole(orc):  # test_fmm.cpp:221-232
    Y = orc.Y(5, 0.0, 0.0)
    for k in range(5):
        for m in range(-k, k):
            exp = 1.0 if m == 0 else 0.0
            assert abs(Y[k, 5 + m] - exp) <= TOL


def test_y_quarter_pi(orc):  # test_fmm.cpp:234-278 (phi = 0, so every value is real)
    Y = orc.Y(5, math.pi / 4, 0.0)
    s2 = math.sqrt(2)
    half = {
        (0, 0): 1,
        (1, 0): s2 / 2, (1, 1): -0.5,
        (2, 0): 0.25, (2, 1): -1.5 / math.sqrt(6), (2, 2): 1.5 / math.sqrt(24),
        (3, 0): s2 / -8, (3, 1): s2 * -9 / 8 / math.sqrt(12), (3, 2): s2 * 15 / 4 / math.sqrt(120),
        (3, 3): s2 * -15 / 4 / math.sqrt(720),
        (4, 0): -13 / 32, (4, 1): -5 / 8 / math.sqrt(20), (4, 2): 75 / 8 / math.sqrt(360),
        (4, 3): -105 / 4 / math.sqrt(5040), (4, 4): 105 / 4 / math.sqrt(40320),
        (5, 0): s2 * -17 / 64, (5, 1): s2 * 45 / 64 / math.sqrt(30), (5, 2): s2 * 105 / 16 / math.sqrt(840),
        (5, 3): s2 * -735 / 16 / math.sqrt(20160), (5, 4): s2 * 945 / 8 / math.sqrt(362880),
        (5, 5): s2 * -945 / 8 / math.sqrt(3628800),
    }
    for (k, m), v in half.items():
        for mm in (m, -m):  # the reference's Y_k^{-m} uses P_k^{|m|} (fmm.hpp:177-179)
            assert abs(Y[k, 5 + mm] - v) <= TOL, (k, mm)


def test_y_half_pi(orc):  # test_fmm.cpp:280-324
    Y = orc.Y(5, math.pi / 2, 0.0)
    half = {(0, 0): 1, (1, 0): 0, (1, 1): -1 / math.sqrt(2), (2, 0): -0.5, (2, 1): 0, (2, 2): 3 / math.sqrt(24),
            (3, 0): 0, (3, 1): 1.5 / math.sqrt(12), (3, 2): 0, (3, 3): -15 / math.sqrt(720),
            (4, 0): 3 / 8, (4, 1): 0, (4, 2): -7.5 / math.sqrt(360), (4, 3): 0, (4, 4): 105 / math.sqrt(40320),
            (5, 0): 0, (5, 1): -15 / 8 / math.sqrt(30), (5, 2): 0, (5, 3): 52.5 / math.sqrt(20160), (5, 4): 0,
            (5, 5): -945 / math.sqrt(3628800)}
    for (k, m), v in half.items():
        for mm in (m, -m):
            assert abs(Y[k, 5 + mm] - v) <= TOL, (k, mm)


# ----------------------------------------------------------- operators ------
def test_p2m_at_centre(orc):  # test_fmm.cpp:326-338
    x = np.array([[1.0, 1.0, 1.0]])
    with orc.build(x, 1) as h:
        M = h.upward(2, 1.0, min_level=0)
    assert abs(M[0, 0, 2] - 1) <= TOL and abs(M[0, 1, 2]) <= TOL and abs(M[0, 1, 3]) <= TOL


def test_m2m_monopole_dipole(orc):  # test_fmm.cpp:340-355
    x = np.array([[1.0, 0.0, 0.0]])
    with orc.build(x, 1) as h:
        h.subdivide(0, 2)
        M = h.upward(2, 1.0, min_level=0)
    assert abs(M[0, 0, 2] - 1) <= TOL and abs(M[0, 1, 2]) <= TOL


def test_m2l_closed_form(orc):  # test_fmm.cpp:357-387: L_{1,1} of cell 9 at p=2
    x = np.array([[1, 1, 1], [1, -1, -1], [-1, -1, -1]], float)
    p = 2
    with orc.build(x, 1) as h:
        h.subdivide(1, 2)
        h.subdivide(8, 2)
        out = h.fmm(p, m=1.0 / 3, G=1.0, eps=0.0, diagnostics=True)
    got = out["L"][9, 1, p + 1]
    e = W
    rho = (1 + e) * 3 / 2 * math.sqrt(3)
    q2, q3, q5, q6 = (math.sqrt(v) for v in (2, 3, 5, 6))
    a = q2 / 24 - q2 / 8 * e
    b = q6 / 96 - q6 / 16 * e + 3 * q6 / 32 * e * e
    terms = [
        complex(1 / 3, 0) * (-1 / q2) * complex(1 / q6, -1 / q6) / (-q2 / 2 * rho ** 2),
        complex(a, a) * 0.5 * complex(0, -1 / q6) / (-q6 / 12 * rho ** 3),
        complex(1 / 12 - e / 4, 0) * (-q2 / 2) * complex(-1 / q6, 1 / q6) / (-q6 / 6 * rho ** 3),
        complex(a, -a) * (-0.5) * 0 / (-0.5 * rho ** 3),
        complex(0, b) * (-q3 / 12) * complex(-q3 * q5 / 18, -q3 * q5 / 18) / (-1 / 12 / q5 * rho ** 4),
        complex(-b, -b) * (-q3 / 6) * complex(0, q2 * q5 / 6) / (-1 / 2 / q2 / q3 / q5 * rho ** 4),
        0,
        complex(-b, b) * (q3 / 6) * complex(2 * q3 / 9, 0) / (-1 / 6 * rho ** 4),
        complex(0, -b) * (q3 / 12) * complex(1 / 6, 1 / 6) / (-q3 / 12 * rho ** 4),
    ]
    exp = sum(terms)
    assert abs(got.real - exp.real) <= TOL and abs(got.imag - exp.imag) <= TOL, (got, exp)


def test_many_points_finite(orc):  # test_fmm.cpp:102-129 (N=1000 Problem, s=5, p=5: no NaN)
    x, v, e = orc.problem_initialize(1000, 0, None)
    out = orc.fmm(x, 5, 5, eps=e, diagnostics=True)
    assert np.isfinite(out["M"]).all() and np.isfinite(out["L"]).all() and np.isfinite(out["a"]).all()


# ---------------------------------------------------- direct sum and RK4 ---
def test_naive_rk4_two_body(orc):  # test_solver_ref_naive.cpp:47-79
    x = np.array([[0, 0, 0], [10, 0, 0]], float)
    v = np.zeros((2, 3))
    eps = 1e-4
    x1, v1 = orc.rk4_direct(x, v, 1.0, m=1.0, G=1.0, eps=eps)
    d = 10.0
    k1 = d / (d * d + eps * eps) ** 1.5
    l2 = k1 / 2
    k2 = k1
    l3 = l2
    k3 = (d - l2) / ((d - l2) ** 2 + eps * eps) ** 1.5
    l4 = k3
    k4 = (d - 2 * l3) / ((d - 2 * l3) ** 2 + eps * eps) ** 1.5
    xd = (0 + 2 * l2 + 2 * l3 + l4) / 6
    vd = (k1 + 2 * k2 + 2 * k3 + k4) / 6
    np.testing.assert_allclose(x1, [[xd, 0, 0], [d - xd, 0, 0]], atol=TOL)
    np.testing.assert_allclose(v1, [[vd, 0, 0], [-vd, 0, 0]], atol=TOL)


def test_naive_acceleration(orc):  # test_solver_ref_naive.cpp:81-105
    eps = 1e-4
    a = orc.direct(np.array([[0, 0, 0], [10, 0, 0]], float), m=1.0, G=1.0, eps=eps)
    np.testing.assert_allclose(a[0], [10 / (100 + eps * eps) ** 1.5, 0, 0], atol=TOL)
    a = orc.direct(np.array([[0, 0, 0], [5, 0, 0], [-5, 0, 0]], float), m=1.0, G=1.0, eps=eps)
    np.testing.assert_allclose(a[0], [0, 0, 0], atol=TOL)
    np.testing.assert_allclose(a[1], [-5 / (25 + eps * eps) ** 1.5 - 10 / (100 + eps * eps) ** 1.5, 0, 0], atol=TOL)


# ------------------------------------------------ problem and benchmark ---
def test_problem_epsilon(orc):  # test_problem.cpp:59-64
    assert abs(orc.problem_initialize(1)[2] - 0.63) <= TOL
    assert abs(orc.problem_initialize(2)[2] - 0.63 * 2 ** -0.22) <= TOL


@pytest.mark.parametrize("e,alpha,c", [
    ([1, 2, 4, 8, 16], 1, 1), ([2, 4, 8, 16, 32], 2, 1), ([2, 8, 32, 128, 512], 2, 2),
    ([2, 0.5, 1 / 8, 1 / 32, 1 / 128], 2, -2)])
def test_curve_fitting(orc, e, alpha, c):  # test_benchmark.cpp:46-83
    a, cc = orc.curve_fit(e, [1, 2, 4, 8, 16])
    assert abs(a - alpha) <= 1e-9 and abs(cc - c) <= 1e-9


def test_curve_fitting_perfect(orc):  # test_benchmark.cpp:46-51
    a, _ = orc.curve_fit([0, 0, 0, 0, 0], [1, 2, 4, 8, 16])
    assert abs(a) <= 1e-9
