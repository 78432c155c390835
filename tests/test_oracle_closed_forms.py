"""Pins of the oracle outputs that the reference has no counterpart for:
per-particle masses, G on the far field, and the potential.  None of them is
covered by a golden vector (the reference has one mass for all particles, no
potential and no G on its far field, fmm.hpp:234,370-372), so each is fixed
here by what the mathematics requires, never by restating the oracle's
formula:

* two-body closed forms with distinct masses and G != 1
  (solver_ref_naive.hpp:103-119 with the pair term of fmm.hpp:531-540);
* Newton's third law: sum_i m_i a_i = 0 for any masses (fails if the pair
  term used m_i for m_j);
* the potential energy U = -G/2 sum_i m_i pot_i is the one whose gradient
  gives the forces: -dU/dx_i = m_i a_i, by central differences (fails if the
  potential misses the softening or weighs the wrong mass);
* exact scalings: a(k m) = k a(m), a(G) with G on the far field = G a(1);
  with the reference's far field (G only in P2P) a(G) is affine in G and its
  G-independent part is the far field alone;
* a uniform mass array gives the scalar-mass result; a zero-mass particle
  pulls nothing;
* the FMM potential and mass-weighted forces converge to the direct sum
  pinned above.
"""
import numpy as np
import pytest


def _two_body(G, eps, m0, m1):
    x = np.array([[0.1, -0.2, 0.3], [0.45, 0.05, -0.15]])
    d = x[1] - x[0]
    r2 = d @ d + eps * eps
    a0 = G * m1 * d / r2 ** 1.5
    a1 = -G * m0 * d / r2 ** 1.5
    return x, np.array([a0, a1]), np.array([m1, m0]) / np.sqrt(r2)


@pytest.mark.parametrize("G,eps", [(1.0, 0.0), (2.5, 0.0), (2.5, 0.1), (0.37, 1e-3)])
def test_direct_two_body_closed_form(orc, G, eps):
    m0, m1 = 0.3, 1.7
    x, a_ref, pot_ref = _two_body(G, eps, m0, m1)
    a, pot = orc.direct(x, G=G, eps=eps, masses=np.array([m0, m1]), potential=True)
    assert np.allclose(a, a_ref, rtol=1e-14, atol=0)
    assert np.allclose(pot, pot_ref, rtol=1e-14, atol=0)
    # the scalar-mass path: every particle of mass m
    a, pot = orc.direct(x, m=m1, G=G, eps=eps, potential=True)
    _, a_ref, pot_ref = _two_body(G, eps, m1, m1)
    assert np.allclose(a, a_ref, rtol=1e-14, atol=0) and np.allclose(pot, pot_ref, rtol=1e-14, atol=0)


def test_direct_subset_two_body(orc):
    x, a_ref, pot_ref = _two_body(2.5, 0.1, 0.3, 1.7)
    a, pot = orc.direct_subset(x, [1], G=2.5, eps=0.1, masses=np.array([0.3, 1.7]), potential=True)
    assert np.allclose(a[0], a_ref[1], rtol=1e-14, atol=0) and np.allclose(pot[0], pot_ref[1], rtol=1e-14, atol=0)


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_direct_third_law_distinct_masses(orc, eps):
    rng = np.random.default_rng(11)
    x = rng.random((300, 3))
    m = rng.uniform(0.1, 3.0, len(x))
    a = orc.direct(x, G=1.9, eps=eps, masses=m)
    net = (m[:, None] * a).sum(0)
    assert np.abs(net).max() <= 1e-12 * (m[:, None] * np.abs(a)).sum()
    # the mistake this catches: each pull weighed by the target's own mass,
    # a_i = G m_i sum_j d_ij / r^3, breaks the law by far more than rounding
    wrong = m[:, None] * orc.direct(x, m=1.0, G=1.9, eps=eps)
    assert np.abs((m[:, None] * wrong).sum(0)).max() > 1e-3 * (m[:, None] * np.abs(wrong)).sum()


def _energy(orc, x, m, G, eps):
    _, pot = orc.direct(x, G=G, eps=eps, masses=m, potential=True)
    return -0.5 * G * (m * pot).sum()


@pytest.mark.parametrize("eps", [0.0, 0.02])
def test_direct_potential_gradient_is_force(orc, eps):
    """-dU/dx_i = m_i a_i by central differences, U = -G/2 sum m_i pot_i."""
    rng = np.random.default_rng(5)
    x = rng.random((40, 3))
    m = rng.uniform(0.2, 2.0, len(x))
    G = 1.3
    a = orc.direct(x, G=G, eps=eps, masses=m)
    h = 1e-6
    for i in (0, 7, 23):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            g = -(_energy(orc, xp, m, G, eps) - _energy(orc, xm, m, G, eps)) / (2 * h)
            assert abs(g - m[i] * a[i, k]) <= 1e-6 * np.abs(m[i] * a[i]).max(), (i, k, g, m[i] * a[i, k])


def test_direct_zero_mass_pulls_nothing(orc):
    rng = np.random.default_rng(2)
    x = rng.random((64, 3))
    m = rng.uniform(0.5, 1.5, len(x))
    m0 = m.copy()
    m0[10] = 0.0
    a, pot = orc.direct(x, eps=1e-3, masses=m0, potential=True)
    keep = np.arange(len(x)) != 10
    b, pb = orc.direct(x[keep], eps=1e-3, masses=m[keep], potential=True)
    assert np.allclose(a[keep], b, rtol=1e-13, atol=0) and np.allclose(pot[keep], pb, rtol=1e-13, atol=0)


# ------------------------------------------------------------------- FMM ---
@pytest.fixture(scope="module")
def cloud():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.random((3000, 3)), 0.3 + 0.05 * rng.random((1000, 3))])
    m = rng.uniform(0.5, 1.5, len(x)) / len(x)
    return x, m


def test_fmm_scales_with_masses(orc, cloud):
    x, m = cloud
    a = orc.fmm(x, 8, 6, eps=1e-3, masses=m)
    b = orc.fmm(x, 8, 6, eps=1e-3, masses=4.0 * m)  # a power of two: the scaling is exact in binary
    assert np.array_equal(b, 4.0 * a)
    c = orc.fmm(x, 8, 6, eps=1e-3, masses=3.0 * m)
    assert np.abs(c - 3.0 * a).max() <= 1e-14 * np.abs(a).max()


def test_fmm_uniform_mass_array_equals_scalar(orc, cloud):
    x, _ = cloud
    n = len(x)
    a = orc.fmm(x, 8, 6, eps=1e-3, m=1.0 / n)
    b = orc.fmm(x, 8, 6, eps=1e-3, masses=np.full(n, 1.0 / n))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("G", [0.5, 2.5])
def test_fmm_far_field_G(orc, cloud, G):
    """With G on the far field the whole force is linear in G.  With the
    reference's far field (G only in P2P, fmm.hpp:539 against fmm.hpp:234,
    370-372) it is affine in G; its part at G = 0 is the far field alone,
    and the part that scales is the P2P sum over the near-field lists."""
    x, m = cloud
    f1 = lambda g: orc.fmm(x, 8, 6, G=g, eps=1e-3, masses=m, far_g=True)
    f0 = lambda g: orc.fmm(x, 8, 6, G=g, eps=1e-3, masses=m, far_g=False)
    one = f1(1.0)
    scale = np.abs(one).max()
    assert np.abs(f1(G) - G * one).max() <= 1e-14 * G * scale
    assert np.array_equal(f0(1.0), one)
    far = f0(0.0)
    assert np.abs(f0(G) - (far + G * (one - far))).max() <= 1e-14 * max(1, G) * scale
    # the G = 0 part is the far field: it differs from the direct sum by
    # exactly the near field, the P2P sum over each leaf's NI cells
    d = orc.direct(x, eps=1e-3, masses=m)
    near = one - far
    with orc.build(x, 8) as h:
        t, lists = h.export(), h.lists()
    rng = np.random.default_rng(0)
    leaf_of = np.empty(len(x), np.int64)
    for c in range(t.ncells):
        if t.child_base[c] < 0:
            leaf_of[t.leaf_points(c)] = c

    def subtree(c):
        if t.child_base[c] < 0:
            return list(t.leaf_points(c))
        return sum((subtree(t.child_base[c] + q) for q in range(8)), [])

    for i in rng.choice(len(x), 20, replace=False):
        src = np.array(sum((subtree(c) for c in lists.get("ni", leaf_of[i])), []), np.int64)
        src = src[src != i]
        pair = orc.direct_subset(np.concatenate([x[i:i + 1], x[src]]), [0], eps=1e-3,
                                 masses=np.concatenate([[m[i]], m[src]]))[0]
        assert np.abs(near[i] - pair).max() <= 1e-12 * np.abs(d[i]).max()
    # and the far field plus the near field is the direct sum to the order-6 truncation
    rel = np.linalg.norm(one - d, axis=1) / np.linalg.norm(d, axis=1)
    assert np.median(rel) < 1e-4


def test_fmm_masses_converge_to_direct(orc, cloud):
    """Per-particle masses through P2M and P2P: the FMM converges with p to
    the direct sum, which the closed forms above pin."""
    x, m = cloud  # no softening: the far field is Newtonian, so the FMM then converges to the direct sum itself
    d = orc.direct(x, G=2.5, eps=0.0, masses=m)
    errs = []
    for p in (4, 8, 12):
        a = orc.fmm(x, 8, p, G=2.5, eps=0.0, masses=m, far_g=True)
        errs.append(np.median(np.linalg.norm(a - d, axis=1) / np.linalg.norm(d, axis=1)))
    assert errs[1] < errs[0] / 30 and errs[2] < errs[1] / 10 and errs[2] < 1e-6, errs


def test_fmm_potential_converges_to_direct(orc, cloud):
    """The FMM potential (softened near field + far-field local expansion)
    converges to the direct potential, which the gradient test pins."""
    x, m = cloud
    _, pd = orc.direct(x, eps=0.0, masses=m, potential=True)
    errs = []
    for p in (4, 8, 12):
        with orc.build(x, 8) as h:
            out = h.fmm(p, m=0.0, eps=0.0, masses=m, diagnostics=True)
        errs.append(np.median(np.abs(out["pot"] - pd) / np.abs(pd)))
    assert errs[1] < errs[0] / 30 and errs[2] < errs[1] / 10 and errs[2] < 1e-6, errs
