"""ctypes binding of the CPU oracle (liboracle.so) and, where built, of the
reference shim (_ref/libvipp_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
Both libraries expose the same call shapes, so `Oracle("orc")` and
`Oracle("ref")` are interchangeable in tests.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"orc": os.path.join(HERE, "liboracle.so"), "ref": os.path.join(HERE, "_ref", "libvipp_ref.so")}

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def available(kind: str) -> bool:
    return os.path.exists(LIBS[kind])


@dataclass
class Tree:
    parent: np.ndarray
    level: np.ndarray
    bmin: np.ndarray
    bmax: np.ndarray
    child_base: np.ndarray
    pt_off: np.ndarray
    pts: np.ndarray

    @property
    def ncells(self) -> int:
        return len(self.parent)

    def leaf_points(self, c: int) -> np.ndarray:
        return self.pts[self.pt_off[c]:self.pt_off[c + 1]]


@dataclass
class Lists:
    ap_off: np.ndarray
    ap: np.ndarray
    ni_off: np.ndarray
    ni: np.ndarray
    nb_off: np.ndarray
    nb: np.ndarray

    def get(self, name: str, c: int) -> np.ndarray:
        off = getattr(self, name + "_off")
        return getattr(self, name)[off[c]:off[c + 1]]


class Oracle:
    def __init__(self, kind: str = "orc"):
        path = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle{' ref' if kind == 'ref' else ''})")
        self.kind = kind
        self.lib = lib = C.CDLL(path)
        pf = kind + "_"
        self.pf = pf

        def fn(name, res, args):
            f = getattr(lib, pf + name)
            f.restype = res
            f.argtypes = args
            return f

        self._last_error = fn("last_error", C.c_char_p, [])
        self._tree_build = fn("tree_build", C.c_int, [_d, C.c_int64, C.c_int64, C.POINTER(_vp)])
        self._tree_free = fn("tree_free", None, [_vp])
        self._tree_ncells = fn("tree_ncells", C.c_int64, [_vp])
        self._tree_export = fn("tree_export", None, [_vp, _i64, _i32, _d, _d, _i64, _i64, _i64])
        self._tree_lists = fn("tree_lists", C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
        self._tree_subdivide = fn("tree_subdivide", C.c_int, [_vp, C.c_int64, C.c_int64])
        self._tree_lists_export = fn("tree_lists_export", None, [_vp, _i64, _i64, _i64, _i64, _i64, _i64])
        if kind == "orc":
            self._tree_fmm = fn("tree_fmm", C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_int,
                                                      _d, _vp, _vp, _vp, _vp, _vp])
            self._rk4_fmm = fn("rk4_fmm", C.c_int, [_d, _d, C.c_int64, _vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                                    C.c_int64, C.c_uint32, C.c_int])
            self._rk4_direct = fn("rk4_direct", C.c_int, [_d, _d, C.c_int64, _vp, C.c_double, C.c_double, C.c_double, C.c_double])
            self._direct = fn("direct_accel", None, [_d, C.c_int64, _vp, C.c_double, C.c_double, C.c_double, _d, _vp])
            self._tree_upward = fn("tree_upward", C.c_int, [_vp, _vp, C.c_double, C.c_uint32, C.c_int32, _d])
            self._direct_subset = fn("direct_accel_subset", None, [_d, C.c_int64, _vp, C.c_double, C.c_double, C.c_double,
                                                                   _i64, C.c_int64, _d, _vp])
        else:
            self._tree_fmm = fn("tree_fmm", C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_uint32, _d, _vp, _vp])
            self._rk4_fmm = fn("rk4_fmm", C.c_int, [_d, _d, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                                    C.c_int64, C.c_uint32])
            self._rk4_direct = fn("rk4_direct", C.c_int, [_d, _d, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double])
            self._direct = fn("direct_accel", None, [_d, C.c_int64, C.c_double, C.c_double, C.c_double, _d])
            self._time_sample = fn("time_sample", C.c_int, [_d, C.c_int64, C.c_int64, C.c_uint32, C.c_double, C.c_double,
                                                            C.c_double, C.c_int64, C.c_int32, _d, _d])
        self._P = fn("P", None, [C.c_uint32, C.c_double, _d])
        self._DP = fn("DP", None, [C.c_uint32, C.c_double, _d])
        self._Y = fn("Y", None, [C.c_uint32, C.c_double, C.c_double, _d])
        self._A = fn("A", C.c_double, [C.c_uint32, C.c_int32])
        self._init = fn("problem_initialize", C.c_double, [C.c_int64, C.c_uint64, C.c_double, _d, _d])
        self._delta_t = fn("problem_delta_t", C.c_double, [C.c_double])
        self._fit = fn("curve_fit", None, [_d, _d, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)])

    def _check(self, rc: int):
        if rc != 0:
            raise RuntimeError(self._last_error().decode())

    # -- tree and lists ----------------------------------------------------
    def build(self, x: np.ndarray, s: int):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        h = _vp()
        self._check(self._tree_build(x, len(x) // 3, int(s), C.byref(h)))
        return _Handle(self, h, len(x) // 3)

    def tree(self, x: np.ndarray, s: int) -> Tree:
        with self.build(x, s) as h:
            return h.export()

    def tree_and_lists(self, x: np.ndarray, s: int):
        with self.build(x, s) as h:
            return h.export(), h.lists()

    # -- forces --------------------------------------------------------------
    def fmm(self, x, s, p, m=None, G=1.0, eps=0.0, masses=None, far_g=False, diagnostics=False):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = len(x)
        if m is None:
            m = 1.0 / n
        with self.build(x, s) as h:
            return h.fmm(p, m=m, G=G, eps=eps, masses=masses, far_g=far_g, diagnostics=diagnostics)

    def direct(self, x, m=None, G=1.0, eps=0.0, masses=None, potential=False):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        n = len(x) // 3
        if m is None:
            m = 1.0 / n
        a = np.zeros(3 * n)
        if self.kind == "ref":
            assert masses is None and not potential
            self._direct(x, n, m, G, eps, a)
            return a.reshape(n, 3)
        pot = np.zeros(n) if potential else None
        ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
        self._direct(x, n, _ptr(ms), m, G, eps, a, _ptr(pot))
        return (a.reshape(n, 3), pot) if potential else a.reshape(n, 3)

    def direct_subset(self, x, idx, m=None, G=1.0, eps=0.0, masses=None, potential=False):
        assert self.kind == "orc"
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        n = len(x) // 3
        if m is None:
            m = 1.0 / n
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = np.zeros(3 * len(idx))
        pot = np.zeros(len(idx)) if potential else None
        ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
        self._direct_subset(x, n, _ptr(ms), m, G, eps, idx, len(idx), a, _ptr(pot))
        return (a.reshape(-1, 3), pot) if potential else a.reshape(-1, 3)

    def rk4_fmm(self, x, v, dt, s, p, m=None, G=1.0, eps=0.0, masses=None, far_g=False):
        x = np.array(x, dtype=np.float64).reshape(-1)
        v = np.array(v, dtype=np.float64).reshape(-1)
        n = len(x) // 3
        if m is None:
            m = 1.0 / n
        if self.kind == "ref":
            assert masses is None
            self._check(self._rk4_fmm(x, v, n, m, G, eps, dt, s, p))
        else:
            ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
            self._check(self._rk4_fmm(x, v, n, _ptr(ms), m, G, eps, dt, s, p, int(far_g)))
        return x.reshape(n, 3), v.reshape(n, 3)

    def rk4_direct(self, x, v, dt, m=None, G=1.0, eps=0.0, masses=None):
        x = np.array(x, dtype=np.float64).reshape(-1)
        v = np.array(v, dtype=np.float64).reshape(-1)
        n = len(x) // 3
        if m is None:
            m = 1.0 / n
        if self.kind == "ref":
            assert masses is None
            self._check(self._rk4_direct(x, v, n, m, G, eps, dt))
        else:
            ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
            self._check(self._rk4_direct(x, v, n, _ptr(ms), m, G, eps, dt))
        return x.reshape(n, 3), v.reshape(n, 3)

    # -- special functions / problem / fitting -------------------------------
    def P(self, nmax, x):
        out = np.zeros((nmax + 1) ** 2)
        self._P(nmax, x, out)
        return out.reshape(nmax + 1, nmax + 1)

    def DP(self, nmax, theta):
        out = np.zeros((nmax + 1) ** 2)
        self._DP(nmax, theta, out)
        return out.reshape(nmax + 1, nmax + 1)

    def Y(self, kmax, theta, phi):
        out = np.zeros(2 * (kmax + 1) * (2 * kmax + 1))
        self._Y(kmax, theta, phi, out)
        return (out[0::2] + 1j * out[1::2]).reshape(kmax + 1, 2 * kmax + 1)

    def A(self, n, m):
        return self._A(n, m)

    def problem_initialize(self, n, seed=0, eps=None):
        x = np.zeros(3 * n)
        v = np.zeros(3 * n)
        e = self._init(n, seed, -1.0 if eps is None else eps, x, v)
        return x.reshape(n, 3), v.reshape(n, 3), e

    def delta_t(self, eps):
        return self._delta_t(eps)

    def curve_fit(self, e_t, t):
        e_t = np.ascontiguousarray(e_t, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float64)
        a = C.c_double()
        c = C.c_double()
        self._fit(e_t, t, len(t), C.byref(a), C.byref(c))
        return a.value, c.value

    def time_sample(self, x, s, p, m, G, eps, per_phase, threads=1):
        """Reference FMM cost of one force evaluation, sampled: (seconds per
        phase, sampled fraction per phase).  P2M, M2M, P2P and L2L/L2P run
        per_phase * threads items on `threads` threads, M2L per_phase items on
        one (the reference's M2L loop is not race-free)."""
        assert self.kind == "ref"
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        secs = np.zeros(7)
        frac = np.zeros(7)
        self._check(self._time_sample(x, len(x) // 3, s, p, m, G, eps, per_phase, int(threads), secs, frac))
        return secs, frac


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _Handle:
    def __init__(self, o: Oracle, h, n: int):
        self.o = o
        self.h = h
        self._n = n

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.h:
            self.o._tree_free(self.h)
            self.h = None

    def export(self) -> Tree:
        nc = self.o._tree_ncells(self.h)
        parent = np.zeros(nc, np.int64)
        level = np.zeros(nc, np.int32)
        bmin = np.zeros(3 * nc)
        bmax = np.zeros(3 * nc)
        cb = np.zeros(nc, np.int64)
        off = np.zeros(nc + 1, np.int64)
        # leaf lists hold every point exactly once
        pts = np.zeros(max(1, self._n), np.int64)
        self.o._tree_export(self.h, parent, level, bmin, bmax, cb, off, pts)
        return Tree(parent, level, bmin.reshape(-1, 3), bmax.reshape(-1, 3), cb, off, pts[:off[-1]])

    def upward(self, p, m, min_level=2, masses=None):
        nc = self.o._tree_ncells(self.h)
        M = np.zeros(2 * nc * (p + 1) * (2 * p + 1))
        ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
        self.o._check(self.o._tree_upward(self.h, _ptr(ms), m, p, min_level, M))
        return (M[0::2] + 1j * M[1::2]).reshape(nc, p + 1, 2 * p + 1)

    def subdivide(self, idx: int, threshold: int):
        self.o._check(self.o._tree_subdivide(self.h, idx, threshold))

    def lists(self) -> Lists:
        nc = self.o._tree_ncells(self.h)
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.o._check(self.o._tree_lists(self.h, C.byref(a), C.byref(b), C.byref(c)))
        ap_off = np.zeros(nc + 1, np.int64)
        ni_off = np.zeros(nc + 1, np.int64)
        nb_off = np.zeros(nc + 1, np.int64)
        ap = np.zeros(max(1, a.value), np.int64)
        ni = np.zeros(max(1, b.value), np.int64)
        nb = np.zeros(max(1, c.value), np.int64)
        self.o._tree_lists_export(self.h, ap_off, ap, ni_off, ni, nb_off, nb)
        return Lists(ap_off, ap[:a.value], ni_off, ni[:b.value], nb_off, nb[:c.value])

    def fmm(self, p, m, G=1.0, eps=0.0, masses=None, far_g=False, diagnostics=False):
        n = self._n
        nc = self.o._tree_ncells(self.h)
        a = np.zeros(3 * n)
        stride = (p + 1) * (2 * p + 1)
        M = np.zeros(2 * nc * stride) if diagnostics else None
        L = np.zeros(2 * nc * stride) if diagnostics else None
        out = {}
        if self.o.kind == "ref":
            assert masses is None and not far_g
            self.o._check(self.o._tree_fmm(self.h, m, G, eps, p, a, _ptr(M), _ptr(L)))
        else:
            an = np.zeros(3 * n) if diagnostics else None
            af = np.zeros(3 * n) if diagnostics else None
            pot = np.zeros(n) if diagnostics else None
            ms = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
            self.o._check(self.o._tree_fmm(self.h, None, _ptr(ms), m, G, eps, p, int(far_g), a, _ptr(an), _ptr(af),
                                           _ptr(pot), _ptr(M), _ptr(L)))
            if diagnostics:
                out.update(a_near=an.reshape(n, 3), a_far=af.reshape(n, 3), pot=pot)
        if not diagnostics:
            return a.reshape(n, 3)
        cz = lambda v: (v[0::2] + 1j * v[1::2]).reshape(nc, p + 1, 2 * p + 1)
        out.update(a=a.reshape(n, 3), M=cz(M), L=cz(L))
        return out
