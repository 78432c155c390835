"""Python binding of libvipp_gravity.so (the C ABI of include/vipp_gravity.h).

Host-side plumbing for tests and bench.py: numpy arrays in, numpy arrays out.
The library is required; importing this module never falls back to anything
else, and every call raises GravityError with the library's message on failure.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvipp_gravity.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "vipp_gravity.h")

F64, F32, F32_SPLIT = 0, 1, 2
PRECISION = {"f64": F64, "f32": F32, "split": F32_SPLIT}
FLAG_REF_FAR_G = 1
PHASES = ["tree", "lists", "upward", "m2l", "p2p", "downward", "rk4", "operators"]

E_INVALID, E_INPUT, E_DEPTH, E_DEVICE, E_STATE = 1, 2, 3, 4, 5
MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_NPARAMS = 28, 128, 25, 23449


class GravityError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("n", C.c_uint64), ("s", C.c_uint64), ("p", C.c_uint32), ("precision", C.c_uint32),
                ("epsilon", C.c_double), ("G", C.c_double), ("mass", C.c_double), ("device", C.c_int32),
                ("flags", C.c_uint32)]


_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built: run python vipp-nbody-bench_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(lib):
    vp, d, i64, i32 = C.c_void_p, C.c_double, C.c_int64, C.c_int32
    P = C.POINTER
    sig = {
        "vipp_grav_create": [P(Config), P(vp)],
        "vipp_grav_destroy": [vp],
        "vipp_grav_upload": [vp, vp, vp, vp],
        "vipp_grav_download": [vp, vp, vp],
        "vipp_grav_step": [vp, d, C.c_uint32],
        "vipp_grav_accelerations": [vp, vp],
        "vipp_grav_energy": [vp, P(d), P(d)],
        "vipp_grav_tree_info": [vp, P(i64), P(i64), P(i32)],
        "vipp_grav_export_tree": [vp, vp, vp, vp, vp, vp, vp, vp],
        "vipp_grav_export_lists_sizes": [vp, P(i64), P(i64), P(i64)],
        "vipp_grav_export_lists": [vp, vp, vp, vp, vp, vp, vp],
        "vipp_grav_export_expansions": [vp, vp, vp],
        "vipp_grav_counts": [vp, vp, i32],
        "vipp_grav_phase_times": [vp, vp, i32],
        "vipp_grav_set_timing": [vp, i32],
        "vipp_grav_device_state": [vp, P(vp), P(vp)],
        "vipp_grav_direct": [vp, vp, i64, vp, vp],
        "vipp_grav_leaf_inputs": [vp, P(i64), vp, vp, vp, vp],
        "vipp_grav_set_closure": [vp, vp, vp, d, C.c_uint32],
        "vipp_mlp_forward": [i32, vp, vp, i64, vp],
        "vipp_mlp_loss_grad": [i32, vp, vp, vp, vp, i64, P(C.c_float), vp],
        "vipp_mlp_train": [i32, vp, vp, vp, vp, i64, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, vp],
        "vipp_grav_problem_initialize": [C.c_uint64, C.c_uint64, d, vp, vp, P(d)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.vipp_grav_last_error.argtypes = [vp]
    lib.vipp_grav_last_error.restype = C.c_char_p


def problem_initialize(n, seed=0, eps=None):
    """The reference's Problem::initialize stream (problem.hpp:188-211) for n
    particles: (x, v, epsilon), x and v (n, 3) float64.  eps None selects the
    reference's default softening 0.63 n^-0.22.  Host code in the library."""
    lib = library()
    x = np.zeros(3 * int(n))
    v = np.zeros(3 * int(n))
    e = C.c_double()
    rc = lib.vipp_grav_problem_initialize(int(n), int(seed), -1.0 if eps is None else float(eps), _ptr(x), _ptr(v), C.byref(e))
    if rc != 0:
        raise GravityError(rc, "problem_initialize: n must be > 0 and epsilon < 1")
    return x.reshape(-1, 3), v.reshape(-1, 3), e.value


def declared_symbols() -> list[str]:
    """Function names declared by include/vipp_gravity.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vipp_grav_[a-z_]+)\s*\(", txt)))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class TreeExport:
    parent: np.ndarray
    level: np.ndarray
    bmin: np.ndarray
    bmax: np.ndarray
    child_base: np.ndarray
    pt_off: np.ndarray
    pts: np.ndarray

    @property
    def ncells(self):
        return len(self.parent)


@dataclass
class ListsExport:
    ap_off: np.ndarray
    ap: np.ndarray
    ni_off: np.ndarray
    ni: np.ndarray
    nb_off: np.ndarray
    nb: np.ndarray


class Gravity:
    """One device context: FP64 state, tree, lists and expansions of N particles."""

    def __init__(self, n, s, p, precision="f32", eps=0.0, G=1.0, mass=None, device=0, ref_far_g=False):
        self.lib = library()
        self.n = int(n)
        self.p = int(p)
        cfg = Config(n=int(n), s=int(s), p=int(p), precision=PRECISION[precision] if isinstance(precision, str) else int(precision),
                     epsilon=float(eps), G=float(G), mass=float(1.0 / n if mass is None else mass), device=int(device),
                     flags=FLAG_REF_FAR_G if ref_far_g else 0)
        h = C.c_void_p()
        rc = self.lib.vipp_grav_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise GravityError(rc, self.lib.vipp_grav_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.vipp_grav_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise GravityError(rc, self.lib.vipp_grav_last_error(self.h).decode())

    def upload(self, x, v=None, masses=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        if x.size != 3 * self.n or v.size != 3 * self.n:
            raise ValueError(f"x and v must hold 3 * {self.n} values, got {x.size} and {v.size}")
        m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64).reshape(-1)
        if m is not None and m.size != self.n:
            raise ValueError(f"masses must hold {self.n} values, got {m.size}")
        self._check(self.lib.vipp_grav_upload(self.h, _ptr(x), _ptr(v), _ptr(m)))

    def download(self):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        self._check(self.lib.vipp_grav_download(self.h, _ptr(x), _ptr(v)))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def step(self, dt, nsteps=1):
        self._check(self.lib.vipp_grav_step(self.h, float(dt), int(nsteps)))

    def accelerations(self, copy=True):
        a = np.zeros(3 * self.n) if copy else None
        self._check(self.lib.vipp_grav_accelerations(self.h, _ptr(a)))
        return None if a is None else a.reshape(-1, 3)

    def direct(self, idx, potential=False):
        """FP64 direct-sum accelerations (and potentials) of particles idx on the device."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        a = np.zeros(3 * len(idx))
        pot = np.zeros(len(idx)) if potential else None
        self._check(self.lib.vipp_grav_direct(self.h, _ptr(idx), len(idx), _ptr(a), _ptr(pot)))
        return (a.reshape(-1, 3), pot) if potential else a.reshape(-1, 3)

    def leaf_inputs(self):
        """Closure view of the last evaluation: dict of x (n x 28), norm, l25 (n x 25), level."""
        n = C.c_int64()
        self._check(self.lib.vipp_grav_leaf_inputs(self.h, C.byref(n), None, None, None, None))
        k = n.value
        x = np.zeros((max(1, k), 28), np.float32)
        norm = np.zeros(max(1, k), np.float32)
        l25 = np.zeros((max(1, k), 25), np.float32)
        lev = np.zeros(max(1, k), np.int32)
        self._check(self.lib.vipp_grav_leaf_inputs(self.h, C.byref(n), _ptr(x), _ptr(norm), _ptr(l25), _ptr(lev)))
        return {"x": x[:k], "norm": norm[:k], "l25": l25[:k], "level": lev[:k]}

    def set_closure(self, params, io=None, bound=float("inf"), p_fallback=12):
        """Enable the learned closure (params: VIPP_MLP_NPARAMS float32; io: 81
        float32 standardisation values or None) or disable it (params None)."""
        if params is None:
            self._check(self.lib.vipp_grav_set_closure(self.h, None, None, 0.0, 0))
            return
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_NPARAMS
        io = None if io is None else np.ascontiguousarray(io, dtype=np.float32)
        assert io is None or io.size == 81
        self._check(self.lib.vipp_grav_set_closure(self.h, _ptr(params), _ptr(io), float(bound), int(p_fallback)))

    def energy(self):
        ke, pe = C.c_double(), C.c_double()
        self._check(self.lib.vipp_grav_energy(self.h, C.byref(ke), C.byref(pe)))
        return ke.value, pe.value

    def tree_info(self):
        nc, nl, dp = C.c_int64(), C.c_int64(), C.c_int32()
        self._check(self.lib.vipp_grav_tree_info(self.h, C.byref(nc), C.byref(nl), C.byref(dp)))
        return nc.value, nl.value, dp.value

    def tree(self) -> TreeExport:
        nc, _, _ = self.tree_info()
        parent = np.zeros(nc, np.int64)
        level = np.zeros(nc, np.int32)
        bmin = np.zeros(3 * nc)
        bmax = np.zeros(3 * nc)
        cb = np.zeros(nc, np.int64)
        off = np.zeros(nc + 1, np.int64)
        pts = np.zeros(max(1, self.n), np.int64)
        self._check(self.lib.vipp_grav_export_tree(self.h, _ptr(parent), _ptr(level), _ptr(bmin), _ptr(bmax), _ptr(cb), _ptr(off), _ptr(pts)))
        return TreeExport(parent, level, bmin.reshape(-1, 3), bmax.reshape(-1, 3), cb, off, pts[:off[-1]])

    def lists(self) -> ListsExport:
        nc, _, _ = self.tree_info()
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.vipp_grav_export_lists_sizes(self.h, C.byref(a), C.byref(b), C.byref(c)))
        ao, no, bo = (np.zeros(nc + 1, np.int64) for _ in range(3))
        ap, ni, nb = np.zeros(max(1, a.value), np.int64), np.zeros(max(1, b.value), np.int64), np.zeros(max(1, c.value), np.int64)
        self._check(self.lib.vipp_grav_export_lists(self.h, _ptr(ao), _ptr(ap), _ptr(no), _ptr(ni), _ptr(bo), _ptr(nb)))
        return ListsExport(ao, ap[:a.value], no, ni[:b.value], bo, nb[:c.value])

    def expansions(self):
        nc, _, _ = self.tree_info()
        stride = (self.p + 1) * (2 * self.p + 1)
        M = np.zeros(2 * nc * stride)
        L = np.zeros(2 * nc * stride)
        self._check(self.lib.vipp_grav_export_expansions(self.h, _ptr(M), _ptr(L)))
        cz = lambda v: (v[0::2] + 1j * v[1::2]).reshape(nc, self.p + 1, 2 * self.p + 1)
        return cz(M), cz(L)

    def counts(self):
        out = np.zeros(8)
        self._check(self.lib.vipp_grav_counts(self.h, _ptr(out), 8))
        return {"p2p_pairs": out[0], "m2l_pairs": out[1], "cells": int(out[2]), "leaves": int(out[3]), "depth": int(out[4]),
                "m2l_live_pairs": out[5], "fallback_leaves": int(out[6]), "list_entries": out[7]}

    def set_timing(self, on=True):
        self._check(self.lib.vipp_grav_set_timing(self.h, 1 if on else 0))

    def phase_times(self, reset=False):
        ms = np.zeros(len(PHASES))
        self._check(self.lib.vipp_grav_phase_times(self.h, _ptr(ms), len(PHASES)))
        if reset:
            self._check(self.lib.vipp_grav_phase_times(self.h, None, 0))
        return dict(zip(PHASES, ms.tolist()))


def _mlp_check(lib, rc):
    if rc != 0:
        raise GravityError(rc, lib.vipp_grav_last_error(None).decode())


def mlp_forward(params, x, device=0):
    """Network outputs (n x 25) on the device."""
    lib = library()
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((len(x), MLP_OUT), np.float32)
    _mlp_check(lib, lib.vipp_mlp_forward(device, _ptr(params), _ptr(x), len(x), _ptr(y)))
    return y


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def mlp_loss_grad(params, x, t, w=None, device=0):
    """((weighted) mean squared error, gradient) on the device."""
    lib = library()
    params, x, t, w = _f32(params), _f32(x), _f32(t), _f32(w)
    loss = C.c_float()
    g = np.zeros(MLP_NPARAMS, np.float32)
    _mlp_check(lib, lib.vipp_mlp_loss_grad(device, _ptr(params), _ptr(x), _ptr(t), _ptr(w), len(x), C.byref(loss), _ptr(g)))
    return float(loss.value), g


def mlp_train(params, x, t, epochs, batch, lr, seed=0, w=None, device=0):
    """Adam training on the device; returns (new params, per-epoch mean loss)."""
    lib = library()
    params = np.array(params, dtype=np.float32, copy=True)
    x, t, w = _f32(x), _f32(t), _f32(w)
    hist = np.zeros(int(epochs), np.float32)
    _mlp_check(lib, lib.vipp_mlp_train(device, _ptr(params), _ptr(x), _ptr(t), _ptr(w), len(x), int(epochs), int(batch), float(lr), int(seed),
                                       _ptr(hist)))
    return params, hist
