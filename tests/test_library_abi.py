"""CPU-side checks of the product library: it loads, exports every function
declared in include/vipp_gravity.h, and validates its configuration before
touching a device (problem.hpp:179-184 rules)."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    import vipp_gravity as vgm
    lib = vgm.library()
    names = vgm.declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("kw,code", [
    (dict(n=0, s=64, p=4), 1),
    (dict(n=100, s=0, p=4), 1),
    (dict(n=100, s=64, p=13), 1),
    (dict(n=100, s=64, p=0), 1),
])
def test_config_validation(kw, code):
    import vipp_gravity as vgm
    with pytest.raises(vgm.GravityError) as e:
        _create_raw(vgm, **kw)
    assert e.value.code == code


def _create_raw(vgm, n, s, p):
    lib = vgm.library()
    cfg = vgm.Config(n=n, s=s, p=p, precision=1, epsilon=0.0, G=1.0, mass=1.0, device=0, flags=0)
    h = ctypes.c_void_p()
    rc = lib.vipp_grav_create(ctypes.byref(cfg), ctypes.byref(h))
    raise vgm.GravityError(rc, lib.vipp_grav_last_error(None).decode())


def test_cpp_wrapper_compiles_cleanly():
    """include/gravity/solver_gpu.hpp with the test's own Problem / Vector types:
    no warnings under -Wall -Wextra."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "solver_gpu_test.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_header_is_plain_c():
    """The C ABI header compiles as C99 (no C++ in the signatures)."""
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-x", "c", os.path.join(ROOT, "include", "vipp_gravity.h")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_problem_initialize_matches_reference_stream(golden):
    """The library's generator of the reference's initial conditions (host
    code) against vectors the reference's own Problem::initialize wrote
    (problem.hpp:188-211, default softening problem.hpp:290-292): bit for bit."""
    import numpy as np
    import vipp_gravity as vgm
    g = golden("problem_init")
    for n, seed, eps in ((1000, 0, None), (257, 7, 1e-3)):
        x, v, e = vgm.problem_initialize(n, seed, eps)
        assert np.array_equal(x, g[f"x_{n}_{seed}"]) and np.array_equal(v, g[f"v_{n}_{seed}"])
        assert e == g[f"eps_{n}_{seed}"]
    with pytest.raises(vgm.GravityError):
        vgm.problem_initialize(0)
