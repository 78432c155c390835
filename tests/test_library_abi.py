"""CPU-side checks of the product library: it loads, exports every function
declared in include/vipp_gravity.h, and validates its configuration before
touching a device (problem.hpp:179-184 rules)."""
import ctypes

import pytest


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
