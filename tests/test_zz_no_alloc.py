"""vipp_grav_direct_noalloc, the allocation-free variant of vipp_grav_direct:
idx, a and pot in caller-owned device memory, partial sums in the library's
static device scratch.  Through the C ABI it must give the forces and
potentials of vipp_grav_direct on the same targets, on two calls in a row
(the second with fewer targets, so stale scratch contents of the first would
show), and reject more targets than its scratch holds and out-of-range
indices.

Tolerance: FP64, 1e-13 of the largest value.  Where both entry points cut the
sources into the same number of slices the results are bit-identical; the
first call below has more targets than the scratch holds at the
unconstrained slice count, so it runs with fewer slices and its sums are
rounded in another order."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """The caller's own device buffers (the entry point under test allocates none)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.lib.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.zeros(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _noalloc(vg, g, hip, idx, potential):
    lib = vg.library()
    f = lib.vipp_grav_direct_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    di = hip.put(idx)
    da = hip.alloc(24 * len(idx))
    dp = hip.alloc(8 * len(idx)) if potential else None
    rc = f(g.h, di, len(idx), da, dp)
    if rc != 0:
        raise vg.GravityError(rc, lib.vipp_grav_last_error(g.h).decode())
    a = hip.get(da, (len(idx), 3))
    pot = hip.get(dp, len(idx)) if potential else None
    return a, pot


@pytest.fixture(scope="module")
def vg():
    import vipp_gravity
    return vipp_gravity


def test_noalloc_direct_matches_direct(vg, orc):
    n = 200000
    x, v, _ = orc.problem_initialize(n, 9, 1e-3)
    m = np.random.default_rng(9).uniform(0.5, 1.5, n) / n
    hip = Hip()
    try:
        with vg.Gravity(n, 64, 8, "split", eps=1e-3, G=1.7) as g:
            g.upload(x, v, m)
            rng = np.random.default_rng(4)
            for k, nt in enumerate((100000, 3000)):  # the second call smaller than the first
                idx = rng.choice(n, nt, replace=False)
                a_ref, p_ref = g.direct(idx, potential=True)
                a, pot = _noalloc(vg, g, hip, idx, potential=True)
                assert np.abs(a - a_ref).max() <= 1e-13 * np.abs(a_ref).max(), k
                assert np.abs(pot - p_ref).max() <= 1e-13 * np.abs(p_ref).max(), k
                if k == 1:  # same slice count as vipp_grav_direct: bit-identical
                    assert np.array_equal(a, a_ref) and np.array_equal(pot, p_ref)
                a2, _ = _noalloc(vg, g, hip, idx, potential=False)
                assert np.array_equal(a2, a)
            with pytest.raises(vg.GravityError) as e:
                _noalloc(vg, g, hip, np.arange(524289) % n, potential=False)
            assert e.value.code == vg.E_INVALID
            with pytest.raises(vg.GravityError) as e:
                _noalloc(vg, g, hip, np.array([0, n, 5]), potential=False)
            assert e.value.code == vg.E_INVALID
            a, _ = _noalloc(vg, g, hip, np.array([0, 1, 5]), potential=False)  # usable after the rejected calls
            assert np.array_equal(a, g.direct([0, 1, 5]))
    finally:
        hip.free()
