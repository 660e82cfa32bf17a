"""hyd_sum_sequential_noalloc (no allocation, no context; its work area is a
static device array) against the existing entry point hyd_sum_sequential
and the oracle's sequential sum, bit for bit, on two calls in a row (the
second with a smaller input), through the C ABI.  The device buffers and the
stream come from the HIP runtime the library itself links (one runtime per
process), called through ctypes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    """Partial sums that wander across binades and through zero, as the
    contour-grid mean does: a random walk of mixed-sign terms of varied
    magnitude."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 3, n)
    a[::7] = -a[::7] * 1.0000001
    return a


class _Hip:
    """The few HIP runtime calls the test needs, from the library's own
    libamdhip64 (already loaded by libhydra_amd.so)."""

    def __init__(self):
        import os
        path = "/opt/rocm/lib/libamdhip64.so.7"
        self.L = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]

    def check(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), nbytes))
        return p


H2D, D2H = 1, 2


def _noalloc_sum(L, hip, stream, a_host):
    n = len(a_host)
    a = hip.malloc(8 * max(n, 1))
    out = hip.malloc(8)
    try:
        hip.check(hip.L.hipMemcpy(a, a_host.ctypes.data, 8 * n, H2D))
        nan = np.array([np.nan])
        hip.check(hip.L.hipMemcpy(out, nan.ctypes.data, 8, H2D))
        rc = L.hyd_sum_sequential_noalloc(a, C.c_int64(n), out, stream)
        assert rc == 0, L.hyd_last_error().decode()
        hip.check(hip.L.hipStreamSynchronize(stream))
        res = np.zeros(1)
        hip.check(hip.L.hipMemcpy(res.ctypes.data, out, 8, D2H))
        return float(res[0])
    finally:
        hip.L.hipFree(a)
        hip.L.hipFree(out)


def test_noalloc_sum_matches_existing_entry_point():
    from hydra_amd import Hydra, lib
    from oracle.hyo import lib as olib
    L = lib()
    L.hyd_sum_sequential_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    h = Hydra(1024)
    hip = _Hip()
    stream = C.c_void_p()
    hip.check(hip.L.hipStreamCreate(C.byref(stream)))
    ol = olib()
    ol.hyo_sum_seq.restype = C.c_double
    ol.hyo_sum_seq.argtypes = [np.ctypeslib.ndpointer(dtype=np.float64), C.c_int64]
    for n, seed in [(1 << 22, 1), (1_000_037, 2), (1, 3)]:
        a = _inputs(n, seed)
        got = _noalloc_sum(L, hip, stream, a)
        want = h.sum_sequential(a)
        assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64), n
        assert np.float64(got).view(np.int64) == np.float64(ol.hyo_sum_seq(a, n)).view(np.int64)
    hip.L.hipStreamDestroy(stream)
    h.close()


def test_noalloc_sum_limits():
    from hydra_amd import lib
    L = lib()
    L.hyd_sum_sequential_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    dummy = C.c_void_p(16)
    assert L.hyd_sum_sequential_noalloc(dummy, C.c_int64((1 << 22) + 1), dummy, None) == -3
    assert L.hyd_sum_sequential_noalloc(dummy, C.c_int64(-1), dummy, None) == -1
    assert L.hyd_sum_sequential_noalloc(dummy, C.c_int64(5), None, None) == -1
