"""Shared fixtures for the parity tests: seeded fields and oracle contours."""
import functools

import numpy as np

from hydra_amd.ic import random_pv


@functools.lru_cache(maxsize=None)
def oracle_state(ng, seed=12345):
    """Oracle model initialised from the seeded field and recontoured
    (caps.f90 initialise + congen recontour); returns (oracle, qq, contours)."""
    from oracle.hyo import Oracle
    qq = random_pv(ng, seed=seed)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    return o, qq, o.contours()


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                                 b.view(np.int64) if b.dtype == np.float64 else b)
