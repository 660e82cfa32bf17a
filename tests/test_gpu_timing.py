"""Per-launch timing (hyd_set_timing) runs the step on one stream in program
order: the results must be the two-stream step's, bit for bit, and the
timing records must be there."""
import numpy as np
import pytest

from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def test_timed_step_is_bit_identical():
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    qq = random_pv(64, seed=31)
    hs = []
    for timing in (0, 1):
        h = Hydra(64)
        h.set_field(qq)
        h.recontour()
        h.evolution_init()
        h.set_timing(timing)
        for _ in range(3):
            h.advance()
        hs.append(h)
    for name in ("qs", "qd", "qq", "uu", "vv", "pp"):
        assert bits_equal(hs[0].field(name), hs[1].field(name)), name
    a, b = hs[0].contours(), hs[1].contours()
    for k in ("x", "y", "nextq"):
        assert bits_equal(a[k], b[k]), k
    recs = hs[1].kernel_times()
    assert any(n.startswith("k:fft") for n, _, _ in recs)
    assert all(ms >= 0.0 for _, ms, _ in recs)
    assert hs[1].event_overhead(20) > 0.0
    for h in hs:
        h.close()
