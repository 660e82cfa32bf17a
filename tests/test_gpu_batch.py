"""The ensemble runner (hyd_run_batch; no reference counterpart: the
reference runs one flow per program): flows of different sizes, seeds and
precisions advanced together give exactly what each gives alone with
hyd_run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLOWS = [(32, 3, 0), (64, 5, 0), (32, 11, 1), (64, 7, 1), (32, 7, 0)]


def _make(ng, seed, prec):
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    h = Hydra(ng, precision=prec)
    h.set_field(random_pv(ng, seed=seed))
    return h


def _state(h):
    c = h.contours()
    return c, h.field("qs"), h.scalar("t")


def test_batch_equals_single_runs():
    steps = 40
    alone = []
    for ng, seed, prec in FLOWS:
        h = _make(ng, seed, prec)
        h.run(steps)
        alone.append(_state(h))
        h.close()
    from hydra_amd import Hydra
    hs = [_make(ng, seed, prec) for ng, seed, prec in FLOWS]
    Hydra.run_batch(hs, steps)
    for (ng, seed, prec), h, (c1, qs1, t1) in zip(FLOWS, hs, alone):
        c2, qs2, t2 = _state(h)
        for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
            assert np.array_equal(c1[k], c2[k]), (ng, seed, prec, k)
        assert np.array_equal(qs1, qs2), (ng, seed, prec)
        assert t1 == t2
        h.close()


def test_batch_rejects_repeated_context():
    from hydra_amd import Hydra, HydraError
    h = _make(32, 3, 0)
    with pytest.raises(HydraError):
        Hydra.run_batch([h, h], 2)
    h.close()
