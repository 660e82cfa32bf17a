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


@pytest.mark.parametrize("ng,seed,prec", [(64, 5, 0), (32, 7, 0), (32, 11, 1)])
def test_run_without_waits_equals_single_steps(ng, seed, prec):
    """hyd_run does not wait for the GPU between its steps; the result, the
    monitor and complexity rows and the advance counts are those of the
    same loop taken one call (and one wait) per step."""
    steps = 40
    h1, h2 = _make(ng, seed, prec), _make(ng, seed, prec)
    h1.run(steps)
    for _ in range(steps):
        h2.run(1)
    c1, qs1, t1 = _state(h1)
    c2, qs2, t2 = _state(h2)
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert np.array_equal(c1[k], c2[k]), k
    assert np.array_equal(qs1, qs2)
    assert t1 == t2
    for kind in ("monitor", "complexity"):
        assert np.array_equal(h1.take_records(kind), h2.take_records(kind)), kind
    a1, u1 = h1.advance_counts()
    a2, u2 = h2.advance_counts()
    assert (a1, u1) == (a2, u2) and a1 == steps
    h1.close()
    h2.close()
