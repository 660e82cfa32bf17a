"""The contour-table checks that guard the device store (hyd_set_contours,
and io.restart which passes file contents through it), run on the host
through the C ABI without a device: a well-formed table passes, and each
way of breaking the store's invariants (contours.f90:12-15,
common.f90:78-88, the level order surgery relies on, contours.f90:372-394)
is rejected with HYD_EARG before anything reaches the device."""
import numpy as np
import pytest

from hydra_amd import HydraError, check_contours
from tests.helpers import TANGLES, tangle_contours

NG = 32


@pytest.fixture()
def good():
    seed, ncirc, rmin, rmax = TANGLES[0]
    dm = 0.0625 * 2 * np.pi / NG
    c = tangle_contours(dm, np.random.default_rng(seed), 3, (-1, 2), rmin, rmax)
    check_contours(c, NG)
    return c


def broken(c, **changes):
    d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    for k, f in changes.items():
        d[k] = f(d[k])
    return d


def test_well_formed_and_empty_pass(good):
    empty = dict(x=np.zeros(0), y=np.zeros(0), nextq=np.zeros(0, np.int32),
                 **{k: np.zeros(0, np.int32) for k in ("i1q", "i2q", "npq", "indq")})
    check_contours(empty, NG)
    check_contours(good, NG)


def _set(a, i, v):
    a[i] = v
    return a


CASES = {
    "i2q past nptq": dict(i2q=lambda a: _set(a, -1, a[-1] + 5)),
    "first contour not at node 1": dict(i1q=lambda a: a + 1),
    "gap between contours": dict(i1q=lambda a: _set(a, 1, a[1] + 1)),
    "npq inconsistent": dict(npq=lambda a: _set(a, 0, a[0] + 1)),
    "levels descending": dict(indq=lambda a: a[::-1].copy()),
    "nextq leaves its contour": dict(nextq=lambda a: _set(a, 0, a[-1])),
    "nextq two cycles": dict(nextq=lambda a: _set(a, 1, 2)),
    "nextq out of range": dict(nextq=lambda a: _set(a, 3, 0)),
    "node at xmax": dict(x=lambda a: _set(a, 0, np.pi)),
    "node below ymin": dict(y=lambda a: _set(a, 2, -np.pi - 1e-9)),
    "NaN node": dict(x=lambda a: _set(a, 4, np.nan)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_malformed_table_rejected(good, name):
    with pytest.raises(HydraError, match="error -1"):
        check_contours(broken(good, **CASES[name]), NG)


def test_count_mismatch_rejected(good):
    c = dict(good, i1q=good["i1q"][:0], i2q=good["i2q"][:0], npq=good["npq"][:0],
             indq=good["indq"][:0])
    with pytest.raises(HydraError):
        check_contours(c, NG)
