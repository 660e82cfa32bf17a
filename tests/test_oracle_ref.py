"""The oracle against the reference build itself (oracle/_ref), on fresh
seeded inputs.  Skipped where the reference build is absent (it is built in
the container that holds /root/reference)."""
import numpy as np
import pytest

from tests.helpers import TANGLES, bits_equal, tangle_contours

NG = 32


@pytest.fixture(scope="module")
def pair():
    from oracle.hyo import Oracle, Ref, ref_available
    if not ref_available(NG):
        pytest.skip("reference build oracle/_ref not present")
    r = Ref(NG)
    o = Oracle(NG)
    o.set("d4small", float(r.consts()[2]))
    return o, r


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fft_and_inversion(pair, seed):
    o, r = pair
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((NG, NG)) * 10.0 ** rng.integers(-3, 4)
    assert bits_equal(o.ptospc(f), r.ptospc(f))
    assert bits_equal(o.spctop(f), r.spctop(f))
    for a, b in zip(o.main_invert(f), r.main_invert(f)):
        assert bits_equal(a, b)


@pytest.mark.parametrize("seed", [11, 12])
def test_recontour_con2grid_surgery(pair, seed):
    from hydra_amd.ic import random_pv
    o, r = pair
    qq = random_pv(NG, seed=seed, k0=4.0)
    qj = (qq.max() - qq.min()) / 40
    empty = dict(x=np.zeros(0), y=np.zeros(0), nextq=np.zeros(0, np.int32),
                 **{k: np.zeros(0, np.int32) for k in ("i1q", "i2q", "npq", "indq")}, qjump=qj)
    o.set_contours(empty)
    r.set_contours(empty)
    o.recontour(qq)
    r.recontour(qq)
    a, b = o.contours(), r.contours()
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert bits_equal(a[k], b[k]), k
    assert bits_equal(o.con2grid(), r.con2grid())
    # shear the contours so that surgery has work to do
    x = a["x"] + 0.05 * np.sin(2 * a["y"] + seed)
    y = a["y"] + 0.05 * np.cos(3 * a["x"])
    L = 2 * np.pi
    x = (1 - 1e-12) * (x - L * np.trunc(x / np.pi))
    y = (1 - 1e-12) * (y - L * np.trunc(y / np.pi))
    c = dict(a, x=x, y=y, qjump=qj)
    o.set_contours(c)
    r.set_contours(c)
    o.surgery()
    r.surgery()
    a, b = o.contours(), r.contours()
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert bits_equal(a[k], b[k]), k


@pytest.mark.parametrize("seed,ncirc,rmin,rmax", TANGLES)
def test_surgery_tangles(pair, seed, ncirc, rmin, rmax):
    """Surgery on tangles of crossing circles (order-dependent reconnections
    and deletions) against the reference build."""
    o, r = pair
    c = tangle_contours(o.getd("dm"), np.random.default_rng(seed), ncirc, (-2, 1, 3), rmin, rmax)
    o.set_contours(c)
    r.set_contours(c)
    o.surgery()
    r.surgery()
    a, b = o.contours(), r.contours()
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert bits_equal(a[k], b[k]), k
