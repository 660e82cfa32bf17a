"""The oracle against the reference build itself (oracle/_ref), on fresh
seeded inputs.  Skipped where the reference build is absent (it is built in
the container that holds /root/reference)."""
import numpy as np
import pytest

from tests.helpers import TANGLES, bits_equal, tangle_contours

NG = 32
_PAIRS = {}


def make_pair(ng):
    from oracle.hyo import Oracle, Ref, ref_available
    if not ref_available(ng):
        pytest.skip("reference build oracle/_ref not present")
    if ng not in _PAIRS:
        r = Ref(ng)
        o = Oracle(ng)
        o.set("d4small", float(r.consts()[2]))
        _PAIRS[ng] = (o, r)
    return _PAIRS[ng]


@pytest.fixture(scope="module")
def pair():
    return make_pair(NG)


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


def same_contours(a, b):
    return all(bits_equal(a[k], b[k]) for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"))


@pytest.mark.parametrize("ng", [32, 64])
def test_recontour_with_contours(ng):
    """congen.f90:40-61 with contours present: con2ugrid (congen.f90:338-441,
    its own index conventions: int() without +1, its own ixbeg) plus the
    bilinear residual, then ugrid2con.  Two inputs: freshly made contours
    sheared by a smooth displacement with a seeded residual, and the
    contours and residual qr = spctop(qs - qc) that prepare
    (evolution.f90:135-162) leaves after the oracle advanced the flow."""
    from oracle.hyo import Oracle
    from tests.helpers import oracle_state
    o, r = make_pair(ng)
    src, qq, c0 = oracle_state(ng)
    qj = src.getd("qjump")
    x = c0["x"] + 0.04 * np.sin(2 * c0["y"] + 1.0)
    y = c0["y"] + 0.04 * np.cos(3 * c0["x"])
    L = 2 * np.pi
    x = (1 - 1e-12) * (x - L * np.trunc(x / np.pi))
    y = (1 - 1e-12) * (y - L * np.trunc(y / np.pi))
    res = np.random.default_rng(ng).standard_normal((ng, ng)) * 0.3 * qj
    cases = [(dict(c0, x=x, y=y, qjump=qj), res)]
    # advected: 6 RK4 steps from the seeded field, then prepare
    a = Oracle(ng)
    a.init_from_field(qq)
    a.recontour(a.d("qr").copy())
    a.advect_start()
    for _ in range(6):
        a.advance()
    a.prepare()
    cases.append((dict(a.contours(), qjump=a.getd("qjump")), a.d("qr").copy()))
    for c, qr in cases:
        o.set_contours(c)
        r.set_contours(c)
        o.recontour(qr)
        r.recontour(qr)
        assert o.get("nq") > 0
        assert same_contours(o.contours(), r.contours())


@pytest.mark.parametrize("ng", [32, 64])
def test_contint(ng):
    """generic.f90:66-79 against the oracle's set-up, caps.f90:166."""
    from oracle.hyo import Oracle
    _, r = make_pair(ng)
    for seed in (1, 2):
        qq = np.random.default_rng(seed).standard_normal((ng, ng)) * 7.0
        o = Oracle(ng, ncontq=80)
        o.init_from_field(qq)
        assert bits_equal(np.float64(o.getd("qjump")), np.float64(r.contint(qq, 80)))
