"""GPU parity of contour surgery (contours.f90:350-683) and of the advect
loop that calls it (evolution.f90:55-80, caps.f90:83-91): the device result
must be bit-identical to the oracle's."""
import numpy as np
import pytest

from helpers import TANGLES, bits_equal, tangle_contours

pytestmark = pytest.mark.gpu


def _same_contours(g, o):
    for k in ("nextq", "i1q", "i2q", "npq", "indq"):
        assert np.array_equal(g[k], o[k]), k
    assert bits_equal(g["x"], o["x"]) and bits_equal(g["y"], o["y"])


@pytest.mark.parametrize("seed,ncirc,rmin,rmax", TANGLES)
def test_surgery_tangle_bitwise(seed, ncirc, rmin, rmax):
    """Crossing circles of one level (many order-dependent reconnections),
    several levels, and tiny circles that surgery deletes."""
    from hydra_amd import Hydra
    from oracle.hyo import Oracle
    ng = 32
    o = Oracle(ng)
    rng = np.random.default_rng(seed)
    c = tangle_contours(o.getd("dm"), rng, ncirc, levels=(-2, 1, 3), rmin=rmin, rmax=rmax)
    o.set_contours(c)
    h = Hydra(ng)
    h.set_field(np.zeros((ng, ng)))
    h.set_contours(c)
    o.surgery()
    h.surgery()
    _same_contours(h.contours(), o.contours())
    h.close()


@pytest.mark.parametrize("ng", [32, 64])
def test_surgery_after_evolution_bitwise(ng):
    """Surgery on contours advected until twist exceeds 2.5 (the state in
    which evolution.f90:58-70 calls it)."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    qq = random_pv(ng, seed=11)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    o.evolution_init()
    while o.getd("twist") <= 2.5:
        o.advance()
    c = o.contours()
    h = Hydra(ng)
    h.set_field(qq)
    h.set_contours(c)
    o.surgery()
    h.surgery()
    g, r = h.contours(), o.contours()
    assert len(r["i1q"]) != len(c["i1q"])  # the topology did change
    _same_contours(g, r)
    h.close()


def test_run_with_surgery_and_recontouring_bitwise():
    """160 steps of caps.f90:83-91 at ng=32: about twenty surgeries and a
    recontouring when the regularisation count reaches 20."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    ng = 32
    qq = random_pv(ng, seed=7)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.run(160)
    h = Hydra(ng)
    h.set_field(qq)
    h.run(160)
    assert h.scalar("t") == o.getd("t")
    _same_contours(h.contours(), o.contours())
    for name in ("qs", "qd", "qc"):
        assert bits_equal(h.field(name), o.d(name)), name
    h.close()


@pytest.mark.parametrize("split", [37, 160])
def test_run_in_two_calls_bitwise(split):
    """run(a) then run(b) equals one run of a+b steps, on the device and in
    the oracle alike (the second call resumes inside advect's loop; split
    160 ends the first call just before the 20th regularisation hands over
    to recontouring), and a batch run split the same way agrees too."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    ng, n = 32, 170
    qq = random_pv(ng, seed=7)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.run(n)
    h = Hydra(ng)
    h.set_field(qq)
    h.run(split)
    h.run(n - split)
    assert h.scalar("t") == o.getd("t")
    _same_contours(h.contours(), o.contours())
    for name in ("qs", "qd", "qc"):
        assert bits_equal(h.field(name), o.d(name)), name
    h.close()
    b = Hydra(ng)
    b.set_field(qq)
    Hydra.run_batch([b], split)
    Hydra.run_batch([b], n - split)
    assert b.scalar("t") == o.getd("t")
    _same_contours(b.contours(), o.contours())
    b.close()
