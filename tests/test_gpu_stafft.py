"""forfft / revfft (stafft.f90:187-406) and xderiv / yderiv
(sta2dfft.f90:169-245) through the C ABI on the device, bit for bit against
the oracle, at lengths other than the context's grid and with odd vector
counts."""
import numpy as np
import pytest

from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

CASES = [(1, 8), (3, 32), (17, 64), (64, 256), (5, 1024), (2, 4096), (1024, 1024), (7, 2048)]


@pytest.fixture(scope="module")
def ctx():
    from hydra_amd import Hydra
    h = Hydra(64)
    yield h
    h.close()


@pytest.mark.parametrize("m,n", CASES)
def test_forfft_revfft_bitwise(ctx, m, n):
    from hydra_amd import initfft
    from oracle.hyo import lib as olib
    f, t = initfft(n)
    x = np.random.default_rng(m * 7919 + n).standard_normal(m * n) * 3.0
    got_f = ctx.forfft(m, n, x, t, f)
    want_f = x.copy()
    olib().hyo_forfft(m, n, want_f, t, f)
    assert bits_equal(got_f, want_f)
    got_r = ctx.revfft(m, n, got_f, t, f)
    want_r = want_f.copy()
    olib().hyo_revfft(m, n, want_r, t, f)
    assert bits_equal(got_r, want_r)
    assert np.allclose(got_r, x, rtol=0, atol=1e-12 * np.abs(x).max())


def test_forfft_rejects_wrong_factors(ctx):
    from hydra_amd import HydraError, initfft
    f, t = initfft(64)
    f = f.copy()
    f[2] += 1
    with pytest.raises(HydraError, match="error -2"):
        ctx.forfft(4, 64, np.zeros(256), t, f)


def test_xderiv_yderiv_bitwise(ctx):
    from oracle.hyo import Oracle, lib as olib
    o = Oracle(64)
    s = np.random.default_rng(5).standard_normal((64, 64))
    for name, fn, k in (("x", ctx.xderiv, "hrkx"), ("y", ctx.yderiv, "hrky")):
        want = np.empty_like(s)
        getattr(olib(), f"hyo_{name}deriv")(64, 64, o.d(k, 64), s, want)
        assert bits_equal(fn(s), want), name
