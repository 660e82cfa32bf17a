"""Pins of the oracle's diagnostics: spec1d (spectral.f90:370-424), l1norm /
l2norm (generic.f90:30-64) and energy (evolution.f90:612-644).

  * against the reference build itself (oracle/_ref), bit for bit, where it
    is present, and against the committed golden vectors of that build;
  * against closed forms: a single Fourier mode lands in its one shell
    k = nint(|k| / delk) with the Parseval weight; 4 * sum(spec) equals the
    physical sum of squares for a field without Nyquist content; the norms of
    constant and single-mode fields.
"""
import os

import numpy as np
import pytest

from oracle.hyo import Oracle
from tests.helpers import bits_equal

NG = 32
L = 2 * np.pi
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_ng32.npz")


def grid(ng=NG):
    x = -np.pi + L / ng * np.arange(ng)
    return np.meshgrid(x, x, indexing="ij")  # physical (ny,nx): numpy [ix, iy]


@pytest.mark.parametrize("mx,my,amp", [(3, 4, 1.5), (5, 0, 2.0), (0, 7, 0.5), (6, 6, 3.0)])
def test_spec1d_single_mode(mx, my, amp):
    """f = A cos(mx x + my y): the spectrum is zero except at shell
    nint(sqrt(mx^2 + my^2) / sqrt 2) (delk = sqrt 2 for L = 2 pi), where it
    holds sum(f^2) / 4 = A^2 N / 8."""
    o = Oracle(NG)
    X, Y = grid()
    f = amp * np.cos(mx * X + my * Y)
    spec, _ = o.spec1d(o.ptospc(f), 0)
    k = int(np.floor(np.hypot(mx, my) / np.sqrt(2) + 0.5))
    assert spec[k] == pytest.approx(amp * amp * NG * NG / 8, rel=1e-12)
    others = np.delete(spec, k)
    assert np.abs(others).max() < 1e-12 * spec[k]


def test_spec1d_parseval_and_iopt():
    """4 sum(spec) = sum(f^2) when the field has no Nyquist content; iopt = 1
    multiplies ss by rksq first (and returns it so)."""
    o = Oracle(NG)
    rng = np.random.default_rng(4)
    f = rng.standard_normal((NG, NG))
    ss = o.ptospc(f)
    ss[NG // 2, :] = 0.0  # the Nyquist row/column of the HC layout
    ss[:, NG // 2] = 0.0
    phys = o.spctop(ss)
    spec, s0 = o.spec1d(ss, 0)
    assert bits_equal(s0, ss)
    assert 4 * spec.sum() == pytest.approx(np.sum(phys * phys), rel=1e-12)
    spec1, s1 = o.spec1d(ss, 1)
    rksq = o.d("rksq")
    assert bits_equal(s1, ss * rksq)
    spec1b, _ = o.spec1d(ss * rksq, 0)
    assert bits_equal(spec1, spec1b)


def test_norms_closed_form():
    o = Oracle(NG)
    X, Y = grid()
    garea = (L / NG) ** 2
    c = -2.75
    l1, l2 = o.norms(np.full((NG, NG), c))
    assert l1 == pytest.approx(abs(c) * L * L, rel=1e-13)
    assert l2 == pytest.approx(c * c * L * L, rel=1e-13)
    amp = 1.7
    l1, l2 = o.norms(amp * np.sin(3 * X))
    assert l2 == pytest.approx(amp * amp * L * L / 2, rel=1e-12)
    assert l1 == pytest.approx(garea * np.abs(amp * np.sin(3 * X)).sum(), rel=1e-13)


def test_diag_against_golden():
    """The golden vectors written by the reference build (tests/golden)."""
    g = np.load(GOLDEN)
    o = Oracle(NG)
    for iopt in (0, 1):
        spec, ss = o.spec1d(g["fft_s"], iopt)
        kmax = int(g["spec_kmax"])
        assert bits_equal(spec[:kmax + 1], g[f"spec{iopt}"][:kmax + 1]), iopt
        assert bits_equal(ss, g[f"spec{iopt}_ss"]), iopt
    l1, l2 = o.norms(g["fft_f"])
    assert l1 == float(g["norm_l1"]) and l2 == float(g["norm_l2"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_diag_against_reference_build(seed):
    from oracle.hyo import Ref, ref_available
    if not ref_available(NG):
        pytest.skip("reference build oracle/_ref not present")
    r = Ref(NG)
    o = Oracle(NG)
    rng = np.random.default_rng(seed)
    ss = rng.standard_normal((NG, NG)) * 10.0 ** rng.integers(-3, 4)
    for iopt in (0, 1):
        a, sa = o.spec1d(ss, iopt)
        b, sb = r.spec1d(ss, iopt)
        assert bits_equal(a, b) and bits_equal(sa, sb), iopt
    assert o.norms(ss) == r.norms(ss)


def test_energy_two_layer_closed_form():
    """kd > 0: epot = (1/2) kd^2 int psi^2 with psi = -A/(k^2 + kd^2) sin(kx)
    for q = A sin(kx); ekin = (1/2) int |u|^2."""
    X, Y = grid()
    k, amp, kd = 2, 1.3, 2.0
    o = Oracle(NG, kd=kd)
    o.init_from_field(amp * np.sin(k * X))
    o.set_contours(dict(x=np.zeros(0), y=np.zeros(0), nextq=np.zeros(0, np.int32),
                        **{n: np.zeros(0, np.int32) for n in ("i1q", "i2q", "npq", "indq")},
                        qjump=1.0))
    o.evolution_init()
    o.L.hyo_inversion(o.m)
    ek, ep = o.energy()
    a = amp / (k * k + kd * kd)  # psi amplitude
    filt = np.exp(-36.0 * (k / (NG / 2)) ** 36)  # Hou-Li filter of the derivative
    assert ep == pytest.approx(0.5 * kd * kd * a * a * L * L / 2, rel=1e-12)
    assert ek == pytest.approx(0.5 * (a * k * filt) ** 2 * L * L / 2, rel=1e-12)
