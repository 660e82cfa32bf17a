"""Pins of the oracle's time step (evolution.f90: adapt 524-608, source
432-490, inversion 494-520, advance 166-326, init 96-131, prepare 135-162,
energy 612-644), which the reference build cannot check (evolution.f90
does not compile here, see oracle/Makefile):

  * adapt, source and energy against closed forms on single- and two-mode
    fields (Fourier modes pass the inversion, the Hou-Li filter and the
    derivative tables unchanged at these wavenumbers);
  * advance against textbook RK4 (cdamp = 0) and Lawson's integrating-factor
    RK4 (cdamp > 0), driven step by step from Python through the oracle's
    own right-hand side (inversion, source at lev 0, velint);
  * init/prepare/inversion as the compositions of pinned pieces they are.
"""
import numpy as np
import pytest

from oracle.hyo import Oracle

NG = 32
L = 2 * np.pi


def grid(ng=NG):
    x = -np.pi + L / ng * np.arange(ng)
    # physical fields are (ny,nx) column-major: numpy [ix, iy]
    return np.meshgrid(x, x, indexing="ij")


def empty_contours(qjump=1.0):
    return dict(x=np.zeros(0), y=np.zeros(0), nextq=np.zeros(0, np.int32),
                **{k: np.zeros(0, np.int32) for k in ("i1q", "i2q", "npq", "indq")}, qjump=qjump)


def model(q, **kw):
    o = Oracle(NG, **kw)
    o.init_from_field(q)
    o.set_contours(empty_contours())
    o.evolution_init()
    return o


@pytest.mark.parametrize("k,amp,tsave", [(1, 4 * np.pi, 1000.0), (8, 4 * np.pi, 1000.0),
                                        (1, 0.5, 0.01)])
def test_adapt_closed_form(k, amp, tsave):
    """q = A sin(kx): zeta = q, |u| = (A/k)f(k)|cos kx| with the Hou-Li
    filter f; dt follows the CFL limit
    (k=1), the vorticity limit (k=8) or dtmax = 0.2 min(tgsave,tcsave)."""
    X, Y = grid()
    q = amp * np.sin(k * X)
    o = model(q, tsave=tsave)
    o.L.hyo_inversion(o.m)
    o.L.hyo_adapt(o.m)
    zzrms = amp / np.sqrt(2.0)
    zzmax = np.abs(q).max() + 1e-12
    # derivatives use the Hou-Li filtered wavenumber (spectral.f90:67-70)
    filt = np.exp(-36.0 * (k / (NG / 2)) ** 36)
    uumax = amp / k * filt * np.abs(np.cos(k * X)).max() + 1e-12
    glx = L / NG
    srwfm = o.getd("srwfm")
    dt = min(glx * 0.7 / uumax, (np.pi / 10) / max(zzmax, srwfm), 0.2 * tsave)
    assert o.getd("zzrms") == pytest.approx(zzrms, rel=1e-13)
    assert o.getd("zzmax") == pytest.approx(zzmax, rel=1e-13)
    assert o.getd("uumax") == pytest.approx(uumax, rel=1e-13)
    assert o.getd("dt") == pytest.approx(dt, rel=1e-13)
    assert o.getd("dt2") == pytest.approx(dt / 2, rel=1e-15)
    assert o.getd("dt3") == pytest.approx(dt / 3, rel=1e-15)
    assert o.getd("dt6") == pytest.approx(dt / 6, rel=1e-15)
    assert o.getd("twist") == pytest.approx(dt * zzmax, rel=1e-13)
    qdiss = o.d("qdiss")
    assert np.allclose(o.d("epq"), np.exp(dt / 2 * zzrms * qdiss), rtol=1e-13, atol=0)
    assert np.allclose(o.d("emq") * o.d("epq"), 1.0, rtol=1e-15, atol=0)


def test_energy_closed_form():
    """ekin = (1/2) int |u|^2 = L^2 A^2 / (4 k^2) for psi = (A/k^2) sin(kx)."""
    X, Y = grid()
    k, amp = 3, 2.0
    o = model(amp * np.sin(k * X))
    o.L.hyo_inversion(o.m)
    ek, ep = o.energy()
    assert ek == pytest.approx(L * L * amp * amp / (4 * k * k), rel=1e-12)
    assert ep == 0.0


def test_source_closed_form():
    """q = a sin x + 4b cos 2y, so psi = -(a sin x + b cos 2y) (green =
    -1/k^2: q is the Laplacian of psi), u = -psi_y = -2b sin 2y,
    v = psi_x = -a cos x and -(u q_x + v q_y) = -6ab cos x sin 2y; for
    qd = fhi*q (no contours) the same with the per-mode fhi weights; lev 1/2
    multiply sqd by epq/epq^2."""
    X, Y = grid()
    a, b = 1.3, -0.7
    q = a * np.sin(X) + 4 * b * np.cos(2 * Y)
    o = model(q)
    o.L.hyo_inversion(o.m)
    o.L.hyo_adapt(o.m)
    fhi = o.d("fhi")   # spectral (nx,ny): numpy [ky, kx]
    fhi1, fhi2 = fhi[0, 1], fhi[2, 0]  # modes (kx=1, ky=0) and (kx=0, ky=2)
    sqs, sqd = np.empty((NG, NG)), np.empty((NG, NG))
    o.L.hyo_source(o.m, sqs, sqd, 0)
    exp_s = o.ptospc(-6 * a * b * np.cos(X) * np.sin(2 * Y))
    exp_d = o.ptospc((2 * fhi1 - 8 * fhi2) * a * b * np.cos(X) * np.sin(2 * Y))
    scale = abs(a * b) * 8
    assert np.abs(sqs - exp_s).max() < 1e-13 * scale
    assert np.abs(sqd - exp_d).max() < 1e-13 * scale
    epq = o.d("epq").copy()
    s1, d1 = np.empty((NG, NG)), np.empty((NG, NG))
    o.L.hyo_source(o.m, s1, d1, 1)
    assert np.array_equal(s1, sqs) and np.array_equal(d1, epq * sqd)
    o.L.hyo_source(o.m, s1, d1, 2)
    assert np.array_equal(s1, sqs) and np.array_equal(d1, epq ** 2 * sqd)


def test_init_prepare_inversion_are_their_compositions():
    """evolution_init: qc = ptospc(con2grid), qd = fhi (qs-qc); inversion:
    qq = flo (qs-qc) + qc + qd then main_invert; prepare: qs <- qq,
    qr = spctop(qs - qc)."""
    from helpers import oracle_state
    o0, qq, c = oracle_state(NG)
    o = Oracle(NG)
    o.init_from_field(qq)
    o.set_contours(c)
    o.evolution_init()
    qc = o.ptospc(o.con2grid())
    assert np.array_equal(o.d("qc"), qc)
    qs = o.d("qs").copy()
    assert np.array_equal(o.d("qd"), o.d("fhi") * (qs - qc))
    # perturb qd so that the three terms of qq all matter
    o.d("qd")[:] = o.d("qd") + 1e-3 * np.cos(np.arange(NG * NG)).reshape(NG, NG)
    qd = o.d("qd").copy()
    o.L.hyo_inversion(o.m)
    qq_s = o.d("flo") * (qs - qc) + qc + qd
    assert np.array_equal(o.d("qq"), qq_s)
    uu, vv, pp = o.main_invert(qq_s)
    assert np.array_equal(o.d("uu"), uu) and np.array_equal(o.d("vv"), vv)
    o.L.hyo_prepare(o.m)
    assert np.array_equal(o.d("qs"), qq_s)
    assert np.array_equal(o.d("qr"), o.spctop(qq_s - qc))


def _rhs(o, qs, qd, x, y):
    """Right-hand side of advance: (sqs, sqd, uq, vq) at the given state."""
    o.d("qs")[:] = qs
    o.d("qd")[:] = qd
    n = len(x)
    o.d("xq", n)[:] = x
    o.d("yq", n)[:] = y
    o.L.hyo_inversion(o.m)
    sqs, sqd = np.empty((NG, NG)), np.empty((NG, NG))
    o.L.hyo_source(o.m, sqs, sqd, 0)
    uq, vq = o.velint(o.d("uu"), o.d("vv"))
    return sqs, sqd, uq, vq


def _wrap(v):
    oms = 1.0 - 1.0e-12
    return oms * (v - L * np.trunc(v / np.pi))


@pytest.mark.parametrize("cdamp", [0.0, 2.0])
def test_advance_is_lawson_rk4(cdamp):
    """One advance from a state just after init equals one step of
    textbook RK4 for (qs, nodes) and Lawson's integrating-factor RK4 for qd
    (factor E = exp(-dt/2 zzrms qdiss) per half step), built on the oracle's
    own right-hand side.  A wrong stage weight, factor power or stage time
    breaks the agreement by O(dt^2)."""
    from hydra_amd.ic import random_pv
    q = random_pv(NG, seed=5, k0=4.0, rms=2.0)
    ref = Oracle(NG)
    ref.init_from_field(q)
    ref.recontour(ref.d("qr").copy())
    c = ref.contours()
    c["qjump"] = c["qjump"] * 0.5  # contour PV differs from the field's
    kw = dict(dt_fixed=0.05, cdamp=cdamp)
    o = Oracle(NG, **kw)
    o.init_from_field(q)
    o.set_contours(c)
    o.evolution_init()
    n = len(c["x"])
    qs0, qd0 = o.d("qs").copy(), o.d("qd").copy()
    x0, y0 = o.d("xq", n).copy(), o.d("yq", n).copy()
    o.advance()
    qs1, qd1 = o.d("qs").copy(), o.d("qd").copy()
    x1, y1 = o.d("xq", n).copy(), o.d("yq", n).copy()
    E = 1.0 / o.d("epq")   # factor of one half step, from this step's adapt
    dt = 0.05

    w = Oracle(NG, **kw)
    w.init_from_field(q)
    w.set_contours(c)
    w.evolution_init()
    k1 = _rhs(w, qs0, qd0, x0, y0)
    s2 = (qs0 + dt / 2 * k1[0], E * (qd0 + dt / 2 * k1[1]),
          _wrap(x0 + dt / 2 * k1[2]), _wrap(y0 + dt / 2 * k1[3]))
    k2 = _rhs(w, *s2)
    k2 = (k2[0], k2[1] / E, k2[2], k2[3])
    s3 = (qs0 + dt / 2 * k2[0], E * (qd0 + dt / 2 * k2[1]),
          _wrap(x0 + dt / 2 * k2[2]), _wrap(y0 + dt / 2 * k2[3]))
    k3 = _rhs(w, *s3)
    k3 = (k3[0], k3[1] / E, k3[2], k3[3])
    s4 = (qs0 + dt * k3[0], E * E * (qd0 + dt * k3[1]),
          _wrap(x0 + dt * k3[2]), _wrap(y0 + dt * k3[3]))
    k4 = _rhs(w, *s4)
    k4 = (k4[0], k4[1] / (E * E), k4[2], k4[3])
    comb = [(a + 2 * b + 2 * cc + d) / 6 for a, b, cc, d in zip(k1, k2, k3, k4)]
    qs_e = qs0 + dt * comb[0]
    qd_e = E * E * (qd0 + dt * comb[1])
    x_e = _wrap(x0 + dt * comb[2])
    y_e = _wrap(y0 + dt * comb[3])

    def per(a, b):
        d = a - b
        return np.abs(d - L * np.round(d / L)).max()

    sc = np.abs(qs0).max()
    assert np.abs(qs1 - qs_e).max() < 1e-12 * sc
    assert np.abs(qd1 - qd_e).max() < 1e-12 * sc
    assert per(x1, x_e) < 1e-10 and per(y1, y_e) < 1e-10
    # the step did move things by far more than the tolerance
    assert np.abs(qs1 - qs0).max() > 1e-4 * sc and per(x1, x0) > 1e-3
    if cdamp > 0:
        assert np.abs(qd1 - qd0).max() > 1e-4 * sc
