"""Restart from saved contours (common.f90 readcont + caps.f90:185-189) and
the grid save (evolution.f90:646-731) on the GPU, against the oracle."""
import os

import numpy as np
import pytest

from hydra_amd import io
from tests.helpers import bits_equal, oracle_state

pytestmark = pytest.mark.gpu

NG = 32


def saved(tmp_path):
    o, qq, c = oracle_state(NG)
    resid = np.random.default_rng(3).standard_normal((NG, NG)) * 0.1
    t = 0.75
    io.savecont(str(tmp_path), 1, t, c, resid)
    return c, resid, t


def test_restart_then_step_bitwise(tmp_path):
    """The restored state, then evolution_init and two steps, bit for bit
    against the oracle restored the same way."""
    from hydra_amd import Hydra
    from oracle.hyo import Oracle
    saved(tmp_path)
    c, qr, t = io.readcont(str(tmp_path), 1, NG)
    h = Hydra(NG)
    assert io.restart(h, str(tmp_path), 1) == t
    o = Oracle(NG)
    o.set_contours(c)
    qa = o.con2grid() + qr
    o.init_from_field(qa, t)
    o.d("qr")[:] = qr
    o.set_contours(c)
    assert bits_equal(h.field("qs"), o.d("qs")) and bits_equal(h.field("qr"), qr)
    hc = h.contours()
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert bits_equal(hc[k], c[k]), k
    assert h.scalar("t") == t and h.scalar("qjump") == c["qjump"] == o.getd("qjump")
    o.evolution_init()
    h.evolution_init()
    for step in range(2):
        o.advance()
        h.advance()
        for name in ("qs", "qd", "qq", "uu", "vv"):
            assert bits_equal(h.field(name), o.d(name)), (step, name)
    h.close()


def test_savegrid_matches_oracle(tmp_path):
    from hydra_amd import Hydra
    from oracle.hyo import Oracle
    from hydra_amd.ic import random_pv
    qq = random_pv(NG, seed=5)
    o = Oracle(NG)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    c = o.contours()
    o.evolution_init()
    h = Hydra(NG)
    h.set_field(qq)
    h.set_contours(c)
    h.evolution_init()
    o.advance()
    h.advance()
    kmax, kmaxred = o.get("kmax"), o.get("kmaxred")
    assert list(h.table("shells")) == [kmax, kmaxred, o.getd("dsumi")]
    assert bits_equal(h.table("spmf")[:kmax + 1], o.d("spmf", NG + 1)[:kmax + 1])
    assert bits_equal(h.table("alk")[1:kmax + 1], o.d("alk", NG + 1)[1:kmax + 1])
    out = str(tmp_path / "run")
    assert io.savegrid(h, out, 0) == 1
    o.inversion()   # savegrid saves the state at t (evolution.f90:87-88)
    for name in ("qq", "uu", "vv", "pp"):
        assert bits_equal(h.field(name), o.d(name)), name
    t = h.scalar("t")
    ek, ep = o.energy()
    assert open(os.path.join(out, "ene.asc")).read() == io.fmt_ene(t, ek, ep) + "\n"
    wka = o.spctop(o.d("qq").copy())
    tr4, f4 = io.read_r4_record(os.path.join(out, "qq.r4"), 1, NG)
    assert tr4 == float(np.float32(t))
    assert bits_equal(f4, (wka + o.d("bety", NG)[None, :]).astype(np.float32))
    ql2 = o.norms(wka)[1]
    assert open(os.path.join(out, "norms.asc")).read() == io.fmt_norms(t, ql2) + "\n"
    spec, _ = o.spec1d(o.d("qq").copy(), 0)
    spmf = o.d("spmf", NG + 1)
    s = 0.0
    for k in range(1, kmax + 1):
        s = s + spec[k]
        spec[k] = spmf[k] * spec[k]
    s = 8.0 * s * o.getd("dsumi")
    lines = open(os.path.join(out, "spectra.asc")).read().splitlines()
    assert lines[0] == io.fmt_spectra_head(t, s, ql2, kmaxred)
    alk = o.d("alk", NG + 1)
    assert lines[1:] == [io.fmt_spectra_line(alk[k], np.log10(spec[k]))
                         for k in range(1, kmaxred + 1)]
    h.close()
