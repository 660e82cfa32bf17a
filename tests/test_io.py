"""The reference's file formats (hydra_amd/io.py): byte layouts, the
formatted lines, and readcont / qq_init.r8 read back by the reference's own
statements (the amdflang build, oracle/_ref) where that build is present;
the format strings also against golden lines written by that build."""
import os

import numpy as np
import pytest

from hydra_amd import io
from tests.helpers import bits_equal, oracle_state

NG = 32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_ng32.npz")

# values written with each reference format by the golden script
FMT_CASES = [
    (1, (0.5, 1.234567890123, 0.0), ()),
    (1, (12.345, 123.456789, 7.0e-5), ()),
    (1, (1000.0, -0.000123, 1e-12), ()),
    (2, (0.25, 3.14159265358979), ()),
    (2, (99.99, 12345.6789012345), ()),
    (3, (1.0, 0.987654321, 42.5, 0.0), (11,)),
    (4, (0.30103, -2.5), ()),
    (4, (1.0, -3.5e-9), ()),
    (5, (0.125, 0.1234567), (264, 7805)),
    (5, (123.456789, 2.5e-7), (0, 0)),
    (4, (1.0, float("-inf")), ()),
    (4, (0.5, float("nan")), ()),
    (2, (12345678.0, 1.0), ()),
    (6, (0.0,), (264, 7805)),
    (6, (1.23456789,), (1321, 473547)),
    (7, (0.0, 0.0, 78.95683520871, 12.56637061436, 40.5, 3.25, 0.25), ()),
    (7, (0.21698, 0.6999999, 123.4567891, 15.7, 61.234567, 3.0000001, 2.5009), ()),
    (7, (1.5, 12.5, 1.0, 1.0, 1.0, 1.0, 12.25), ()),
]


def python_line(kind, v, iv):
    if kind == 1:
        return io.fmt_ene(v[0], v[1], v[2])
    if kind == 2:
        return io.fmt_norms(v[0], v[1])
    if kind == 3:
        return io.fmt_spectra_head(v[0], v[1], v[2], iv[0])
    if kind == 4:
        return io.fmt_spectra_line(v[0], v[1])
    if kind == 6:
        return io.fmt_complexity(v[0], iv[0], iv[1])
    if kind == 7:
        return io.fmt_monitor(*v[:7])
    return io.fmt_synopsis(iv[0], iv[1], v[0], v[1])


def ref_or_skip():
    from oracle.hyo import Ref, ref_available
    if not ref_available(NG):
        pytest.skip("reference build oracle/_ref not present")
    return Ref(NG)


def test_record_layouts(tmp_path):
    """Direct access: no markers, RECL bytes per record; sequential: the
    byte count before and after the payload."""
    rng = np.random.default_rng(1)
    qq = rng.standard_normal((NG, NG))
    p = str(tmp_path / "qq_init.r8")
    io.write_qq_init(p, 0.75, qq)
    raw = open(p, "rb").read()
    assert len(raw) == 8 * (NG * NG + 1)
    assert np.frombuffer(raw[:8], "<f8")[0] == 0.75
    assert bits_equal(np.frombuffer(raw[8:], "<f8").reshape(NG, NG), qq)
    t, back = io.read_qq_init(p, NG)
    assert t == 0.75 and bits_equal(back, qq)
    p4 = str(tmp_path / "qq.r4")
    io.write_r4_record(p4, 3, 2.5, qq)
    io.write_r4_record(p4, 1, 0.5, -qq)
    assert os.path.getsize(p4) == 3 * 4 * (NG * NG + 1)
    t3, f3 = io.read_r4_record(p4, 3, NG)
    t1, f1 = io.read_r4_record(p4, 1, NG)
    assert (t3, t1) == (2.5, 0.5)
    assert bits_equal(f3, qq.astype(np.float32)) and bits_equal(f1, (-qq).astype(np.float32))
    ps = str(tmp_path / "seq")
    with open(ps, "wb") as f:
        io.write_record(f, np.arange(3, dtype="<i4"), np.array([1.5], "<f8"))
    raw = open(ps, "rb").read()
    assert len(raw) == 4 + 20 + 4
    assert np.frombuffer(raw[:4], "<i4")[0] == 20 and raw[:4] == raw[-4:]
    with open(ps, "rb") as f:
        pl = io.read_record(f)
    assert np.array_equal(np.frombuffer(pl[:12], "<i4"), [0, 1, 2])


def test_formats_golden():
    g = np.load(GOLDEN)
    lines = [str(s) for s in g["fmt_lines"]]
    assert len(lines) == len(FMT_CASES)
    for (kind, v, iv), want in zip(FMT_CASES, lines):
        assert python_line(kind, v, iv) == want, (kind, v, iv)


def test_formats_against_reference_build():
    r = ref_or_skip()
    rng = np.random.default_rng(9)
    for kind, v, iv in FMT_CASES:
        assert python_line(kind, v, iv) == r.format(kind, v, iv), (kind, v, iv)
    for _ in range(200):
        v = tuple(rng.standard_normal(4) * 10.0 ** rng.integers(-6, 4, size=4))
        t = abs(v[0])
        assert io.fmt_ene(t, v[1], v[2]) == r.format(1, (t, v[1], v[2]))
        assert io.fmt_norms(t, abs(v[1])) == r.format(2, (t, abs(v[1])))
        assert io.fmt_spectra_line(v[2], v[3]) == r.format(4, (v[2], v[3]))
        assert io.fmt_synopsis(7, 1234, t, abs(v[3])) == r.format(5, (t, abs(v[3])), (7, 1234))
        assert io.fmt_complexity(t, 17, 123456) == r.format(6, (t,), (17, 123456))
        m = tuple(abs(x) for x in rng.standard_normal(7) * 10.0 ** rng.integers(-3, 2, size=7))
        assert io.fmt_monitor(*m) == r.format(7, m), m


def test_qq_init_read_by_reference(tmp_path):
    r = ref_or_skip()
    qq = np.random.default_rng(2).standard_normal((NG, NG))
    io.write_qq_init(str(tmp_path / "qq_init.r8"), 1.25, qq)
    t, back = r.read_qq_init(str(tmp_path))
    assert t == 1.25 and bits_equal(back, qq)


def save_oracle_state(tmp_path, nsave=2):
    """Two contour saves of the seeded oracle state (the second after a
    shift of the nodes), as savecont writes them."""
    o, qq, c = oracle_state(NG)
    rng = np.random.default_rng(3)
    saves = []
    for k in range(nsave):
        ck = dict(c)
        ck["x"] = c["x"] + 1e-3 * k
        resid = rng.standard_normal((NG, NG)) * 0.1
        t = 0.5 * (k + 1) + 1.0 / 3.0
        io.savecont(str(tmp_path), k + 1, t, ck, resid)
        saves.append((ck, resid, t))
    return saves


def test_readcont_matches_reference(tmp_path):
    """readcont of each save: the reference's own readcont (common.f90:30-100)
    and ours give the same contours, residual, time and qs, bit for bit."""
    from oracle.hyo import Oracle
    r = ref_or_skip()
    saves = save_oracle_state(tmp_path)
    for iind, (ck, resid, t) in enumerate(saves, start=1):
        c, qr, tt = io.readcont(str(tmp_path), iind, NG)
        qs_r, qr_r, t_r = r.readcont(str(tmp_path), iind)
        rc = r.contours()
        for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
            assert bits_equal(c[k], rc[k]), (iind, k)
        assert bits_equal(c["x"], ck["x"]) and bits_equal(c["nextq"], ck["nextq"])
        assert bits_equal(qr, qr_r) and bits_equal(qr, resid.astype(np.float32).astype(np.float64))
        assert tt == t_r == round(t, 5)
        # qs = ptospc(con2grid + qr), the oracle's way (con2grid with the
        # qjump parsed from the synopsis: equal qs means equal qjump too)
        o = Oracle(NG)
        o.set_contours(c)
        qs = o.ptospc(o.con2grid() + qr)
        assert bits_equal(qs, qs_r), iind


def test_readcont_rebuilds_links(tmp_path):
    """Without the reference build: the round trip of savecont / readcont
    keeps every array, and nextq closes each contour on itself."""
    saves = save_oracle_state(tmp_path, 1)
    ck, resid, t = saves[0]
    c, qr, tt = io.readcont(str(tmp_path), 1, NG)
    for k in ("x", "y", "nextq", "i1q", "i2q", "npq", "indq"):
        assert bits_equal(c[k], ck[k]), k
    assert tt == round(t, 5) and c["qjump"] == float(f"{ck['qjump']:.12f}")
