"""The drop-in Fortran modules (hydra_amd/fortran/dropin: contours,
spectral, sta2dfft, stafft) under the reference's own, unmodified callers.

oracle/_ref/dropin_ng32 is congen.f90 and common.f90 of the reference
compiled against the drop-in modules and linked with libhydra_amd.so
(oracle/Makefile target dropin; built where the reference sources are).  It
recontours a seeded field (the reference's congen code calling the drop-in
renode and reading the drop-in tables), recontours again with those contours
and a residual (the con2ugrid branch), and restarts from a contour save
through the reference's readcont (drop-in con2grid and ptospc).  Every
result must equal the oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_ng32")
NG = 32
KEYS = ("x", "y", "nextq", "i1q", "i2q", "npq", "indq")


def read_dump(path):
    with open(path, "rb") as f:
        b = f.read()
    nq, nptq = np.frombuffer(b, "<i4", 2)
    qjump = np.frombuffer(b, "<f8", 1, 8)[0]
    o = 16
    x = np.frombuffer(b, "<f8", nptq, o)
    y = np.frombuffer(b, "<f8", nptq, o + 8 * nptq)
    o += 16 * nptq
    nextq = np.frombuffer(b, "<i4", nptq, o)
    o += 4 * nptq
    i1q, i2q, npq, indq = (np.frombuffer(b, "<i4", nq, o + 4 * nq * k) for k in range(4))
    return dict(x=x, y=y, nextq=nextq, i1q=i1q, i2q=i2q, npq=npq, indq=indq, qjump=qjump)


def same(a, b):
    for k in KEYS:
        assert bits_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_reference_callers_over_dropin_modules(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/dropin_ng32 not built (needs the reference sources)")
    from hydra_amd import io
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    qq = random_pv(NG, seed=31)
    res = 0.4 * random_pv(NG, seed=32) * (qq.max() - qq.min()) / 80.0
    io.write_qq_init(str(tmp_path / "qq_init.r8"), 0.0, qq)
    io.write_qq_init(str(tmp_path / "qq_res.r8"), 0.0, res)
    # oracle: caps.f90 set-up and recont, then recontouring with a residual
    o = Oracle(NG)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    first = o.contours()
    o.recontour(res)
    second = o.contours()
    # a contour save of the first contours with the residual, for readcont
    io.savecont(str(tmp_path), 1, 0.5, dict(first, qjump=o.getd("qjump")), res)
    out = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0 and "dropin_check done" in out.stdout
    d1 = read_dump(tmp_path / "dropin_first.bin")
    assert d1["qjump"] == o.getd("qjump")
    same(d1, first)
    same(read_dump(tmp_path / "dropin_second.bin"), second)
    # readcont: contours and qjump as saved, qs = ptospc(con2grid + qr)
    c, qr, t = io.readcont(str(tmp_path), 1, NG)
    d3 = read_dump(tmp_path / "dropin_readcont.bin")
    same(d3, c)
    assert d3["qjump"] == c["qjump"]
    r = Oracle(NG)
    r.set_contours(c)
    qa = r.con2grid() + qr
    want = r.ptospc(qa)
    qs = np.fromfile(tmp_path / "dropin_qs.bin", dtype="<f8").reshape(NG, NG)
    assert bits_equal(qs, want)
