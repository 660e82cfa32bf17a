"""Write the golden fixtures of the reference itself.

Runs the amdflang build of the reference's own modules (oracle/_ref, built
from /root/reference by oracle/Makefile) on seeded inputs and stores inputs
and outputs in tests/golden/ref_ng32.npz.  Run in the build container only:

    make -C oracle ref NG=32 && python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hydra_amd.ic import random_pv  # noqa: E402
from oracle.hyo import Ref  # noqa: E402

NG = 32
NCONTQ = 20


def main():
    r = Ref(NG)
    out = {}
    qq = random_pv(NG, seed=2024)
    out["qq"] = qq
    qjump = (qq.max() - qq.min()) / NCONTQ
    out["qjump"] = np.float64(qjump)
    empty = dict(x=np.zeros(0), y=np.zeros(0), nextq=np.zeros(0, np.int32),
                 **{k: np.zeros(0, np.int32) for k in ("i1q", "i2q", "npq", "indq")}, qjump=qjump)
    r.set_contours(empty)
    r.recontour(qq)                      # congen.f90:24-128, first call (no contours)
    c = r.contours()
    for k, v in c.items():
        out["rc_" + k] = v
    out["c2g_qc"] = r.con2grid()         # contours.f90:173-346
    rng = np.random.default_rng(5)
    uu, vv = rng.standard_normal((2, NG, NG))
    out["vi_uu"], out["vi_vv"] = uu, vv
    out["vi_uq"], out["vi_vq"] = r.velint(uu, vv)   # contours.f90:136-169
    f = rng.standard_normal((NG, NG))
    s = rng.standard_normal((NG, NG))
    out["fft_f"], out["fft_s"] = f, s
    out["ptospc"] = r.ptospc(f)          # sta2dfft.f90:103-132
    out["spctop"] = r.spctop(s)          # sta2dfft.f90:136-165
    out["mi_uu"], out["mi_vv"], out["mi_pp"] = r.main_invert(s)   # spectral.f90:304-342
    out["gr_x"], out["gr_y"] = r.gradient(s)                      # spectral.f90:345-366
    for tid, name in [(1, "green"), (2, "flo"), (3, "fhi"), (4, "qdiss"), (7, "hrkx"),
                      (9, "hrky"), (11, "xtrig")]:
        out["tab_" + name] = r.table(tid)
    out["d4small"] = np.float64(r.consts()[2])
    # diagnostics: spec1d spectral.f90:370-424, l1norm/l2norm generic.f90:30-64
    out["spec_kmax"] = np.int64(r.kmax())
    for iopt in (0, 1):
        out[f"spec{iopt}"], out[f"spec{iopt}_ss"] = r.spec1d(s, iopt)
    out["norm_l1"], out["norm_l2"] = (np.float64(v) for v in r.norms(f))
    # the output formats of evolution.f90:676-728 and 750, written by the
    # reference's format statements (cases listed in tests/test_io.py)
    from tests.test_io import FMT_CASES
    out["fmt_lines"] = np.array([r.format(k, v, iv) for k, v, iv in FMT_CASES])
    # renode of the first three contours, contours.f90:687-933
    for j in range(3):
        i1, i2 = c["i1q"][j], c["i2q"][j]
        xr, yr = r.renode(c["x"][i1 - 1:i2], c["y"][i1 - 1:i2])
        out[f"rn{j}_x"], out[f"rn{j}_y"] = xr, yr
    # surgery on a sheared copy, contours.f90:350-683
    x = c["x"] + 0.03 * np.sin(3 * c["y"])
    y = c["y"] + 0.03 * np.cos(2 * c["x"])
    L = 2 * np.pi
    x = (1 - 1e-12) * (x - L * np.trunc(x / np.pi))
    y = (1 - 1e-12) * (y - L * np.trunc(y / np.pi))
    sc = dict(c, x=x, y=y, qjump=qjump)
    out["sg_in_x"], out["sg_in_y"] = x, y
    r.set_contours(sc)
    r.surgery()
    s2 = r.contours()
    for k, v in s2.items():
        out["sg_" + k] = v
    # contint, generic.f90:66-79 (the contour interval of the initial field)
    out["contint"] = np.float64(r.contint(qq, NCONTQ))
    # recontouring with contours present (congen.f90:40-61: con2ugrid of the
    # contours, congen.f90:338-441, plus the bilinear residual): the sheared
    # contours above and a seeded residual of about half a contour interval
    res = random_pv(NG, seed=77)
    res = 0.5 * qjump * res / np.abs(res).max()
    out["rx_res"] = res
    r.set_contours(sc)
    r.recontour(res)
    rx = r.contours()
    for k, v in rx.items():
        out["rx_" + k] = v
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"ref_ng{NG}.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", len(c["x"]), "nodes,", len(c["i1q"]), "contours;",
          "surgery ->", len(s2["x"]), len(s2["i1q"]), "; recontour with contours ->",
          len(rx["x"]), len(rx["i1q"]))


if __name__ == "__main__":
    main()
