"""The whole-run driver (hydra_amd.run.caps_run) and the library's save
machinery against the oracle, driven here by the reference's text
(caps.f90:83-91, evolution.f90:28-131, 166-204, 524-608, 648-778): the
recontouring/advect cycle with surgery, the monitor and complexity rows,
gsave/csave from adapt, the interpolated grid save at the next step's
start (and the pending saves init drops after a recontouring), the contour
saves with their residual.  Every file the run writes must equal the one
written from the oracle's numbers."""
import filecmp
import os

import numpy as np
import pytest

from hydra_amd import io

pytestmark = pytest.mark.gpu

NG = 32
TWISTMAX, NREGMAX = 2.5, 20


def oracle_files(qq, outdir, tsim, tgsave, tcsave):
    """evolution.f90's advect loop and saves written out over the oracle."""
    from oracle.hyo import Oracle
    o = Oracle(NG, tgsave=tgsave, tcsave=tcsave)
    o.init_from_field(qq)
    os.makedirs(os.path.join(outdir, "cont"), exist_ok=True)
    cpx = open(os.path.join(outdir, "complexity.asc"), "w")
    mon = open(os.path.join(outdir, "monitor.asc"), "w")
    glx = o.getd("glx")
    tinit, igrec, icrec = 0.0, 0, 1
    qspre = ekpre = eppre = tgrid = dtprev = None

    class HView:   # the transforms and tables savegrid needs, from the oracle
        def table(self, name):
            if name == "shells":
                return np.array([o.get("kmax"), o.get("kmaxred"), o.getd("dsumi")])
            return o.d(name, NG + 1) if name in ("spmf", "alk") else o.d(name, NG)

        def spec1d(self, ss, iopt):
            return o.spec1d(ss.copy(), iopt)

        def spctop(self, s):
            return o.spctop(s.copy())

        def norms(self, q):
            return o.norms(q)

    hv = HView()
    while o.getd("t") <= tsim:
        o.recontour(o.d("qr").copy())                    # recont
        o.evolution_init()                               # init
        cpx.write(io.fmt_complexity(o.getd("t"), o.get("nq"), o.get("nptq")) + "\n")
        gsave = csave = False                            # evolution.f90:115-116
        qspre = o.d("qs").copy()                         # evolution.f90:126
        o.set("twist", 0.0)
        ireg = 0
        while o.getd("t") <= tsim:
            twist = o.getd("twist")
            if twist > TWISTMAX:
                ireg += 1
                if ireg == NREGMAX:
                    o.prepare()
                    break
                o.surgery()
                cpx.write(io.fmt_complexity(o.getd("t"), o.get("nq"), o.get("nptq")) + "\n")
                o.set("twist", twist - TWISTMAX)
            # advance: inversion, reset, the saves, adapt (the oracle's
            # advance repeats the same inversion and reset itself)
            t = o.getd("t")
            o.inversion()
            qq = o.d("qq").copy()
            if gsave:
                pt = (t - tgrid) / dtprev
                ptc = 1.0 - pt
                ek, ep = o.energy()
                io.savegrid_slot(hv, outdir, dict(tgrid=tgrid, ekin=pt * ekpre + ptc * ek,
                                                  epot=pt * eppre + ptc * ep, igrec=igrec,
                                                  qqs=pt * qspre + ptc * qq))
                igrec += 1
            if csave:
                res = o.spctop(qq - o.d("qc"))
                c = o.contours()
                io.savecont(outdir, icrec, t, dict(c, qjump=o.getd("qjump")), res)
                icrec += 1
            flo, fhi = o.d("flo"), o.d("fhi")
            qs_reset = flo * (o.d("qs") - o.d("qc")) + o.d("qc") + o.d("qd")
            ek_now, ep_now = o.energy()
            o.advance()
            dt = o.getd("dt")
            zr = o.getd("zzrms")
            mon.write(io.fmt_monitor(t, o.getd("uumax") * dt / glx, 0.5 * (zr * zr), zr,
                                     o.getd("zzmax"), o.getd("uumax"), o.getd("twist")) + "\n")
            tg = tinit + tgsave * igrec
            gsave = t + dt >= tg
            if gsave:
                tgrid, qspre, ekpre, eppre = tg, qs_reset, ek_now, ep_now
            csave = t + dt >= tinit + tcsave * icrec
            dtprev = dt
    # evolution.f90:85-89: a save time that coincides with tsim
    t, dt = o.getd("t"), o.getd("dt")
    gs = abs(tinit + tgsave * igrec - tsim) < 1e-6 * dt
    cs = abs(tinit + tcsave * icrec - tsim) < 1e-6 * dt
    if gs or cs:
        o.inversion()
        qq = o.d("qq").copy()
    if gs:
        tgrid = tinit + tgsave * igrec
        pt = (t - tgrid) / dt
        ptc = 1.0 - pt
        ek, ep = o.energy()
        io.savegrid_slot(hv, outdir, dict(tgrid=tgrid, ekin=pt * ekpre + ptc * ek,
                                          epot=pt * eppre + ptc * ep, igrec=igrec,
                                          qqs=pt * qspre + ptc * qq))
    if cs:
        res = o.spctop(qq - o.d("qc"))
        io.savecont(outdir, icrec, t, dict(o.contours(), qjump=o.getd("qjump")), res)
    cpx.close()
    mon.close()
    return o


def test_caps_run_files_equal_oracle(tmp_path):
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from hydra_amd.run import caps_run
    qq = random_pv(NG, seed=7)
    # tsim falls on a grid save time: the end-of-run save is made too
    tsim, tgsave, tcsave = 0.6, 0.05, 0.12
    a, b = tmp_path / "gpu", tmp_path / "oracle"
    h = Hydra(NG, tgsave=tgsave, tcsave=tcsave)
    steps = caps_run(h, qq, str(a), tsim)
    h.close()
    oracle_files(qq, str(b), tsim, tgsave, tcsave)
    names = ["complexity.asc", "monitor.asc", "ene.asc", "norms.asc", "spectra.asc", "qq.r4",
             "cont/qqsynopsis.asc", "cont/qqresi.r4"]
    names += sorted("cont/" + f for f in os.listdir(b / "cont") if f.startswith("qq") and
                    f[-3:].isdigit())
    assert steps > 40
    assert len(open(b / "ene.asc").readlines()) >= 8        # grid saves happened
    assert len(open(b / "cont/qqsynopsis.asc").readlines()) >= 4
    assert len(open(b / "complexity.asc").readlines()) >= 4  # surgeries happened
    for n in names:
        if not filecmp.cmp(a / n, b / n, shallow=False) and n.endswith(".asc"):
            import difflib
            print("".join(list(difflib.unified_diff(open(b / n).readlines(),
                                                    open(a / n).readlines(), "oracle",
                                                    "gpu"))[:40]))
        assert filecmp.cmp(a / n, b / n, shallow=False), n
