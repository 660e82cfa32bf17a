"""ctypes bindings of the oracle (liboracle.so) and of the amdflang build of
the reference modules (_ref/libhydra_ref_ngNN.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker.

Array conventions (as in the reference): a physical field (ny,nx) is a numpy
array of shape (nx, ny) in C order (index [ix, iy]); a spectral field (nx,ny)
is shape (ny, nx) (index [ky, kx]).  Contour index arrays are 1-based.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class HyoParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("ncontq", C.c_int),
        ("ellx", C.c_double), ("elly", C.c_double), ("kd", C.c_double), ("beta", C.c_double),
        ("rtherm", C.c_double), ("rekman", C.c_double), ("cdamp", C.c_double), ("nnu", C.c_double),
        ("tgsave", C.c_double), ("tcsave", C.c_double), ("dt_fixed", C.c_double),
        ("npm", C.c_int64), ("ell_cells", C.c_double),
    ]


def build(ref_ngs=()):
    """Compile liboracle.so (and, when the reference is present, _ref builds)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/basic_hydra"):
        for ng in ref_ngs:
            subprocess.run(["make", "-s", "-C", HERE, "ref", f"NG={ng}"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.hyo_create.restype = C.c_void_p
        L.hyo_create.argtypes = [C.POINTER(HyoParams)]
        L.hyo_destroy.argtypes = [C.c_void_p]
        L.hyo_dfield.restype = C.POINTER(C.c_double)
        L.hyo_dfield.argtypes = [C.c_void_p, C.c_char_p]
        L.hyo_ifield.restype = C.POINTER(C.c_int)
        L.hyo_ifield.argtypes = [C.c_void_p, C.c_char_p]
        L.hyo_get_d.restype = C.c_double
        L.hyo_get_d.argtypes = [C.c_void_p, C.c_char_p]
        L.hyo_set_d.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.hyo_get_i.restype = C.c_long
        L.hyo_get_i.argtypes = [C.c_void_p, C.c_char_p]
        L.hyo_set_i.argtypes = [C.c_void_p, C.c_char_p, C.c_long]
        for fn in ("hyo_forfft", "hyo_revfft"):
            getattr(L, fn).argtypes = [C.c_int, C.c_int, _D, _D, _I]
        L.hyo_initfft.argtypes = [C.c_int, _I, _D]
        for fn in ("hyo_xderiv", "hyo_yderiv"):
            getattr(L, fn).argtypes = [C.c_int, C.c_int, _D, _D, _D]
        L.hyo_ptospc.argtypes = [C.c_void_p, _D, _D]
        L.hyo_spctop.argtypes = [C.c_void_p, _D, _D]
        L.hyo_main_invert.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.hyo_gradient.argtypes = [C.c_void_p, _D, _D, _D]
        L.hyo_velint.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.hyo_con2grid.argtypes = [C.c_void_p, _D]
        L.hyo_con2grid_counts.argtypes = [C.c_void_p, _D, _I]
        L.hyo_surgery.argtypes = [C.c_void_p]
        L.hyo_renode.argtypes = [C.c_void_p, _D, _D, C.c_int, _D, _D, C.POINTER(C.c_int)]
        L.hyo_recontour.argtypes = [C.c_void_p, _D]
        L.hyo_init_from_field.argtypes = [C.c_void_p, _D, C.c_double]
        for fn in ("hyo_evolution_init", "hyo_advect_start", "hyo_prepare", "hyo_inversion",
                   "hyo_adapt"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.hyo_source.argtypes = [C.c_void_p, _D, _D, C.c_int]
        L.hyo_advance.argtypes = [C.c_void_p]
        L.hyo_advect_step.argtypes = [C.c_void_p]
        L.hyo_run.argtypes = [C.c_void_p, C.c_int]
        L.hyo_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hyo_spec1d.argtypes = [C.c_void_p, _D, _D, C.c_int]
        for fn in ("hyo_l1norm", "hyo_l2norm"):
            getattr(L, fn).argtypes = [C.c_void_p, _D]
            getattr(L, fn).restype = C.c_double
        _lib = L
    return _lib


def default_params(ng, kd=0.0, ncontq=80, dt_fixed=0.0, npm=0, tsave=1000.0, cdamp=2.0,
                   ell_cells=0.0, tgsave=None, tcsave=None):
    return HyoParams(nx=ng, ny=ng, ncontq=ncontq, ellx=2 * np.pi, elly=2 * np.pi, kd=kd,
                     beta=0.0, rtherm=0.0, rekman=0.0, cdamp=cdamp, nnu=3.0,
                     tgsave=tsave if tgsave is None else tgsave,
                     tcsave=tsave if tcsave is None else tcsave, dt_fixed=dt_fixed, npm=npm,
                     ell_cells=ell_cells)


class Oracle:
    """One oracle model (constants, tables, contour store and fields)."""

    def __init__(self, ng, **kw):
        self.L = lib()
        self.params = default_params(ng, **kw)
        self.m = self.L.hyo_create(C.byref(self.params))
        if not self.m:
            raise ValueError("hyo_create failed")
        self.ng = ng

    def __del__(self):
        if getattr(self, "m", None):
            self.L.hyo_destroy(self.m)
            self.m = None

    # -- views -------------------------------------------------------------
    def d(self, name, n=None, shape=None):
        p = self.L.hyo_dfield(self.m, name.encode())
        if not p:
            raise KeyError(name)
        if shape is None:
            shape = (n,) if n is not None else (self.ng, self.ng)
        return np.ctypeslib.as_array(p, shape=shape)

    def i(self, name, n):
        p = self.L.hyo_ifield(self.m, name.encode())
        return np.ctypeslib.as_array(p, shape=(n,))

    def get(self, name):
        v = self.L.hyo_get_i(self.m, name.encode())
        if v != -1 or name in ("nq", "nptq"):
            return int(v)
        return self.L.hyo_get_d(self.m, name.encode())

    def getd(self, name):
        return self.L.hyo_get_d(self.m, name.encode())

    def set(self, name, v):
        if isinstance(v, (int, np.integer)) and name in ("nq", "nptq", "contex", "ireg", "phase"):
            self.L.hyo_set_i(self.m, name.encode(), int(v))
        else:
            self.L.hyo_set_d(self.m, name.encode(), float(v))

    # -- contour store -----------------------------------------------------
    def contours(self):
        n, nc = self.get("nptq"), self.get("nq")
        return dict(x=self.d("xq", n).copy(), y=self.d("yq", n).copy(),
                    nextq=self.i("nextq", n).copy(), i1q=self.i("i1q", nc).copy(),
                    i2q=self.i("i2q", nc).copy(), npq=self.i("npq", nc).copy(),
                    indq=self.i("indq", nc).copy(), qjump=self.getd("qjump"))

    def set_contours(self, c):
        n, nc = len(c["x"]), len(c["i1q"])
        self.d("xq", n)[:] = c["x"]
        self.d("yq", n)[:] = c["y"]
        self.i("nextq", n)[:] = c["nextq"]
        for k in ("i1q", "i2q", "npq", "indq"):
            self.i(k, nc)[:] = c[k]
        self.set("nptq", n)
        self.set("nq", nc)
        self.set("qjump", c["qjump"])
        self.set("phase", 0)  # a restart recontours (caps.f90 recont) first

    # -- routines ----------------------------------------------------------
    def ptospc(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        s = np.empty_like(r)
        self.L.hyo_ptospc(self.m, r, s)
        return s

    def spctop(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64).copy()
        r = np.empty_like(s)
        self.L.hyo_spctop(self.m, s, r)
        return r

    def main_invert(self, qq):
        qq = np.ascontiguousarray(qq, dtype=np.float64)
        uu, vv, pp = (np.empty_like(qq) for _ in range(3))
        self.L.hyo_main_invert(self.m, qq, uu, vv, pp)
        return uu, vv, pp

    def gradient(self, ff):
        ff = np.ascontiguousarray(ff, dtype=np.float64)
        fx, fy = np.empty_like(ff), np.empty_like(ff)
        self.L.hyo_gradient(self.m, ff, fx, fy)
        return fx, fy

    def velint(self, uu, vv):
        n = self.get("nptq")
        uq, vq = np.empty(n), np.empty(n)
        self.L.hyo_velint(self.m, np.ascontiguousarray(uu), np.ascontiguousarray(vv), uq, vq)
        return uq, vq

    def con2grid(self, with_counts=False):
        qc = np.empty((self.ng, self.ng))
        if with_counts:
            nf = 4 * self.ng
            cnt = np.empty((nf, nf), dtype=np.int32)
            self.L.hyo_con2grid_counts(self.m, qc, cnt)
            return qc, cnt
        self.L.hyo_con2grid(self.m, qc)
        return qc

    def surgery(self):
        rc = self.L.hyo_surgery(self.m)
        if rc:
            raise RuntimeError(f"hyo_surgery failed ({rc})")

    def renode(self, xd, yd, cap=None):
        xd = np.ascontiguousarray(xd, dtype=np.float64)
        yd = np.ascontiguousarray(yd, dtype=np.float64)
        cap = cap or (4 * len(xd) + 4096)
        xr, yr = np.empty(cap), np.empty(cap)
        npr = C.c_int(0)
        rc = self.L.hyo_renode(self.m, xd, yd, len(xd), xr, yr, C.byref(npr))
        if rc:
            raise RuntimeError("hyo_renode failed")
        return xr[: npr.value].copy(), yr[: npr.value].copy()

    def recontour(self, qq):
        qq = np.ascontiguousarray(qq, dtype=np.float64).copy()
        rc = self.L.hyo_recontour(self.m, qq)
        if rc:
            raise RuntimeError(f"hyo_recontour failed ({rc})")

    def init_from_field(self, qq, t0=0.0):
        self.L.hyo_init_from_field(self.m, np.ascontiguousarray(qq, dtype=np.float64), t0)

    def evolution_init(self):
        self.L.hyo_evolution_init(self.m)

    def advect_start(self):
        """evolution.f90:44-51: init, twist = 0, ireg = 0 (run resumes
        inside the advect loop afterwards)."""
        self.L.hyo_advect_start(self.m)

    def prepare(self):
        """evolution.f90:135-162: qs folded, qd = qs - qc, qr = spctop(qd)."""
        self.L.hyo_prepare(self.m)

    def inversion(self):
        """evolution.f90:494-520."""
        self.L.hyo_inversion(self.m)

    def advance(self):
        rc = self.L.hyo_advance(self.m)
        if rc:
            raise RuntimeError("hyo_advance failed")

    def advect_step(self):
        return self.L.hyo_advect_step(self.m)

    def run(self, nsteps):
        if self.L.hyo_run(self.m, nsteps):
            raise RuntimeError("hyo_run failed")

    def energy(self):
        a, b = C.c_double(), C.c_double()
        self.L.hyo_energy(self.m, C.byref(a), C.byref(b))
        return a.value, b.value

    def spec1d(self, ss, iopt=0):
        """(spec[0..max(nx,ny)], ss after the call); entries above kmax stay 0."""
        ss = np.ascontiguousarray(ss, dtype=np.float64).copy()
        spec = np.zeros(max(ss.shape) + 1)
        self.L.hyo_spec1d(self.m, ss, spec, int(iopt))
        return spec, ss

    def norms(self, qq):
        qq = np.ascontiguousarray(qq, dtype=np.float64)
        return self.L.hyo_l1norm(self.m, qq), self.L.hyo_l2norm(self.m, qq)


# ---------------------------------------------------------------------------
# The reference itself (amdflang build of its modules + ref_shim.f90).

def ref_path(ng):
    return os.path.join(HERE, "_ref", f"libhydra_ref_ng{ng}.so")


def ref_available(ng):
    return os.path.exists(ref_path(ng))


def run_big_stack(fn, *a, **kw):
    """Run fn in a thread with a 1 GiB stack (reference routines keep large
    automatic arrays on the stack)."""
    out = {}

    def tgt():
        try:
            out["r"] = fn(*a, **kw)
        except BaseException as e:  # pragma: no cover - re-raised below
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(1 << 30)
    try:
        t = threading.Thread(target=tgt)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


class Ref:
    """The reference's modules at grid ng (compile-time size)."""

    def __init__(self, ng):
        self.ng = ng
        self.L = C.CDLL(ref_path(ng))
        L = self.L
        L.ref_set_contours.argtypes = [C.c_int, C.c_int, _D, _D, _I, _I, _I, _I, _I,
                                       C.POINTER(C.c_double)]
        L.ref_get_contours.argtypes = [_D, _D, _I, _I, _I, _I, _I]
        L.ref_counts.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ref_velint.argtypes = [_D, _D, _D, _D]
        L.ref_con2grid.argtypes = [_D]
        L.ref_renode.argtypes = [_D, _D, C.c_int, _D, _D, C.POINTER(C.c_int)]
        L.ref_ptospc.argtypes = [_D, _D]
        L.ref_spctop.argtypes = [_D, _D]
        L.ref_forfft.argtypes = [C.c_int, _D]
        L.ref_revfft.argtypes = [C.c_int, _D]
        L.ref_xderiv.argtypes = [_D, _D]
        L.ref_yderiv.argtypes = [_D, _D]
        L.ref_main_invert.argtypes = [_D, _D, _D, _D]
        L.ref_gradient.argtypes = [_D, _D, _D]
        L.ref_table.argtypes = [C.c_int, _D]
        L.ref_consts.argtypes = [_D]
        L.ref_recontour.argtypes = [_D]
        run_big_stack(L.ref_init)

    def table(self, tid):
        n = self.ng
        size = {**{k: n * n for k in range(1, 7)}, 7: n, 8: n, 9: n, 10: n, 11: 2 * n, 12: 2 * n}[tid]
        out = np.zeros(size)
        self.L.ref_table(tid, out)
        return out

    def consts(self):
        out = np.zeros(12)
        self.L.ref_consts(out)
        return out

    def set_contours(self, c):
        q = C.c_double(c["qjump"])
        self.L.ref_set_contours(len(c["x"]), len(c["i1q"]), np.ascontiguousarray(c["x"]),
                                np.ascontiguousarray(c["y"]),
                                np.ascontiguousarray(c["nextq"], dtype=np.int32),
                                *(np.ascontiguousarray(c[k], dtype=np.int32)
                                  for k in ("i1q", "i2q", "npq", "indq")), C.byref(q))

    def contours(self):
        n, nc = C.c_int(), C.c_int()
        self.L.ref_counts(C.byref(n), C.byref(nc))
        n, nc = n.value, nc.value
        x, y = np.empty(n), np.empty(n)
        nxt = np.empty(n, dtype=np.int32)
        arrs = [np.empty(nc, dtype=np.int32) for _ in range(4)]
        self.L.ref_get_contours(x, y, nxt, *arrs)
        return dict(x=x, y=y, nextq=nxt, i1q=arrs[0], i2q=arrs[1], npq=arrs[2], indq=arrs[3])

    def velint(self, uu, vv):
        n, nc = C.c_int(), C.c_int()
        self.L.ref_counts(C.byref(n), C.byref(nc))
        uq, vq = np.empty(n.value), np.empty(n.value)
        run_big_stack(self.L.ref_velint, np.ascontiguousarray(uu), np.ascontiguousarray(vv), uq, vq)
        return uq, vq

    def con2grid(self):
        qc = np.empty((self.ng, self.ng))
        run_big_stack(self.L.ref_con2grid, qc)
        return qc

    def surgery(self):
        run_big_stack(self.L.ref_surgery)

    def renode(self, xd, yd):
        xr = np.empty(4 * len(xd) + 4096)
        yr = np.empty_like(xr)
        npr = C.c_int(0)
        run_big_stack(self.L.ref_renode, np.ascontiguousarray(xd), np.ascontiguousarray(yd),
                      len(xd), xr, yr, C.byref(npr))
        return xr[: npr.value].copy(), yr[: npr.value].copy()

    def ptospc(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        s = np.empty_like(r)
        run_big_stack(self.L.ref_ptospc, r, s)
        return s

    def spctop(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64).copy()
        r = np.empty_like(s)
        run_big_stack(self.L.ref_spctop, s, r)
        return r

    def forfft(self, m, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        run_big_stack(self.L.ref_forfft, m, x)
        return x

    def revfft(self, m, x):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        run_big_stack(self.L.ref_revfft, m, x)
        return x

    def main_invert(self, qq):
        qq = np.ascontiguousarray(qq, dtype=np.float64).copy()
        uu, vv, pp = (np.empty_like(qq) for _ in range(3))
        run_big_stack(self.L.ref_main_invert, qq, uu, vv, pp)
        return uu, vv, pp

    def gradient(self, ff):
        ff = np.ascontiguousarray(ff, dtype=np.float64).copy()
        fx, fy = np.empty_like(ff), np.empty_like(ff)
        run_big_stack(self.L.ref_gradient, ff, fx, fy)
        return fx, fy

    def spec1d(self, ss, iopt=0):
        """As Oracle.spec1d; entries above kmax are left 0."""
        self.L.ref_spec1d.argtypes = [_D, _D, C.c_int]
        ss = np.ascontiguousarray(ss, dtype=np.float64).copy()
        spec = np.zeros(max(ss.shape) + 1)
        run_big_stack(self.L.ref_spec1d, ss, spec, int(iopt))
        return spec, ss

    def norms(self, qq):
        self.L.ref_norms.argtypes = [_D, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        qq = np.ascontiguousarray(qq, dtype=np.float64).copy()
        a, b = C.c_double(), C.c_double()
        run_big_stack(self.L.ref_norms, qq, C.byref(a), C.byref(b))
        return a.value, b.value

    def contint(self, qq, n):
        """generic.f90:66-79: (max(qq) - min(qq)) / n."""
        self.L.ref_contint.argtypes = [_D, C.c_int, C.POINTER(C.c_double)]
        dq = C.c_double()
        run_big_stack(self.L.ref_contint, np.ascontiguousarray(qq, dtype=np.float64), int(n),
                      C.byref(dq))
        return dq.value

    def kmax(self):
        k = C.c_int()
        self.L.ref_kmax(C.byref(k))
        return k.value

    def readcont(self, directory, iind):
        """common.f90 readcont(iind) run in directory: (qs, qr, t); the
        contours are then the module's (contours())."""
        self.L.ref_readcont.argtypes = [C.c_int, _D, _D, C.POINTER(C.c_double)]
        n = self.ng
        qs, qr, t = np.empty((n, n)), np.empty((n, n)), C.c_double()
        cwd = os.getcwd()
        os.chdir(directory)
        try:
            run_big_stack(self.L.ref_readcont, int(iind), qs, qr, C.byref(t))
        finally:
            os.chdir(cwd)
        return qs, qr, t.value

    def read_qq_init(self, directory):
        """caps.f90:150-153's open/read of directory/qq_init.r8: (t, qq)."""
        self.L.ref_read_qq_init.argtypes = [C.POINTER(C.c_double), _D]
        n = self.ng
        qq, t = np.empty((n, n)), C.c_double()
        cwd = os.getcwd()
        os.chdir(directory)
        try:
            run_big_stack(self.L.ref_read_qq_init, C.byref(t), qq)
        finally:
            os.chdir(cwd)
        return t.value, qq

    def format(self, kind, v=(), iv=()):
        """The reference's output formats (see ref_shim.f90 ref_format)."""
        self.L.ref_format.argtypes = [C.c_int, _D, _I, C.c_char_p]
        va = np.zeros(8)
        va[:len(v)] = v
        ia = np.zeros(2, np.int32)
        ia[:len(iv)] = iv
        buf = C.create_string_buffer(128)
        self.L.ref_format(int(kind), va, ia, buf)
        return buf.raw.decode().rstrip()

    def recontour(self, qq):
        run_big_stack(self.L.ref_recontour, np.ascontiguousarray(qq, dtype=np.float64).copy())
