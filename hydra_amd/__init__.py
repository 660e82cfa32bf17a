"""hydra_amd -- MI355X-native contour-advection step (Hydra ca/plane/qg/caps).

Host-side mirror of the reference's module interface over the C ABI of
libhydra_amd.so (include/hydra_amd.h).  Method names and argument meaning
follow the Fortran routines they replace:

    ptospc / spctop / xderiv / yderiv      lib/stafft/sta2dfft.f90:103-245
    main_invert / gradient                 ca/plane/qg/caps/spectral.f90:304-366
    velint / con2grid                      ca/plane/qg/caps/contours.f90:136-346
    set_field                              ca/plane/qg/caps/caps.f90:143-183
    evolution_init / inversion / advance   ca/plane/qg/caps/evolution.f90

Errors raise HydraError (the reference stops the program instead).  There is
no CPU fallback: if the HIP library is missing or no GPU is present, the
constructor raises.

Array layouts are the reference's: a physical field (ny,nx) is a float64
numpy array of shape (nx, ny) in C order; a spectral field (nx,ny) has
shape (ny, nx).  Contour index arrays are 1-based int32.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhydra_amd.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

FIELDS = {"qs": 0, "qd": 1, "qc": 2, "qq": 3, "pp": 4, "uu": 5, "vv": 6, "qr": 7, "epq": 8,
          "emq": 9}
SCALARS = {"t": 0, "dt": 1, "twist": 2, "zzrms": 3, "zzmax": 4, "uumax": 5, "qjump": 6,
           "dtmax": 7}
TABLES = {"green": 1, "flo": 2, "fhi": 3, "qdiss": 4, "rksq": 5, "hrkx": 6, "rkx": 7, "hrky": 8,
          "rky": 9, "xtrig": 10, "ytrig": 11, "spmf": 12, "alk": 13, "shells": 14, "bety": 15,
          "rk": 16, "kmag": 17, "xgf": 18, "ygf": 19, "xgu": 20, "ygu": 21, "wcongen": 22,
          "ixp": 23, "iyp": 24, "ixw": 25, "iyw": 26, "ibx": 27, "iby": 28}


class HydraError(RuntimeError):
    pass


class HydParams(C.Structure):
    _fields_ = [
        ("ng", C.c_int32), ("ncontq", C.c_int32),
        ("ellx", C.c_double), ("elly", C.c_double), ("kd", C.c_double), ("beta", C.c_double),
        ("rtherm", C.c_double), ("rekman", C.c_double), ("cdamp", C.c_double), ("nnu", C.c_double),
        ("tgsave", C.c_double), ("tcsave", C.c_double), ("dt_fixed", C.c_double),
        ("node_capacity", C.c_int64), ("precision", C.c_int32), ("device", C.c_int32),
        ("ell_cells", C.c_double),
    ]


_lib = None


def lib():
    """Load libhydra_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HydraError(f"{LIB_PATH} is missing: run `python -m hydra_amd.build` first "
                             "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        L.hyd_last_error.restype = C.c_char_p
        L.hyd_init.argtypes = [C.POINTER(HydParams), C.POINTER(C.c_void_p)]
        L.hyd_destroy.argtypes = [C.c_void_p]
        L.hyd_initfft.argtypes = [C.c_int32, _I, _D]
        for fn in ("hyd_forfft", "hyd_revfft"):
            getattr(L, fn).argtypes = [C.c_void_p, C.c_int32, C.c_int32, _D, _D, _I]
        for fn in ("hyd_ptospc", "hyd_spctop", "hyd_xderiv", "hyd_yderiv"):
            getattr(L, fn).argtypes = [C.c_void_p, _D, _D]
        L.hyd_main_invert.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.hyd_gradient.argtypes = [C.c_void_p, _D, _D, _D]
        L.hyd_spec1d.argtypes = [C.c_void_p, _D, _D, C.c_int32]
        L.hyd_norms.argtypes = [C.c_void_p, _D, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hyd_energy.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hyd_get_table.argtypes = [C.c_void_p, C.c_int32, _D]
        L.hyd_velint.argtypes = [C.c_void_p, _D, _D, _D, _D, C.c_int32, _D, _D]
        L.hyd_con2grid.argtypes = [C.c_void_p, _D, _D, _I, C.c_int32, C.c_double, _D, C.c_void_p]
        L.hyd_sum_sequential.argtypes = [C.c_void_p, _D, C.c_int64, C.POINTER(C.c_double)]
        L.hyd_set_field.argtypes = [C.c_void_p, _D, C.c_double]
        L.hyd_set_contours.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _D, _D, _I, _I, _I, _I,
                                       _I, C.c_double]
        L.hyd_check_contours.argtypes = [C.POINTER(HydParams), C.c_int32, C.c_int32, _D, _D, _I,
                                         _I, _I, _I, _I]
        L.hyd_get_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.hyd_get_contours.argtypes = [C.c_void_p, _D, _D, _I, _I, _I, _I, _I]
        L.hyd_get_field.argtypes = [C.c_void_p, C.c_int32, _D]
        L.hyd_put_field.argtypes = [C.c_void_p, C.c_int32, _D]
        L.hyd_get_scalar.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        L.hyd_set_scalar.argtypes = [C.c_void_p, C.c_int32, C.c_double]
        for fn in ("hyd_evolution_init", "hyd_inversion", "hyd_advance", "hyd_sync",
                   "hyd_prepare", "hyd_recontour", "hyd_surgery", "hyd_advect_step"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.hyd_renode.argtypes = [C.c_void_p, _D, _D, C.c_int32, _D, _D, C.c_int32,
                                 C.POINTER(C.c_int32)]
        L.hyd_run.argtypes = [C.c_void_p, C.c_int32]
        L.hyd_phase_times.argtypes = [C.c_void_p, _D, C.c_int32]
        L.hyd_advance_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), C.c_int32]
        L.hyd_take_records.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _D]
        L.hyd_set_saves.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_int32]
        L.hyd_pending_saves.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.hyd_take_grid_save.argtypes = [C.c_void_p, _D, _D]
        L.hyd_contour_save_size.argtypes = [C.c_void_p, C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]
        L.hyd_take_contour_save.argtypes = [C.c_void_p, _D, _D, _D, _D, _I, _I, _I]
        L.hyd_save_now.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.hyd_save_counters.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.hyd_run_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_int32]
        L.hyd_get_tuning.argtypes = [C.c_void_p, _I]
        L.hyd_copy_calibration.argtypes = [C.c_void_p, C.c_int64]
        L.hyd_event_overhead.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        L.hyd_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.hyd_kernel_times.argtypes = [C.c_void_p, C.c_int32, C.c_char_p,
                                       np.ctypeslib.ndpointer(dtype=np.float64),
                                       np.ctypeslib.ndpointer(dtype=np.int64)]
        L.hyd_cnn_create.argtypes = [C.c_int32, C.c_double, C.c_uint64, C.c_int32,
                                     C.POINTER(C.c_void_p)]
        L.hyd_cnn_destroy.argtypes = [C.c_void_p]
        _F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.hyd_cnn_forward.argtypes = [C.c_void_p, _F, _D, _D, C.c_int32, _F, _F]
        L.hyd_cnn_gradients.argtypes = [C.c_void_p, _F, _D, _D, C.c_int32, _F, _F, _F,
                                        C.POINTER(C.c_float)]
        L.hyd_cnn_train_step.argtypes = [C.c_void_p, _F, _D, _D, C.c_int32, _F, _F, C.c_float,
                                         C.POINTER(C.c_float)]
        L.hyd_cnn_num_weights.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.hyd_cnn_get_weights.argtypes = [C.c_void_p, _F]
        L.hyd_cnn_set_weights.argtypes = [C.c_void_p, _F]
        _U16 = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
        L.hyd_cnn_mfma_probe.argtypes = [_U16, _U16, _F]
        L.hyd_set_correction.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                         C.c_double]
        L.hyd_contour_velocity.argtypes = [C.c_void_p, C.c_int32, _D, _D, _I, C.c_double,
                                           C.c_int32, _D, _D, _D, _D, C.c_void_p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HydraError(f"hydra_amd error {rc}: {lib().hyd_last_error().decode()}")


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _params(ng, ncontq=80, ellx=2 * np.pi, elly=None, kd=0.0, beta=0.0, rekman=0.0, cdamp=2.0,
            nnu=3.0, tgsave=1000.0, tcsave=1000.0, dt_fixed=0.0, node_capacity=0, device=0,
            precision=0, ell_cells=0.0):
    return HydParams(ng=ng, ncontq=ncontq, ellx=ellx, elly=ellx if elly is None else elly, kd=kd,
                     beta=beta, rtherm=0.0, rekman=rekman, cdamp=cdamp, nnu=nnu, tgsave=tgsave,
                     tcsave=tcsave, dt_fixed=dt_fixed, node_capacity=node_capacity,
                     precision=precision, device=device, ell_cells=ell_cells)


def _contour_args(c):
    ii = {k: np.ascontiguousarray(c[k], dtype=np.int32) for k in ("nextq", "i1q", "i2q", "npq", "indq")}
    return (len(c["x"]), len(c["i1q"]), _f64(c["x"]), _f64(c["y"]), ii["nextq"], ii["i1q"],
            ii["i2q"], ii["npq"], ii["indq"])


def initfft(n):
    """stafft.f90:63-125 initfft: (factors[5], trig[2n]) on the host."""
    f = np.zeros(5, np.int32)
    t = np.zeros(2 * n)
    _check(lib().hyd_initfft(int(n), f, t))
    return f, t


def check_contours(c, ng, **kw):
    """The contour-table checks of set_contours, on the host without a
    device (contours numbered consecutively from node 1, ascending levels,
    one nextq cycle per contour, nodes inside the domain); raises
    HydraError naming the first that fails."""
    p = _params(ng, **kw)
    _check(lib().hyd_check_contours(C.byref(p), *_contour_args(c)))


class Hydra:
    """One flow on one GPU: the reference's module state, device resident."""

    def __init__(self, ng, ncontq=80, ellx=2 * np.pi, elly=None, kd=0.0, beta=0.0, rekman=0.0,
                 cdamp=2.0, nnu=3.0, tgsave=1000.0, tcsave=1000.0, dt_fixed=0.0,
                 node_capacity=0, device=0, precision=0, ell_cells=0.0):
        self.L = lib()
        self.ng = ng
        self.params = _params(ng, ncontq=ncontq, ellx=ellx, elly=elly, kd=kd, beta=beta,
                              rekman=rekman, cdamp=cdamp, nnu=nnu, tgsave=tgsave, tcsave=tcsave,
                              dt_fixed=dt_fixed, node_capacity=node_capacity, device=device,
                              precision=precision, ell_cells=ell_cells)
        h = C.c_void_p()
        _check(self.L.hyd_init(C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.hyd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- sta2dfft / spectral ---------------------------------------------------
    def _grid_op(self, fn, a, nout=1):
        a = _f64(a, (self.ng, self.ng))
        outs = [np.empty_like(a) for _ in range(nout)]
        _check(fn(self.h, a, *outs))
        return outs[0] if nout == 1 else tuple(outs)

    def forfft(self, m, n, x, trig=None, factors=None):
        """stafft.f90:187-287: m vectors of length n, vector index fastest
        (x has m*n values), transformed forward; returns the new array."""
        if trig is None:
            factors, trig = initfft(n)
        x = np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        _check(self.L.hyd_forfft(self.h, int(m), int(n), x, _f64(trig),
                                 np.ascontiguousarray(factors, dtype=np.int32)))
        return x

    def revfft(self, m, n, x, trig=None, factors=None):
        """stafft.f90:290-406, the inverse of forfft."""
        if trig is None:
            factors, trig = initfft(n)
        x = np.ascontiguousarray(x, dtype=np.float64).ravel().copy()
        _check(self.L.hyd_revfft(self.h, int(m), int(n), x, _f64(trig),
                                 np.ascontiguousarray(factors, dtype=np.int32)))
        return x

    def ptospc(self, rvar):
        return self._grid_op(self.L.hyd_ptospc, rvar)

    def spctop(self, svar):
        return self._grid_op(self.L.hyd_spctop, svar)

    def xderiv(self, var):
        return self._grid_op(self.L.hyd_xderiv, var)

    def yderiv(self, var):
        return self._grid_op(self.L.hyd_yderiv, var)

    def main_invert(self, qq):
        """returns (uu, vv, pp)"""
        return self._grid_op(self.L.hyd_main_invert, qq, 3)

    def gradient(self, ff):
        return self._grid_op(self.L.hyd_gradient, ff, 2)

    # -- diagnostics ---------------------------------------------------------
    def spec1d(self, ss, iopt=0):
        """spectral.f90:370-424: (spec[0..max(nx,ny)], ss after the call);
        entries above kmax are 0 here."""
        ss = _f64(ss, (self.ng, self.ng)).copy()
        spec = np.zeros(self.ng + 1)
        _check(self.L.hyd_spec1d(self.h, ss, spec, int(iopt)))
        return spec, ss

    def norms(self, qq):
        """generic.f90:30-64: (l1norm, l2norm) of a physical field"""
        a, b = C.c_double(), C.c_double()
        _check(self.L.hyd_norms(self.h, _f64(qq, (self.ng, self.ng)), C.byref(a), C.byref(b)))
        return a.value, b.value

    def energy(self):
        """evolution.f90:612-644: (ekin, epot) from the current uu, vv, pp"""
        a, b = C.c_double(), C.c_double()
        _check(self.L.hyd_energy(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def table(self, name):
        n = self.ng
        u = 16 * n
        size = {1: n * n, 2: n * n, 3: n * n, 4: n * n, 5: n * n, 10: 2 * n, 11: 2 * n,
                12: n + 1, 13: n + 1, 14: 3, 16: n * n, 17: n * n, 18: 4 * n, 19: 4 * n,
                20: u, 21: u, 22: 1024, 25: 3 * u, 26: 3 * u, 27: u + 1,
                28: u + 1}.get(TABLES[name], n)
        out = np.empty(size)
        _check(self.L.hyd_get_table(self.h, TABLES[name], out))
        return out

    # -- contours ----------------------------------------------------------------
    def velint(self, uu, vv, xq, yq):
        xq, yq = _f64(xq), _f64(yq)
        n = len(xq)
        uq, vq = np.empty(n), np.empty(n)
        _check(self.L.hyd_velint(self.h, _f64(uu, (self.ng, self.ng)), _f64(vv, (self.ng, self.ng)),
                                 xq, yq, n, uq, vq))
        return uq, vq

    def con2grid(self, xq, yq, nextq, qjump, with_counts=False):
        xq, yq = _f64(xq), _f64(yq)
        nextq = np.ascontiguousarray(nextq, dtype=np.int32)
        qc = np.empty((self.ng, self.ng))
        cnt = np.empty((4 * self.ng, 4 * self.ng), dtype=np.int32) if with_counts else None
        _check(self.L.hyd_con2grid(self.h, xq, yq, nextq, len(xq), float(qjump), qc,
                                   cnt.ctypes.data_as(C.c_void_p) if with_counts else None))
        return (qc, cnt) if with_counts else qc

    def contour_velocity(self, xq, yq, nextq, qjump, xs, ys, fields=False):
        """Velocity of the contours' PV on this context's grid (con2grid,
        ptospc, main_invert) at the points (xs, ys); with fields=True also
        the FP32 (u, v, q) grid [ng*ng, 3].  The context's state is not
        touched."""
        xq, yq, xs, ys = _f64(xq), _f64(yq), _f64(xs), _f64(ys)
        nextq = np.ascontiguousarray(nextq, dtype=np.int32)
        m = len(xs)
        us, vs = np.empty(m), np.empty(m)
        f = np.empty((self.ng * self.ng, 3), np.float32) if fields else None
        _check(self.L.hyd_contour_velocity(self.h, len(xq), xq, yq, nextq, float(qjump), m, xs,
                                           ys, us, vs,
                                           f.ctypes.data_as(C.c_void_p) if fields else None))
        return (us, vs, f) if fields else (us, vs)

    def set_correction(self, net, su=1.0, sq=1.0, sd=1.0):
        """Attach a trained VelocityCNN (None detaches): each RK4 stage adds
        its (du, dv) at the nodes to velint's velocity; inputs are scaled by
        1/su (velocity) and 1/sq (PV), the output by sd."""
        _check(self.L.hyd_set_correction(self.h, net.h if net is not None else None,
                                         float(su), float(sq), float(sd)))
        self._net = net

    def renode(self, xd, yd, cap=None):
        """contours.f90:687-933: returns (xr, yr); empty when the contour is dropped."""
        xd, yd = _f64(xd), _f64(yd)
        cap = cap or (4 * len(xd) + 4096)
        xr, yr = np.empty(cap), np.empty(cap)
        npr = C.c_int32(0)
        _check(self.L.hyd_renode(self.h, xd, yd, len(xd), xr, yr, cap, C.byref(npr)))
        return xr[: npr.value].copy(), yr[: npr.value].copy()

    def recontour(self):
        """congen.f90:24-128 on the device state (residual qr)."""
        _check(self.L.hyd_recontour(self.h))

    def prepare(self):
        """evolution.f90:135-162."""
        _check(self.L.hyd_prepare(self.h))

    def sum_sequential(self, a):
        a = _f64(a).ravel()
        out = C.c_double()
        _check(self.L.hyd_sum_sequential(self.h, a, len(a), C.byref(out)))
        return out.value

    # -- device-resident state ----------------------------------------------------
    def set_field(self, qq, t0=0.0):
        _check(self.L.hyd_set_field(self.h, _f64(qq, (self.ng, self.ng)), float(t0)))

    def set_contours(self, c):
        _check(self.L.hyd_set_contours(self.h, *_contour_args(c), float(c["qjump"])))

    def counts(self):
        a, b = C.c_int32(), C.c_int32()
        _check(self.L.hyd_get_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def contours(self):
        n, nc = self.counts()
        x, y = np.empty(max(n, 1)), np.empty(max(n, 1))
        nxt = np.empty(max(n, 1), dtype=np.int32)
        arrs = [np.empty(max(nc, 1), dtype=np.int32) for _ in range(4)]
        _check(self.L.hyd_get_contours(self.h, x, y, nxt, *arrs))
        return dict(x=x[:n], y=y[:n], nextq=nxt[:n], i1q=arrs[0][:nc], i2q=arrs[1][:nc],
                    npq=arrs[2][:nc], indq=arrs[3][:nc], qjump=self.scalar("qjump"))

    def field(self, name):
        out = np.empty((self.ng, self.ng))
        _check(self.L.hyd_get_field(self.h, FIELDS[name], out))
        return out

    def put_field(self, name, a):
        _check(self.L.hyd_put_field(self.h, FIELDS[name], _f64(a, (self.ng, self.ng))))

    def scalar(self, name):
        out = C.c_double()
        _check(self.L.hyd_get_scalar(self.h, SCALARS[name], C.byref(out)))
        return out.value

    def set_scalar(self, name, v):
        _check(self.L.hyd_set_scalar(self.h, SCALARS[name], float(v)))

    def evolution_init(self):
        _check(self.L.hyd_evolution_init(self.h))

    def take_records(self, kind):
        """Rows since the last call: "monitor" (t, cfl, zzrms^2/2, zzrms,
        zzmax, uumax, twist) or "complexity" (t, nq, nptq)."""
        k, w = {"monitor": (0, 7), "complexity": (1, 3)}[kind]
        rows = []
        while True:
            buf = np.zeros(256 * w)
            n = self.L.hyd_take_records(self.h, k, 256, buf)
            if n < 0:
                _check(n)
            rows.append(buf[:n * w].reshape(n, w))
            if n < 256:
                return np.concatenate(rows)

    def set_saves(self, on=True, tinit=0.0, igrec=0, icrec=1):
        """Enable the reference's save schedule (evolution.f90:587-599)."""
        _check(self.L.hyd_set_saves(self.h, int(bool(on)), float(tinit), int(igrec), int(icrec)))

    def save_counters(self):
        """(igrec, icrec): the next grid and contour record numbers."""
        g, c = C.c_int32(), C.c_int32()
        _check(self.L.hyd_save_counters(self.h, C.byref(g), C.byref(c)))
        return g.value, c.value

    def pending_saves(self):
        g, c = C.c_int32(), C.c_int32()
        _check(self.L.hyd_pending_saves(self.h, C.byref(g), C.byref(c)))
        return g.value, c.value

    def take_grid_save(self):
        """dict(tgrid, ekin, epot, igrec, qqs) of the oldest grid save."""
        info = np.zeros(4)
        qqs = np.empty((self.ng, self.ng))
        _check(self.L.hyd_take_grid_save(self.h, info, qqs))
        return dict(tgrid=info[0], ekin=info[1], epot=info[2], igrec=int(info[3]), qqs=qqs)

    def take_contour_save(self):
        """dict(t, qjump, icrec, contex, residual, x, y, npq, i1q, indq)."""
        nq, npt = C.c_int32(), C.c_int32()
        _check(self.L.hyd_contour_save_size(self.h, C.byref(nq), C.byref(npt)))
        nq, npt = nq.value, npt.value
        info = np.zeros(4)
        res = np.empty((self.ng, self.ng))
        x, y = np.empty(max(npt, 1)), np.empty(max(npt, 1))
        a = [np.empty(max(nq, 1), np.int32) for _ in range(3)]
        _check(self.L.hyd_take_contour_save(self.h, info, res, x, y, *a))
        return dict(t=info[0], qjump=info[1], icrec=int(info[2]), contex=bool(info[3]),
                    residual=res, x=x[:npt], y=y[:npt], npq=a[0][:nq], i1q=a[1][:nq],
                    indq=a[2][:nq])

    def save_now(self, grid=False, cont=False):
        _check(self.L.hyd_save_now(self.h, int(bool(grid)), int(bool(cont))))

    def phase_times(self, reset=False):
        """Surgery and recontouring calls and wall seconds so far:
        {"surgery": (calls, s), "recontour": (calls, s)}."""
        out = np.zeros(4)
        _check(self.L.hyd_phase_times(self.h, out, int(bool(reset))))
        return {"surgery": (int(out[0]), float(out[1])), "recontour": (int(out[2]), float(out[3]))}

    def advance_counts(self, reset=False):
        """(RK4 advances, node-updates) taken so far; reset zeroes both."""
        a, u = C.c_int64(), C.c_int64()
        _check(self.L.hyd_advance_counts(self.h, C.byref(a), C.byref(u), int(bool(reset))))
        return a.value, u.value

    def inversion(self):
        _check(self.L.hyd_inversion(self.h))

    def advance(self):
        """evolution.f90:166-326: one RK4 step."""
        _check(self.L.hyd_advance(self.h))

    def surgery(self):
        """contours.f90:350-683 on the device contour store."""
        _check(self.L.hyd_surgery(self.h))

    def advect_step(self):
        """evolution.f90:55-80 loop body; returns 1 when the caller must
        recontour (prepare has run), 0 after an advance."""
        rc = self.L.hyd_advect_step(self.h)
        if rc < 0:
            _check(rc)
        return rc

    def run(self, nsteps):
        """caps.f90:83-91 for nsteps advances (recontour, init, advect)."""
        _check(self.L.hyd_run(self.h, int(nsteps)))

    @staticmethod
    def run_batch(flows, nsteps):
        """hyd_run for each flow of an ensemble (distinct Hydra objects),
        their steps running concurrently; same results as flow.run(nsteps)."""
        arr = (C.c_void_p * len(flows))(*[f.h for f in flows])
        _check(lib().hyd_run_batch(arr, len(flows), int(nsteps)))

    def tuning(self):
        """FFT batch per layout class picked by the autotuner at init."""
        v = np.zeros(4, dtype=np.int32)
        _check(self.L.hyd_get_tuning(self.h, v))
        return v.tolist()

    def copy_calibration(self, nbytes):
        """One streaming copy kernel of nbytes (traffic-counter reference)."""
        _check(self.L.hyd_copy_calibration(self.h, int(nbytes)))

    def event_overhead(self, reps=200):
        """ms the per-launch timing events report for an empty kernel"""
        v = C.c_double()
        _check(self.L.hyd_event_overhead(self.h, int(reps), C.byref(v)))
        return v.value

    def sync(self):
        _check(self.L.hyd_sync(self.h))

    # -- timing -----------------------------------------------------------------
    def set_timing(self, level=2):
        """0 off, 1 single-kernel records only, 2 all records."""
        _check(self.L.hyd_set_timing(self.h, int(level)))

    def kernel_times(self, max_records=1 << 16):
        names = C.create_string_buffer(32 * max_records)
        ms = np.zeros(max_records)
        by = np.zeros(max_records, dtype=np.int64)
        n = self.L.hyd_kernel_times(self.h, max_records, names, ms, by)
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), ms[i], int(by[i]))
                for i in range(n)]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class VelocityCNN:
    """The learned sub-grid velocity correction (hyd_cnn_*): a 4-layer 3x3
    CNN on the n x n coarse grid (inputs u, v, q), output (du, dv) sampled
    at node positions; BF16 matrix-core kernels, FP32 accumulation, Adam."""

    def __init__(self, n, ell=2 * np.pi, seed=1, device=0):
        self.L = lib()
        self.n = n
        h = C.c_void_p()
        _check(self.L.hyd_cnn_create(int(n), float(ell), int(seed), int(device), C.byref(h)))
        self.h = h
        nw = C.c_int64()
        _check(self.L.hyd_cnn_num_weights(self.h, C.byref(nw)))
        self.nw = nw.value

    def close(self):
        if getattr(self, "h", None):
            self.L.hyd_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _in(self, in3):
        a = _f32(in3)
        if a.shape != (self.n * self.n, 3):
            a = a.reshape(self.n * self.n, 3)
        return a

    def forward(self, in3, xn, yn):
        m = len(xn)
        du, dv = np.empty(m, np.float32), np.empty(m, np.float32)
        _check(self.L.hyd_cnn_forward(self.h, self._in(in3), _f64(xn), _f64(yn), m, du, dv))
        return du, dv

    def gradients(self, in3, xn, yn, tu, tv):
        g = np.empty(self.nw, np.float32)
        loss = C.c_float()
        _check(self.L.hyd_cnn_gradients(self.h, self._in(in3), _f64(xn), _f64(yn), len(xn),
                                        _f32(tu), _f32(tv), g, C.byref(loss)))
        return g, loss.value

    def train_step(self, in3, xn, yn, tu, tv, lr=1e-3):
        loss = C.c_float()
        _check(self.L.hyd_cnn_train_step(self.h, self._in(in3), _f64(xn), _f64(yn), len(xn),
                                         _f32(tu), _f32(tv), float(lr), C.byref(loss)))
        return loss.value

    def weights(self):
        w = np.empty(self.nw, np.float32)
        _check(self.L.hyd_cnn_get_weights(self.h, w))
        return w

    def set_weights(self, w):
        _check(self.L.hyd_cnn_set_weights(self.h, _f32(w)))


def mfma_probe(a_bf16, b_bf16):
    """D = A B of one 16x16x32 BF16 matrix-core block (A row-major
    [16][32], b_bf16[n][k] [16][32]); returns D (16x16 FP32)."""
    d = np.empty(256, np.float32)
    _check(lib().hyd_cnn_mfma_probe(np.ascontiguousarray(a_bf16, np.uint16).ravel(),
                                    np.ascontiguousarray(b_bf16, np.uint16).ravel(), d))
    return d.reshape(16, 16)
