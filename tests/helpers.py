"""Shared fixtures for the parity tests: seeded fields and oracle contours."""
import functools

import numpy as np

from hydra_amd.ic import random_pv


@functools.lru_cache(maxsize=None)
def oracle_state(ng, seed=12345):
    """Oracle model initialised from the seeded field and recontoured
    (caps.f90 initialise + congen recontour); returns (oracle, qq, contours)."""
    from oracle.hyo import Oracle
    qq = random_pv(ng, seed=seed)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    return o, qq, o.contours()


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                                 b.view(np.int64) if b.dtype == np.float64 else b)


# (seed, circles per level, smallest and largest radius) of the surgery
# tangles: crossing circles, a dense tangle, and circles small enough that
# renode leaves three nodes or fewer and surgery deletes them
TANGLES = [(1, 6, 0.02, 1.0), (2, 40, 0.02, 0.6), (3, 30, 0.001, 0.006)]


def tangle_contours(dm, rng, ncirc, levels, rmin, rmax):
    """Closed polygons on circles of random centre and radius, nodes about
    0.7 dm apart, grouped by level in ascending order: a dense tangle of
    crossing segments of the same level."""
    hl = np.pi
    xs, ys, nxt, i1, i2, npq, ind = [], [], [], [], [], [], []
    n = 0
    for lev in levels:
        for _ in range(ncirc):
            cx, cy = rng.uniform(-hl, hl, 2)
            r = rng.uniform(rmin, rmax)
            m = max(4, int(2 * np.pi * r / (0.7 * dm)))
            th = np.sort(rng.uniform(0, 2 * np.pi, m))
            x = cx + r * np.cos(th)
            y = cy + r * np.sin(th)
            x = x - 2 * hl * np.floor((x + hl) / (2 * hl))
            y = y - 2 * hl * np.floor((y + hl) / (2 * hl))
            xs.append(x)
            ys.append(y)
            nxt.append(np.r_[np.arange(n + 2, n + m + 1), n + 1])
            i1.append(n + 1)
            i2.append(n + m)
            npq.append(m)
            ind.append(lev)
            n += m
    cat = np.concatenate
    return dict(x=cat(xs), y=cat(ys), nextq=cat(nxt).astype(np.int32),
                i1q=np.array(i1, np.int32), i2q=np.array(i2, np.int32),
                npq=np.array(npq, np.int32), indq=np.array(ind, np.int32), qjump=1.0)


def state_digest(c, qs, t):
    """SHA-256 of a flow's state: contour store (xq, yq as float64 bits,
    nextq and the contour table as int32) and the spectral PV qs, plus t's
    bits and the counts.  Equal digests mean bit-identical states."""
    import hashlib
    hc = hashlib.sha256()
    for k in ("x", "y"):
        hc.update(np.ascontiguousarray(c[k], dtype="<f8").tobytes())
    for k in ("nextq", "i1q", "i2q", "npq", "indq"):
        hc.update(np.ascontiguousarray(c[k], dtype="<i4").tobytes())
    hq = hashlib.sha256(np.ascontiguousarray(qs, dtype="<f8").tobytes())
    return {"t": float(t).hex(), "nptq": int(len(c["x"])), "nq": int(len(c["i1q"])),
            "contours": hc.hexdigest()[:32], "qs": hq.hexdigest()[:32]}
