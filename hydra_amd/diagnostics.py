"""Error of a flow against the plain FP64 flow advanced from the same state.

The plain FP64 step is bit-identical to the reference's arithmetic, so its
state stands for the reference's; a mixed-precision (or otherwise modified)
flow is measured against it with the reference's own diagnostics, evaluated
in FP64 on both states:

  * RMS and maximum node displacement (periodic minimum image, domain units),
    defined when the two contour topologies are equal;
  * maximum PV-grid difference: max |q - q64| over the physical grid of
    q = spctop(qq) after an inversion (evolution.f90:494-520), relative to
    max |q64|;
  * energy (evolution.f90:612-644) and enstrophy (l2norm of q,
    generic.f90:48-64) of each state, their relative difference at the same
    time, and each flow's drift from the common initial values.

The other flow's state (contours, qs, qd) is evaluated by loading it into
the plain context, so call this last: the plain context's own state is
replaced.
"""
from __future__ import annotations

import numpy as np


def fp64_diagnostics(h64):
    """(q grid, energy, enstrophy) of the plain context's current state."""
    h64.inversion()
    q = h64.spctop(h64.field("qq"))
    ekin, epot = h64.energy()
    return q, ekin + epot, h64.norms(q)[1]


def error_vs_fp64(h64, hx, e0=None, z0=None):
    """Errors of context hx against the plain FP64 context h64 (same ng,
    both advanced over the same steps from the same state).  e0, z0: the
    common initial energy and enstrophy, for the drifts."""
    c64, cx = h64.contours(), hx.contours()
    same = len(c64["x"]) == len(cx["x"]) and all(
        np.array_equal(c64[k], cx[k]) for k in ("nextq", "i1q", "i2q", "npq", "indq"))
    out = {"topology_equal": bool(same), "nodes": int(len(c64["x"]))}
    if same:
        L = h64.params.ellx
        d = np.r_[cx["x"] - c64["x"], cx["y"] - c64["y"]]
        d = d - L * np.round(d / L)
        out["rms_node_displacement"] = float(np.sqrt(np.mean(d * d))) if len(d) else 0.0
        out["max_node_displacement"] = float(np.abs(d).max()) if len(d) else 0.0
    q64, e64, z64 = fp64_diagnostics(h64)
    # load hx's state into the plain context and evaluate it the same way
    qjump = hx.scalar("qjump")
    h64.set_contours(dict(cx, qjump=qjump))
    for name in ("qs", "qd"):
        h64.put_field(name, hx.field(name))
    qx, ex, zx = fp64_diagnostics(h64)
    out["max_pv_grid_diff"] = float(np.abs(qx - q64).max() / np.abs(q64).max())
    # the scale of that difference: a node moving across a fine grid line
    # changes one fine cell's level count by one, worth at most one contour
    # interval qjump on the coarse grid (the averaging weights sum to one)
    out["qjump_over_max_pv"] = float(qjump / np.abs(q64).max())
    out["energy_rel_diff"] = float(abs(ex - e64) / abs(e64))
    out["enstrophy_rel_diff"] = float(abs(zx - z64) / abs(z64))
    if e0 is not None:
        out["energy_drift_fp64"] = float(e64 / e0 - 1.0)
        out["energy_drift"] = float(ex / e0 - 1.0)
    if z0 is not None:
        out["enstrophy_drift_fp64"] = float(z64 / z0 - 1.0)
        out["enstrophy_drift"] = float(zx / z0 - 1.0)
    return out
