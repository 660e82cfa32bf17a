"""A whole run as the reference's caps program makes it (caps.f90:71-96 with
initialise 100-246 for a new run and evolution.f90:28-92): the
device-resident advect loop on one flow, with every output file the
reference writes -- complexity.asc, monitor.asc, ene.asc, norms.asc,
spectra.asc, qq.r4 and the contour saves in cont/ (qqsynopsis.asc,
qqresi.r4, qqindexNNN, qqnodesNNN).

The library advances the flow and computes what each file holds (the save
schedule of adapt, the interpolation of savegrid, the residual of
savecont, the monitor and complexity rows); this module writes the lines
and records in the reference's formats (hydra_amd.io)."""
from __future__ import annotations

import os

from hydra_amd import io


def _drain(h, outdir):
    with open(os.path.join(outdir, "complexity.asc"), "a") as f:
        for t, nq, nptq in h.take_records("complexity"):
            f.write(io.fmt_complexity(t, int(nq), int(nptq)) + "\n")
    with open(os.path.join(outdir, "monitor.asc"), "a") as f:
        for row in h.take_records("monitor"):
            f.write(io.fmt_monitor(*row) + "\n")
    g, c = h.pending_saves()
    for _ in range(g):
        io.savegrid_slot(h, outdir, h.take_grid_save())
    for _ in range(c):
        io.savecont_slot(outdir, h.take_contour_save())


def caps_run(h, qq_init, outdir, tsim, t0=0.0):
    """A new simulation (loopinit = 0) of the flow h from the full PV field
    qq_init (ny,nx) at time t0 to tsim, saving every tgsave / tcsave of h's
    parameters into outdir.  Returns the number of RK4 steps taken."""
    os.makedirs(os.path.join(outdir, "cont"), exist_ok=True)
    for name in ("complexity.asc", "ene.asc", "norms.asc", "monitor.asc", "spectra.asc",
                 "qq.r4", os.path.join("cont", "qqsynopsis.asc"),
                 os.path.join("cont", "qqresi.r4")):
        open(os.path.join(outdir, name), "wb").close()   # status='replace'
    h.set_field(qq_init, t0)
    h.set_saves(True, tinit=t0, igrec=0, icrec=1)
    steps = 0
    while h.scalar("t") <= tsim:   # caps.f90:83 and evolution.f90:55
        h.run(1)
        steps += 1
        _drain(h, outdir)
    # evolution.f90:85-89: a save time that coincides with tsim
    dt = h.scalar("dt")
    p = h.params
    gsave = abs(t0 + p.tgsave * _records(h, "grid") - tsim) < 1e-6 * dt
    csave = abs(t0 + p.tcsave * _records(h, "cont") - tsim) < 1e-6 * dt
    if gsave or csave:
        h.save_now(grid=gsave, cont=csave)
        _drain(h, outdir)
    return steps


def _records(h, kind):
    """The next grid (igrec) or contour (icrec) record number."""
    return h.save_counters()[0 if kind == "grid" else 1]
