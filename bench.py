"""Contour node-updates/s of the plain FP64 contour-advection step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--ng 1024] [--gpus 1]

Workload: the doubly periodic 2D Euler CAPS step at ng=1024 (contour grid
16ng for recontouring, 4ng inside con2grid), a seeded random PV field with
1D spectrum ~ k^3 exp(-2 (k/6)^2) and rms 4*pi, contoured into 80 levels
on the GPU (congen), surgery scale dm = 0.01*(2*pi/ng) (ell = glx; the
reference's own dm = 0.0625*glx is timed beside it), then K passes of the
advect loop (evolution.f90:55-80: surgery when the twist exceeds 2.5,
recontouring after 20 regularisations, one RK4 advance) on the
device-resident state.  One node-update = one contour node advanced through
one full RK4 step (four stages: con2grid, FFT inversion, sources, velint,
node and spectral updates).  Inputs are resident in HBM before timing.

Besides the JSON line the script prints nothing on stdout.  Roofline
entries count algorithmic bytes as the step's dataflow reads and writes
each logical field once per use (a 2D transform: its field once in and once
out, shared by its two 1D passes; node arrays at FP64):
  * roofline: the kernel with the largest share of the step (the inverse
    FFT pass), timed per launch with HIP events on the library's stream in
    steps run on one stream, so that each launch runs alone;
  * step_roofline: 137 s_g N + 336 P bytes per RK4 step (N = ng^2, s_g = 8)
    over the measured ms per step.
cpu_baseline is the oracle (C restatement of the reference, gcc -O3, one
core) on a bounded sample of the same state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HIP-event record name -> kernel symbol as rocprofv3 reports it ({N} = ng)
KERNEL_SYMBOL = {"k:fft_fwd": "fft_kernel<{N}, true, double, double, double>",
                 "k:fft_rev": "fft_kernel<{N}, false, double, double, double>",
                 "k:fft32_fwd": "fft_kernel<{N}, true, float, float, float|double>",
                 "k:fft32_rev": "fft_kernel<{N}, false, double|float, float, float>"}


def traffic_from_profile(kernel, alg_bytes):
    """HBM bytes per launch of `kernel`: its measured HBM bytes per
    algorithmic byte (committed rocprofv3 --pmc summary: FETCH_SIZE and
    WRITE_SIZE scaled by a copy kernel of known size) times this run's
    algorithmic bytes per launch; None without a summary for the kernel."""
    path = os.path.join(ROOT, "profiles", "fft_kernel_pmc_traffic_ng1024.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    rec = d.get(kernel)
    return None if rec is None else int(round(rec["hbm_per_algorithmic_byte"] * alg_bytes))


COPY_BYTES = 96 << 20  # each of read and write of the calibration copy (<= 128 ng^2 B)


def copy_calibration(h):
    """A plain device copy of known size (read COPY_BYTES, write COPY_BYTES):
    the reference point that scales the FETCH_SIZE/WRITE_SIZE counters."""
    h.copy_calibration(COPY_BYTES)


def pmc_summary(fetch_dir, write_dir, ng, out_json):
    """HBM traffic of the FFT kernels from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by the same
    counter on the calibration copy, as bytes per algorithmic byte
    (aggregated over the batch variants VT).  A dispatch's algorithmic bytes
    are (sizeof in + sizeof out) x n x vectors; the vectors are its blocks
    (grid size / workgroup size, both in the record) times the VT in the
    kernel's name."""
    import csv
    import glob
    import re
    size = {"double": 8, "float": 4}

    def rows(d, counter):
        out = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") == counter:
                        out.append(row)
        return out

    def summarise(d, counter):
        copy, per = [], {}
        for row in rows(d, counter):
            k = row["Kernel_Name"]
            v = float(row["Counter_Value"])
            if "k_copy_calib" in k:
                copy.append(v)
                continue
            m = re.search(r"fft_kernel<(\d+), (true|false), (\w+), (\w+), (\w+), (\d+)>", k)
            if not m:
                continue
            n = int(m.group(1))
            ti, to = size[m.group(3)], size[m.group(5)]
            # its blocks, VT vectors each (all fields of the launch)
            vectors = int(row["Grid_Size"]) // int(row["Workgroup_Size"]) * int(m.group(6))
            alg = (ti + to) * n * vectors
            key = "fft_kernel<%s, %s, %s, %s, %s>" % m.group(1, 2, 3, 4, 5)
            rec = per.setdefault(key, [0.0, 0.0, 0])
            rec[0] += v
            rec[1] += alg
            rec[2] += 1
        return float(np.mean(copy)), per

    cf, fe = summarise(fetch_dir, "FETCH_SIZE")
    cw, wr = summarise(write_dir, "WRITE_SIZE")
    kib = COPY_BYTES / 1024.0
    ff, fw = kib / cf, kib / cw
    run = ("python3 $REPO/bench.py --steps 3 --warmup 1 --no-cpu --no-mixed --ref-dm-steps 0 "
           "--copy-calib")
    out = {"commands": ["cd /tmp && export TMPDIR=/tmp",
                        "timeout -s KILL 300 rocprofv3 --pmc FETCH_SIZE --output-format csv "
                        "--kernel-include-regex 'fft_kernel|k_copy_calib' -d FETCH_DIR -- " + run,
                        "timeout -s KILL 300 rocprofv3 --pmc WRITE_SIZE --output-format csv "
                        "--kernel-include-regex 'fft_kernel|k_copy_calib' -d WRITE_DIR -- " + run,
                        "python3 $REPO/bench.py --pmc-summary FETCH_DIR WRITE_DIR"],
           "calibration": {"copy_bytes_each_way": COPY_BYTES, "fetch_kib": cf, "write_kib": cw,
                           "fetch_factor": ff, "write_factor": fw}}
    for k in sorted(fe):
        if k not in wr:
            continue
        hbm = (fe[k][0] * ff + wr[k][0] * fw) * 1024.0
        alg = fe[k][1]
        out[k] = {"dispatches": fe[k][2], "fetch_kib_total": fe[k][0], "write_kib_total": wr[k][0],
                  "algorithmic_bytes_total": alg, "hbm_per_algorithmic_byte": hbm / alg}
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def stats_summary(db, out, title, command):
    """Per-kernel table (calls, total and average duration in microseconds,
    share of the kernel time) from a rocprofv3 --kernel-trace database."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(end - start) from kernels "
                       "group by name order by 3 desc").fetchall()
    tot = sum(r[2] for r in rows)
    lines = [f"# rocprofv3 kernel statistics: {title}", "#",
             "# command (from the repository root on the MI355X box):"]
    lines += ["#   " + c for c in command]
    lines += ["# durations in microseconds", "#",
              f"{'kernel':<72} {'calls':>6} {'total_us':>12} {'avg_us':>10} {'pct':>6}"]
    for name, n, t in rows:
        nm = name if len(name) <= 72 else name[:69] + "..."
        lines.append(f"{nm:<72} {n:>6} {t / 1e3:>12.1f} {t / 1e3 / n:>10.2f} {100.0 * t / tot:>6.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def step_bytes(ng, nodes, s_g=8):
    """Algorithmic HBM bytes of one RK4 step: 137 field passes of s_g-byte
    grid elements (33 per stage x 4, adapt 5) and 42 FP64 node words."""
    return 137 * s_g * ng * ng + 336 * nodes


def make_flow(ng, qq, precision=0, ell_cells=0.0, ncontq=80):
    from hydra_amd import Hydra
    h = Hydra(ng, ncontq=ncontq, precision=precision, ell_cells=ell_cells,
              node_capacity=max(16 * ng * ng, 1 << 22))
    t0 = time.perf_counter()
    h.set_field(qq)
    h.recontour()
    h.evolution_init()
    h.sync()
    return h, time.perf_counter() - t0


def advect_steps(h, k):
    """k passes of the advect loop (evolution.f90:55-80): surgery when the
    twist exceeds 2.5, recontouring after 20 regularisations (caps.f90:
    83-91), then one RK4 advance; one library call, as the reference's main
    loop is one loop.  Returns the node-updates (the library counts each
    advance's nodes)."""
    h.advance_counts(reset=True)
    h.run(k)
    adv, updates = h.advance_counts(reset=True)
    if adv != k:
        raise RuntimeError(f"advect loop took {adv} advances, expected {k}")
    return updates


def timed_run(h, steps, warmup):
    for _ in range(warmup):
        advect_steps(h, 1)
    h.sync()
    h.phase_times(reset=True)
    t0 = time.perf_counter()
    updates = advect_steps(h, steps)
    h.sync()
    wall = time.perf_counter() - t0
    return updates, wall, h.phase_times()


def fft_roofline(h, steps, ng):
    """Per-launch timing of the FFT passes (the step's largest kernel by
    time) and the step's breakdown by kernel group."""
    kr = min(steps, 10)
    h.set_timing(1)
    advect_steps(h, kr)
    h.sync()
    recs = h.kernel_times()
    h.set_timing(0)
    tot, cnt, pbytes = {}, {}, {}
    for name, ms, by in recs:
        tot[name] = tot.get(name, 0.0) + ms
        cnt[name] = cnt.get(name, 0) + 1
        if name.startswith("k:"):
            pbytes.setdefault(name, []).append(by)
    nb = 2
    h.set_timing(2)
    advect_steps(h, nb)
    h.sync()
    gt, gc, gb = {}, {}, {}
    for name, ms, by in h.kernel_times():
        gt[name] = gt.get(name, 0.0) + ms
        gc[name] = gc.get(name, 0) + 1
        gb[name] = gb.get(name, 0) + by
    h.set_timing(0)
    groups = {}
    for name, ms in sorted(gt.items(), key=lambda kv: -kv[1]):
        log(f"  {name:16s} {ms / nb:8.3f} ms/step  {gc[name] // nb} calls/step")
        groups[name] = {"ms_per_step": round(ms / nb, 4), "calls_per_step": gc[name] // nb,
                        "algorithmic_gbs": round(gb[name] / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
    dom = max(pbytes, key=lambda k: tot[k])
    ev_ms = h.event_overhead(200)
    avg_ms = tot[dom] / cnt[dom]
    # a launch is one 1D pass of F fields; the 2D transform's fields count
    # once in and once out, shared by its two passes: half the pass's bytes
    pass_bytes = float(np.mean(pbytes[dom]))
    alg = pass_bytes / 2
    achieved = alg / (avg_ms * 1e-3) / 1e9
    ksym = KERNEL_SYMBOL.get(dom, dom).replace("{N}", str(ng))
    traffic = traffic_from_profile(ksym, pass_bytes)
    roof = {"bound": "hbm", "kernel": ksym, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": int(alg),
            "counting": "2D transform = its fields once in + once out, half per 1D-pass launch",
            "avg_launch_ms": round(avg_ms, 4), "empty_launch_ms": round(ev_ms, 4),
            "launches_per_step": cnt[dom] / kr}
    if traffic is not None:
        roof["traffic_over_algorithmic"] = round(traffic / alg, 3)
        roof["hbm_gbs_measured"] = round(traffic / (avg_ms * 1e-3) / 1e9, 1)
    return roof, groups


def regularisation_costs(h):
    """One surgery and one prepare + recontouring of the current state
    (the nq > 0 branch of congen.f90:40-61) timed on their own, then the
    loop restarted (advect's start): the amortised extras of a run."""
    h.phase_times(reset=True)
    p0 = h.counts()[0]
    h.surgery()
    h.prepare()
    h.recontour()
    h.evolution_init()
    h.sync()
    pt = h.phase_times(reset=True)
    return {"nodes": p0, "surgery_s": round(pt["surgery"][1], 4),
            "recontour_with_contours_s": round(pt["recontour"][1], 4),
            "nodes_after": h.counts()[0]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ng", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--ell-cells", type=float, default=1.0,
                    help="surgery/renode scale ell in grid lengths: 1.0 gives dm = 0.01 glx "
                         "(the workload's), 6.25 the reference's dm = 0.0625 glx")
    ap.add_argument("--ref-dm-steps", type=int, default=100,
                    help="steps of the same loop with the reference's dm (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision timing and its error against the FP64 step")
    ap.add_argument("--err-steps", type=int, default=100)
    ap.add_argument("--ensemble", type=int, default=0, metavar="FLOWS",
                    help="also time an ensemble of FLOWS independent mixed-precision flows "
                         "(ng from {256,512,1024}, rms in [2pi,6pi], 40-120 levels, seeded) "
                         "run together (hyd_run_batch) and one after another")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--precision", type=int, default=0,
                    help="0 plain FP64 (bit-exact), 1 mixed FP32 grid / FP64 state")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run the calibration copy (for rocprofv3 --pmc runs)")
    ap.add_argument("--stats-summary", nargs=4, metavar=("DB", "OUT", "TITLE", "COMMAND"),
                    help="write a per-kernel table from a rocprofv3 --kernel-trace database "
                         "(COMMAND lines separated by ' ;; ')")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs into profiles/fft_kernel_pmc_traffic_ng<ng>.json")
    args = ap.parse_args()
    if args.stats_summary:
        db, out, title, cmd = args.stats_summary
        stats_summary(db, out, title, cmd.split(" ;; "))
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.ng,
                    os.path.join(ROOT, "profiles", f"fft_kernel_pmc_traffic_ng{args.ng}.json"))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs one flow on one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from hydra_amd.ic import random_pv

    ng = args.ng
    qq = random_pv(ng, seed=args.seed)
    h, t_setup = make_flow(ng, qq, args.precision, args.ell_cells)
    P, nq = h.counts()
    tuning = h.tuning()
    dm = 0.04 * (args.ell_cells if args.ell_cells > 0 else 6.25) * (2 * np.pi / ng) / 4
    log(f"setup: ng={ng} ell={args.ell_cells} glx contours={nq} nodes={P} "
        f"recontour+init {t_setup:.2f}s")
    updates, wall, ph = timed_run(h, args.steps, args.warmup)
    P_end, nq_end = h.counts()
    ms_step = 1e3 * wall / args.steps
    value = updates / wall
    p_avg = updates / args.steps
    log(f"timed: {args.steps} steps {wall:.3f}s, nodes {P} -> {P_end}, contours {nq} -> {nq_end}, "
        f"surgeries {ph['surgery']}, recontours {ph['recontour']}")
    s_g = 8 if args.precision == 0 else 4
    bstep = step_bytes(ng, p_avg, s_g)
    step_roof = {"bound": "hbm", "bytes_per_step": int(bstep),
                 "formula": "137*s_g*N + 336*P (N=ng^2, s_g=%d, P=mean nodes)" % s_g,
                 "achieved": round(bstep / (ms_step * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": round(bstep / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "ms_per_step_at_peak": round(bstep / (HBM_PEAK_GBS * 1e9) * 1e3, 4)}
    roofline, groups = fft_roofline(h, args.steps, ng)
    run = {"setup_recontour_init_s": round(t_setup, 4), "timed_s": round(wall, 4),
           "surgeries": ph["surgery"][0], "surgery_s": round(ph["surgery"][1], 4),
           "recontours": ph["recontour"][0], "recontour_s": round(ph["recontour"][1], 4),
           "whole_run_s": round(t_setup + wall, 4),
           "whole_run_node_updates_per_s": round(updates / (t_setup + wall), 1)}
    if args.copy_calib:
        copy_calibration(h)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(h, qq, ng, args.cpu_steps, args.ell_cells)
        except Exception as e:  # the baseline is a report, not the product
            log(f"cpu baseline failed: {e}")
    regs = regularisation_costs(h)
    log(f"regularisation: {regs}")
    h.close()
    mixed = None
    if args.precision == 0 and not args.no_mixed:
        mixed = mixed_report(qq, ng, args.steps, args.warmup, args.err_steps, value,
                             args.ell_cells)
    refdm = None
    if args.ref_dm_steps > 0 and args.ell_cells != 6.25:
        g, ts = make_flow(ng, qq, args.precision, 6.25)
        u, w, phr = timed_run(g, args.ref_dm_steps, args.warmup)
        refdm = {"ell_cells": 6.25, "dm": "0.0625*glx (contours.f90:38)", "nodes": g.counts()[0],
                 "steps": args.ref_dm_steps, "value": round(u / w, 1),
                 "ms_per_step": round(1e3 * w / args.ref_dm_steps, 4),
                 "surgeries": phr["surgery"][0], "setup_s": round(ts, 3)}
        log(f"reference dm: {refdm}")
        g.close()
    out = {
        "metric": "contour node-updates/s at ng=1024: plain FP64 vs FP32/batched at matched error",
        "value": round(value, 1),
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32/f64",
        "data": "synthetic",
        "config": {"workload": ("plain FP64" if args.precision == 0 else "mixed FP32-grid/FP64-state")
                   + f" CAPS advect loop (RK4 step + surgery + recontouring), 2D Euler, ng={ng}, "
                     f"80 PV levels, seeded k^3 exp(-2(k/6)^2) field (rms 4pi), "
                     f"dm={dm / (2 * np.pi / ng):.4g}*glx",
                   "ng": ng, "nodes": P, "nodes_end": P_end, "nodes_mean": round(p_avg, 1),
                   "contours": nq, "ell_cells": args.ell_cells,
                   "precision": "plain-fp64" if args.precision == 0 else "mixed-fp32-grid",
                   "fft_batch_autotuned": tuning,
                   "dt": "adaptive (reference adapt)", "seed": args.seed},
        "roofline": roofline,
        "step_roofline": step_roof,
        "run": run,
        "regularisation": regs,
        "kernel_groups": groups,
        "cpu_baseline": cpu,
    }
    if mixed is not None:
        out["mixed"] = mixed
    if refdm is not None:
        out["reference_dm"] = refdm
    if args.ensemble > 0:
        out["ensemble"] = ensemble_report(args.ensemble, args.ensemble_steps, args.seed)
    print(json.dumps(out), flush=True)


def mixed_report(qq, ng, steps, warmup, err_steps, plain_value, ell_cells):
    """The mixed-precision step (FP32 grids and transforms, FP64 state and
    nodes) on the same workload: its throughput over the same loop, and its
    error against the FP64 step after err_steps RK4 steps from the same
    contours (no surgery in between, so the nodes correspond one to one),
    evaluated with the FP64 diagnostics (hydra_amd.diagnostics)."""
    from hydra_amd.diagnostics import error_vs_fp64, fp64_diagnostics
    h, _ = make_flow(ng, qq, 1, ell_cells)
    updates, wall, ph = timed_run(h, steps, warmup)
    h.close()
    h64, _ = make_flow(ng, qq, 0, ell_cells)
    c = h64.contours()
    h32, _ = make_flow(ng, qq, 1, ell_cells)
    h32.set_contours(c)
    h32.evolution_init()
    _, e0, z0 = fp64_diagnostics(h64)
    h64.evolution_init()
    for _ in range(err_steps):
        h64.advance()
        h32.advance()
    err = error_vs_fp64(h64, h32, e0, z0)
    err["steps"] = err_steps
    h32.close()
    h64.close()
    log(f"mixed: {steps} steps {wall:.3f}s; after {err_steps} steps {err}")
    return {"value": round(updates / wall, 1), "unit": "node-updates/s",
            "ms_per_step": round(1e3 * wall / steps, 4),
            "speedup_vs_plain": round(updates / wall / plain_value, 3),
            "surgeries": ph["surgery"][0],
            "dtype": "f32 grids/f64 nodes and state",
            "step_roofline_frac": round(step_bytes(ng, updates / steps, 4)
                                        / (wall / steps) / 1e9 / HBM_PEAK_GBS, 4),
            "error_vs_fp64": err}


def ensemble_report(nflows, steps, seed):
    """Aggregate node-updates/s of an ensemble of independent mixed-precision
    flows of uneven sizes: all advanced together by hyd_run_batch, and the
    same flows advanced one after another with hyd_run.  Both start each
    flow with a recontouring (hyd_run's loop), included in the time.  Nodes
    per step are taken as the mean of each flow's counts before and after."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    rng = np.random.default_rng(seed + 1)
    spec = [(int(rng.choice([256, 512, 1024])), float(rng.uniform(2 * np.pi, 6 * np.pi)),
             int(rng.integers(40, 121)), int(rng.integers(1 << 30))) for _ in range(nflows)]

    def make():
        hs = []
        for ng, rms, lev, sd in spec:
            h = Hydra(ng, ncontq=lev, precision=1, node_capacity=max(8 * ng * ng, 1 << 20))
            h.set_field(random_pv(ng, seed=sd, rms=rms))
            hs.append(h)
        return hs

    def timed(run):
        hs = make()
        run(hs, 2)  # warm-up: first contours, graph capture
        for h in hs:
            h.sync()
            h.advance_counts(reset=True)
        t0 = time.perf_counter()
        run(hs, steps)
        for h in hs:
            h.sync()
        wall = time.perf_counter() - t0
        p1 = [h.counts()[0] for h in hs]
        updates = sum(h.advance_counts()[1] for h in hs)
        for h in hs:
            h.close()
        return updates / wall, wall, p1

    def one_by_one(hs, k):
        for h in hs:
            h.run(k)

    v_batch, w_batch, nodes = timed(lambda hs, k: Hydra.run_batch(hs, k))
    v_seq, w_seq, _ = timed(one_by_one)
    log(f"ensemble: {nflows} flows, {steps} steps: batched {w_batch:.3f}s, one by one {w_seq:.3f}s")
    return {"flows": [{"ng": ng, "rms": round(rms, 3), "levels": lev, "nodes": p}
                      for (ng, rms, lev, _), p in zip(spec, nodes)],
            "steps": steps, "value_batched": round(v_batch, 1),
            "value_one_by_one": round(v_seq, 1), "unit": "node-updates/s",
            "precision": "mixed-fp32-grid"}


def cpu_baseline(h, qq, ng, nsteps, ell_cells):
    """The oracle (C restatement, gcc -O3, one core) advancing the same
    state through the same loop for a bounded number of steps; returns
    node-updates/s."""
    from oracle.hyo import Oracle, build
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build()
    c = h.contours()
    o = Oracle(ng, ell_cells=ell_cells)
    o.init_from_field(qq)
    o.set_contours(c)
    o.advect_start()
    P = len(c["x"])
    updates = 0
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.run(1)
        updates += o.get("nptq")
    wall = time.perf_counter() - t0
    return {"value": round(updates / wall, 1), "unit": "node-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{nsteps} steps of the advect loop (surgery + RK4 advance) of the oracle "
                      f"(C restatement of evolution.f90/contours.f90, gcc -O3 -ffp-contract=off, "
                      f"one thread) from the same ng={ng} contours ({P} nodes), {wall:.1f} s"}


if __name__ == "__main__":
    main()
