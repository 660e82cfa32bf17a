"""Contour node-updates/s of the plain FP64 contour-advection step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--ng 1024] [--gpus 1]

Workload: the doubly periodic 2D Euler CAPS step at ng=1024 (contour grid
16ng for recontouring, 4ng inside con2grid), a seeded random PV field with
1D spectrum ~ k^3 exp(-2 (k/6)^2) and rms 4*pi, contoured into 80 levels
on the GPU (congen), then K RK4 steps (evolution.f90 advance) of the
device-resident state.  One node-update = one contour node advanced through
one full RK4 step (four stages: con2grid, FFT inversion, sources, velint,
node and spectral updates).  Inputs are resident in HBM before timing.

Besides the JSON line the script prints nothing on stdout.  The roofline
entry is for the kernel with the largest share of the timed region, timed
with HIP events on the library's stream; cpu_baseline is the oracle (C
restatement of the reference, one core) on a bounded sample of the same
state.
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


# HIP-event record name -> kernel symbol as rocprofv3 reports it
KERNEL_SYMBOL = {"k:fft_fwd": "fft_kernel<N, true>", "k:fft_rev": "fft_kernel<N, false>"}


def traffic_from_profile(kernel, ng):
    """HBM bytes per launch of `kernel` at this ng, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel
    of known size; see profiles/), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    rec = d.get(f"{kernel}@ng{ng}")
    return None if rec is None else rec["traffic_bytes"]


COPY_BYTES = 256 << 20  # each of read and write of the calibration copy


def copy_calibration():
    """A plain device copy of known size (read COPY_BYTES, write COPY_BYTES):
    the reference point that scales the FETCH_SIZE/WRITE_SIZE counters."""
    import torch
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()


def pmc_summary(fetch_dir, write_dir, ng, out_json):
    """Per-launch HBM traffic of the FFT kernels from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by the same
    counter on the calibration copy."""
    import csv
    import glob

    def per_kernel(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    k = row["Kernel_Name"]
                    key = ("copy" if "elementwise" in k or "mul" in k.lower() else
                           "fft_kernel<N, true>" if "fft_kernel<" in k and "true>" in k else
                           "fft_kernel<N, false>" if "fft_kernel<" in k and "false>" in k else None)
                    if key:
                        vals.setdefault(key, []).append(float(row["Counter_Value"]))
        return {k: float(np.mean(v)) for k, v in vals.items()}

    fe = per_kernel(fetch_dir, "FETCH_SIZE")
    wr = per_kernel(write_dir, "WRITE_SIZE")
    kib = COPY_BYTES / 1024.0
    ff = kib / fe["copy"]
    fw = kib / wr["copy"]
    out = {"calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": fe["copy"], "write_kib": wr["copy"],
                           "fetch_factor": ff, "write_factor": fw}}
    N = ng * ng
    for k in ("fft_kernel<N, true>", "fft_kernel<N, false>"):
        if k in fe and k in wr:
            t = (fe[k] * ff + wr[k] * fw) * 1024.0
            out[f"{k}@ng{ng}"] = {"fetch_kib": fe[k], "write_kib": wr[k], "traffic_bytes": round(t),
                                  "algorithmic_bytes": 16 * N}
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ng", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", type=int, default=0,
                    help="0 plain FP64 (bit-exact), 1 mixed FP32 grid / FP64 state")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run the calibration copy (for rocprofv3 --pmc runs)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs into profiles/pmc_traffic.json")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.ng, os.path.join(ROOT, "profiles", "pmc_traffic.json"))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs one flow on one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv

    ng = args.ng
    qq = random_pv(ng, seed=args.seed)
    h = Hydra(ng, ncontq=80, precision=args.precision)
    t0 = time.perf_counter()
    h.set_field(qq)
    h.recontour()
    h.evolution_init()
    h.sync()
    t_setup = time.perf_counter() - t0
    P, nq = h.counts()
    tuning = h.tuning()
    log(f"setup: ng={ng} contours={nq} nodes={P} recontour+init {t_setup:.2f}s")
    N = ng * ng

    def step():
        """One pass of the advect loop (evolution.f90:55-80): surgery when the
        twist exceeds 2.5, recontouring after 20 regularisations
        (caps.f90:83-91), then one RK4 advance.  Returns the nodes advanced."""
        while h.advect_step() == 1:
            h.recontour()
            h.evolution_init()
        return h.counts()[0]

    for _ in range(args.warmup):
        step()
    h.sync()
    # timed region: K steps; HIP events around each launch of the single
    # kernels ("k:" records) on the library stream
    h.set_timing(1)
    h.sync()
    t0 = time.perf_counter()
    updates = 0
    for _ in range(args.steps):
        updates += step()
    h.sync()
    wall = time.perf_counter() - t0
    recs = h.kernel_times()
    h.set_timing(0)
    P_end, nq_end = h.counts()
    ms_step = 1e3 * wall / args.steps
    value = updates / wall
    log(f"timed: {args.steps} steps {wall:.3f}s, nodes {P} -> {P_end}, contours {nq} -> {nq_end}")

    tot = {}
    cnt = {}
    for name, ms, by in recs:
        tot[name] = tot.get(name, 0.0) + ms
        cnt[name] = cnt.get(name, 0) + 1
    # breakdown by kernel group (after the timed region, all records on)
    nb = 2
    h.set_timing(2)
    for _ in range(nb):
        step()
    h.sync()
    gt, gc = {}, {}
    for name, ms, by in h.kernel_times():
        gt[name] = gt.get(name, 0.0) + ms
        gc[name] = gc.get(name, 0) + 1
    h.set_timing(0)
    for name, ms in sorted(gt.items(), key=lambda kv: -kv[1]):
        log(f"  {name:16s} {ms / nb:8.3f} ms/step  {gc[name] // nb} calls/step")
    # roofline: the single kernel ("k:" records, one record per launch) with
    # the largest share of the timed region; its algorithmic bytes per launch
    # are recorded with each launch
    kbytes = {}
    for name, ms, by in recs:
        if name.startswith("k:"):
            kbytes.setdefault(name, []).append(by)
    dom = max(kbytes, key=lambda k: tot[k])
    avg_ms = tot[dom] / cnt[dom]
    alg = float(np.mean(kbytes[dom]))
    achieved = alg / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": KERNEL_SYMBOL.get(dom, dom), "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profile(KERNEL_SYMBOL.get(dom, dom), ng),
                "avg_launch_ms": round(avg_ms, 4), "launches_per_step": cnt[dom] / args.steps,
                "algorithmic_bytes_per_launch": int(alg)}

    if args.copy_calib:
        copy_calibration()
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(h, qq, ng, args.cpu_steps)
        except Exception as e:  # the baseline is a report, not the product
            log(f"cpu baseline failed: {e}")
    h.close()
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
        "config": {"workload": f"plain FP64 CAPS RK4 step, 2D Euler, ng={ng}, 80 PV levels, "
                               f"seeded k^3 exp(-2(k/6)^2) field (rms 4pi)",
                   "ng": ng, "nodes": P, "nodes_end": P_end, "contours": nq,
                   "precision": "plain-fp64" if args.precision == 0 else "mixed-fp32-grid",
                   "fft_batch_autotuned": tuning,
                   "dt": "adaptive (reference adapt)", "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(h, qq, ng, nsteps):
    """The oracle (C restatement, one core) advancing the same state for a
    bounded number of steps; returns node-updates/s."""
    import subprocess
    from oracle.hyo import Oracle, build
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build()
    c = h.contours()
    o = Oracle(ng)
    o.init_from_field(qq)
    o.set_contours(c)
    o.evolution_init()
    o.set("twist", 0.0)
    o.set("ireg", 0)
    P = len(c["x"])
    updates = 0
    t0 = time.perf_counter()
    for _ in range(nsteps):
        while o.advect_step() == 1:
            o.recontour(o.d("qr").copy())
            o.evolution_init()
            o.set("twist", 0.0)
            o.set("ireg", 0)
        updates += o.get("nptq")
    wall = time.perf_counter() - t0
    return {"value": round(updates / wall, 1), "unit": "node-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{nsteps} steps of the advect loop (surgery + RK4 advance) of the oracle "
                      f"(C restatement of evolution.f90/contours.f90, gcc -O2, one thread) from "
                      f"the same ng={ng} contours ({P} nodes), {wall:.1f} s"}


if __name__ == "__main__":
    main()
