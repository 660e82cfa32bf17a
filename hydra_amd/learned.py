"""Training of the learned sub-grid velocity correction (no reference
counterpart: the reference always uses bilinear velint, contours.f90:136-169).

The coarse flow runs at ng; its contours' PV is put on a grid `factor`
times finer and inverted there by the library's own con2grid, FFT and
spectral inversion (Hydra.contour_velocity on a second context): that
velocity at the nodes is the judge.  The target of the network is the
judge minus the coarse grid's velocity of the same contours at the same
nodes, i.e. the part of the node velocity the coarse grid cannot resolve;
the network's inputs are the coarse step's own fields (u, v and the
physical PV), scaled by the flow's rms velocity su and rms PV sq, which is
exactly what the time step feeds it when the network is attached with
Hydra.set_correction(net, su, sq, sd); targets and output are in units of
sd, the rms of the error.
"""
import numpy as np

from . import Hydra, VelocityCNN


class Snapshot:
    """One training pair: in3 [ng*ng, 3] inputs (u/su, v/su, q/sq), node
    positions (xn, yn) and targets (tu, tv) in units of sd."""

    __slots__ = ("in3", "xn", "yn", "tu", "tv")

    def __init__(self, in3, xn, yn, tu, tv):
        self.in3, self.xn, self.yn, self.tu, self.tv = in3, xn, yn, tu, tv


def judge_for(flow, factor=4, node_capacity=None):
    """A context on the grid `factor` times finer than the flow's."""
    ng = flow.ng * factor
    return Hydra(ng, node_capacity=node_capacity or max(flow.counts()[0] * 4, 1 << 22))


class Scales:
    """su: rms velocity, sq: rms PV of the flow's grid (the inputs are
    divided by them); sd: rms of the coarse grid's node-velocity error (the
    targets are divided by it, the network's output multiplied)."""

    __slots__ = ("su", "sq", "sd")

    def __init__(self, su, sq, sd):
        self.su, self.sq, self.sd = su, sq, sd


def _inputs(flow, sc):
    flow.inversion()
    uu, vv = flow.field("uu"), flow.field("vv")
    q = flow.spctop(flow.field("qq"))
    if sc is None:
        return uu, vv, q
    return np.stack([uu / sc.su, vv / sc.su, q / sc.sq], axis=-1).reshape(-1, 3).astype(np.float32)


def _errors(flow, judge, max_nodes, rng):
    c = flow.contours()
    x, y, nxt, qjump = c["x"], c["y"], c["nextq"], c["qjump"]
    xs, ys = x, y
    if max_nodes is not None and len(x) > max_nodes:
        pick = np.sort(rng.choice(len(x), max_nodes, replace=False))
        xs, ys = np.ascontiguousarray(x[pick]), np.ascontiguousarray(y[pick])
    uf, vf = judge.contour_velocity(x, y, nxt, qjump, xs, ys)
    uc, vc = flow.contour_velocity(x, y, nxt, qjump, xs, ys)
    return xs, ys, uf - uc, vf - vc


def flow_scales(flow, judge, max_nodes=20000):
    """The Scales of the flow's current state."""
    uu, vv, q = _inputs(flow, None)
    _, _, eu, ev = _errors(flow, judge, max_nodes, np.random.default_rng(0))
    return Scales(float(np.sqrt(np.mean(uu * uu + vv * vv))), float(np.sqrt(np.mean(q * q))),
                  float(np.sqrt(np.mean(eu * eu + ev * ev))))


def snapshot(flow, judge, sc, max_nodes=None, rng=None):
    """The training pair of the flow's current state (see the module text).
    With max_nodes, a random subset of that many nodes is kept."""
    xs, ys, eu, ev = _errors(flow, judge, max_nodes, rng or np.random.default_rng(0))
    return Snapshot(_inputs(flow, sc), xs, ys, (eu / sc.sd).astype(np.float32),
                    (ev / sc.sd).astype(np.float32))


def collect(flow, judge, nsnap, steps_between, sc, max_nodes=None, seed=0):
    """nsnap snapshots of a running flow, steps_between advect-loop passes
    apart (surgery and recontouring included, as in Hydra.run)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsnap):
        if i:
            flow.run(steps_between)
        out.append(snapshot(flow, judge, sc, max_nodes, rng))
    return out


def rescale(snaps, sc, extra=()):
    """Re-fix the output scale sd as the rms target over the training
    snapshots (snapshots of evolved flows carry larger errors than the first
    state the scale came from); every snapshot's targets, and those of the
    extra (held-out) sets, are rescaled to match.  Returns the new Scales."""
    ss = sum(float(np.sum(s.tu.astype(np.float64) ** 2 + s.tv.astype(np.float64) ** 2))
             for s in snaps)
    cnt = sum(len(s.tu) for s in snaps)
    r = float(np.sqrt(ss / max(cnt, 1)))
    if not (r > 0.0):
        return sc
    for s in list(snaps) + [t for group in extra for t in group]:
        s.tu = (s.tu / r).astype(np.float32)
        s.tv = (s.tv / r).astype(np.float32)
    return Scales(sc.su, sc.sq, sc.sd * r)


def train(net, snaps, iters, lr=2e-3, lr_final=5e-4, decay_at=0.67):
    """Adam on the snapshots in turn; returns the loss history."""
    hist = []
    for it in range(iters):
        s = snaps[it % len(snaps)]
        hist.append(net.train_step(s.in3, s.xn, s.yn, s.tu, s.tv,
                                   lr=lr if it < decay_at * iters else lr_final))
    return hist


def residual_ratio(net, snaps):
    """||target - network|| / ||target|| over the snapshots: the fraction of
    the coarse grid's node-velocity error left after the correction."""
    before = after = 0.0
    for s in snaps:
        du, dv = net.forward(s.in3, s.xn, s.yn)
        tu, tv = s.tu.astype(np.float64), s.tv.astype(np.float64)
        before += np.sum(tu * tu + tv * tv)
        after += np.sum((tu - du) ** 2 + (tv - dv) ** 2)
    return float(np.sqrt(after / before))


def new_network(flow, seed=7):
    return VelocityCNN(flow.ng, ell=flow.params.ellx, seed=seed)


# 3x3 convolutions of 3 -> 32 -> 32 -> 32 -> 2 channels: multiply-adds per pixel
FLOPS_PER_PIXEL = 2 * 9 * (3 * 32 + 32 * 32 + 32 * 32 + 32 * 2)


def main(argv=None):
    """Train a correction for an ng flow against the factor-times finer judge,
    report the held-out error left, and time the step with and without it."""
    import argparse
    import json
    import time
    from .ic import random_pv

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--ng", type=int, default=512)
    ap.add_argument("--factor", type=int, default=4)
    ap.add_argument("--snapshots", type=int, default=6, help="training snapshots in all")
    ap.add_argument("--flows", type=int, default=1,
                    help="seeded flows the training snapshots are drawn from")
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--nodes", type=int, default=60000)
    ap.add_argument("--iters", type=int, default=360)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=5e-4)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--save", default=None, help="write the trained weights (.npy)")
    a = ap.parse_args(argv)

    def flow(seed):
        h = Hydra(a.ng, node_capacity=max(16 * a.ng * a.ng, 1 << 22))
        h.set_field(random_pv(a.ng, seed=seed))
        h.recontour()
        h.evolution_init()
        return h

    f = flow(a.seed)
    judge = judge_for(f, a.factor)
    sc = flow_scales(f, judge)
    t0 = time.perf_counter()
    per = -(-a.snapshots // a.flows)
    train_set = []
    for k in range(a.flows):
        fk = f if k == 0 else flow(a.seed + 1000 + k)
        train_set += collect(fk, judge, min(per, a.snapshots - len(train_set)), a.every, sc,
                             a.nodes, seed=1 + k)
        if fk is not f:
            fk.close()
    test_set = collect(flow(a.seed + 1), judge, 2, a.every, sc, a.nodes, seed=2)
    sd_first = sc.sd
    sc = rescale(train_set, sc, extra=(test_set,))
    t1 = time.perf_counter()
    net = new_network(f)
    hist = train(net, train_set, a.iters, lr=a.lr, lr_final=a.lr_final)
    t2 = time.perf_counter()
    out = dict(ng=a.ng, judge_ng=judge.ng, su=sc.su, sq=sc.sq, sd=sc.sd, sd_first_state=sd_first,
               snapshots=len(train_set), flows=a.flows, snapshot_s=t1 - t0,
               train_s=t2 - t1, iters=a.iters, loss_first=hist[0], loss_last=hist[-1],
               train_left=residual_ratio(net, train_set[:2]),
               held_out_left=residual_ratio(net, test_set),
               gflop_per_forward=FLOPS_PER_PIXEL * a.ng * a.ng / 1e9)
    if a.save:
        np.save(a.save, net.weights())
    for tag in ("plain", "corrected"):
        g = flow(a.seed + 2)
        if tag == "corrected":
            g.set_correction(net, sc.su, sc.sq, sc.sd)
        g.run(2)
        g.sync()
        t = time.perf_counter()
        g.run(a.steps)
        g.sync()
        out[f"ms_per_step_{tag}"] = 1e3 * (time.perf_counter() - t) / a.steps
        out[f"nodes_{tag}"] = g.counts()[0]
        g.set_correction(None)
        g.close()
    print(json.dumps(out), flush=True)
    net.close()


if __name__ == "__main__":
    main()
