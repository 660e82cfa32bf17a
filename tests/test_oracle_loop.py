"""Pins of the oracle's loop control: one pass of advect's time loop
(evolution.f90:55-80) and the caps.f90:83-91 cycle of recontouring and
advecting, each against the sequence of pinned pieces (surgery, advance,
prepare, recontour, init) that the reference's text prescribes, run step by
step from here.  Also: a run split into two calls equals one run of the
summed step count (a run resumes inside advect's loop)."""
import numpy as np
import pytest

from hydra_amd.ic import random_pv
from oracle.hyo import Oracle
from tests.helpers import bits_equal

NG = 32
TWISTMAX = 2.5   # evolution.f90:36
NREGMAX = 20     # evolution.f90:38


def started(ng=NG, steps=3, seed=12345):
    """caps.f90 initialise + recont + advect's start, then a few steps."""
    o = Oracle(ng)
    o.init_from_field(random_pv(ng, seed=seed))
    o.recontour(o.d("qr").copy())
    o.advect_start()
    for _ in range(steps):
        o.advance()
    return o


def state(o):
    n = o.get("nptq")
    nc = o.get("nq")
    out = {k: o.d(k).copy() for k in ("qs", "qd", "qc", "qr")}
    out.update(x=o.d("xq", n).copy(), y=o.d("yq", n).copy(), nextq=o.i("nextq", n).copy(),
               **{k: o.i(k, nc).copy() for k in ("i1q", "i2q", "npq", "indq")})
    for k in ("t", "dt", "twist"):
        out[k] = np.float64(o.getd(k))
    out["ireg"] = o.get("ireg")
    return out


def assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert bits_equal(a[k], b[k]), k


@pytest.mark.parametrize("twist", [TWISTMAX, TWISTMAX - 1e-3])
def test_no_regularisation_at_or_below_twistmax(twist):
    """twist <= twistmax (the test is strict, evolution.f90:58): the pass is
    one advance and nothing else."""
    a, b = started(), started()
    for o in (a, b):
        o.set("twist", twist)
        o.set("ireg", 3)
    assert a.advect_step() == 0
    b.advance()
    assert_same(state(a), state(b))
    assert a.get("ireg") == 3


def test_surgery_above_twistmax():
    """twist > twistmax: ireg+1, surgery, twist - twistmax, then advance."""
    a, b = started(), started()
    tw = TWISTMAX + 0.7
    for o in (a, b):
        o.set("twist", tw)
        o.set("ireg", 5)
    assert a.advect_step() == 0
    b.surgery()
    b.set("twist", tw - TWISTMAX)
    b.set("ireg", 6)
    b.advance()
    assert_same(state(a), state(b))
    assert a.get("ireg") == 6
    # the advance adds dt * zzmax to what surgery left (adapt, evolution.f90:578)
    assert a.getd("twist") == pytest.approx(tw - TWISTMAX + a.getd("dt") * a.getd("zzmax"),
                                            rel=1e-15)


def test_last_regularisation_hands_over():
    """The nregmax-th regularisation does no surgery and no advance: prepare,
    then the loop returns for recontouring (evolution.f90:62-67)."""
    a, b = started(), started()
    for o in (a, b):
        o.set("twist", TWISTMAX + 0.1)
        o.set("ireg", NREGMAX - 1)
    t0 = a.getd("t")
    assert a.advect_step() == 1
    b.prepare()
    b.set("ireg", NREGMAX)
    assert_same(state(a), state(b))
    assert a.getd("t") == t0
    assert a.get("phase") == 0


def reference_cycle(o, nsteps):
    """caps.f90:83-91 (recont, evolve) with advect's loop (evolution.f90:
    44-80) written out from the reference's text, stopping after nsteps
    advances; returns the step numbers at which surgery and recontouring
    happened."""
    done, surg, recont = 0, [], []
    while done < nsteps:
        o.recontour(o.d("qr").copy())
        recont.append(done)
        o.evolution_init()
        twist, ireg = 0.0, 0
        o.set("twist", 0.0)
        while done < nsteps:
            twist = o.getd("twist")
            if twist > TWISTMAX:
                ireg += 1
                if ireg == NREGMAX:
                    o.prepare()
                    break
                o.surgery()
                surg.append(done)
                o.set("twist", twist - TWISTMAX)
            o.advance()
            done += 1
    return surg, recont


@pytest.fixture(scope="module")
def long_run():
    """The reference's cycle over enough steps for regularisations and a
    second recontouring at ng=32 (the 20th regularisation comes at step 160),
    and the oracle's run over the same steps."""
    n = 175
    o = Oracle(NG)
    o.init_from_field(random_pv(NG, seed=12345))
    surg, recont = reference_cycle(o, n)
    r = Oracle(NG)
    r.init_from_field(random_pv(NG, seed=12345))
    r.run(n)
    return n, state(o), state(r), surg, recont


def test_run_follows_the_reference_cycle(long_run):
    n, want, got, surg, recont = long_run
    assert len(surg) >= 5 and len(recont) >= 2   # both branches are exercised
    want = dict(want, ireg=got["ireg"])  # reference_cycle keeps ireg locally
    assert_same(got, want)


@pytest.mark.parametrize("split", [1, 37, 153, 160])
def test_run_in_two_calls_equals_one_run(long_run, split):
    """run(a) then run(b) is run(a+b): the second call resumes inside the
    loop instead of recontouring from a stale residual (split 160 ends the
    first call just before the 20th regularisation hands over)."""
    n, _, whole, _, _ = long_run
    a = Oracle(NG)
    a.init_from_field(random_pv(NG, seed=12345))
    a.run(split)
    a.run(n - split)
    assert_same(state(a), whole)
