"""The learned velocity correction in the time step (no reference
counterpart): trained on an ng=512 flow against the velocity its contours
have on an ng=2048 grid (the library's own con2grid, FFT and inversion),
it must remove part of the coarse grid's node-velocity error on a flow it
has not seen; attached to a flow, the step must move the nodes by dt times
the correction, and detached, the step must be the reference's again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NG = 512


def _flow(seed):
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    h = Hydra(NG, node_capacity=1 << 22)
    h.set_field(random_pv(NG, seed=seed))
    h.recontour()
    h.evolution_init()
    return h


@pytest.fixture(scope="module")
def trained():
    from hydra_amd import learned as LN
    flow = _flow(11)
    judge = LN.judge_for(flow)
    assert judge.ng == 4 * NG
    sc = LN.flow_scales(flow, judge)
    train = LN.collect(flow, judge, 6, 8, sc, max_nodes=60000, seed=1)
    other = _flow(12)
    test = LN.collect(other, judge, 2, 8, sc, max_nodes=60000, seed=2)
    net = LN.new_network(flow)
    LN.train(net, train, 360)
    yield dict(net=net, sc=sc, test=test, train=train)
    net.close()


def test_judge_targets_are_subgrid():
    """The coarse grid's error is small against the velocity itself (the
    judge and the coarse grid agree to a few per cent) but not zero."""
    from hydra_amd import learned as LN
    flow = _flow(13)
    judge = LN.judge_for(flow)
    sc = LN.flow_scales(flow, judge)
    err = sc.sd / sc.su
    print(f"rms coarse-grid node-velocity error / rms velocity: {err:.4f}")
    assert 1e-4 < err < 0.2, err


def test_training_reduces_held_out_error(trained):
    from hydra_amd import learned as LN
    r_train = LN.residual_ratio(trained["net"], trained["train"][:2])
    r_test = LN.residual_ratio(trained["net"], trained["test"])
    print(f"error left after the correction: train {r_train:.3f}, held-out {r_test:.3f}")
    assert r_test < 0.95, r_test


def test_attached_step_moves_nodes_by_the_correction(trained):
    """One RK4 step with the network attached against the same step without:
    the node displacement differs by about dt * (du, dv) and is finite; after
    detaching, the step is bitwise the plain one."""
    net, sc = trained["net"], trained["sc"]
    start = _flow(21)
    c0 = start.contours()
    start.inversion()
    uu, vv = start.field("uu"), start.field("vv")
    q = start.spctop(start.field("qq"))
    du, dv = net.forward(np.stack([uu / sc.su, vv / sc.su, q / sc.sq], axis=-1).reshape(-1, 3),
                         c0["x"], c0["y"])
    a, b = _flow(21), _flow(21)
    a.set_correction(net, sc.su, sc.sq, sc.sd)
    a.advance()
    b.advance()
    ca, cb = a.contours(), b.contours()
    assert len(ca["x"]) == len(cb["x"]) == len(c0["x"])
    dx, dy = ca["x"] - cb["x"], ca["y"] - cb["y"]
    assert np.all(np.isfinite(dx)) and np.all(np.isfinite(dy))
    dt = b.scalar("dt")
    expect = dt * sc.sd * np.sqrt(np.mean(du.astype(np.float64) ** 2 + dv.astype(np.float64) ** 2))
    got = np.sqrt(np.mean(dx ** 2 + dy ** 2))
    print(f"rms displacement from the correction {got:.3e}, dt*rms(du,dv) {expect:.3e}")
    assert expect > 0
    assert 0.5 * expect < got < 2.0 * expect
    d = _flow(21)
    d.set_correction(net, sc.su, sc.sq, sc.sd)
    d.set_correction(None)
    d.advance()
    assert np.array_equal(d.contours()["x"], cb["x"])
    assert np.array_equal(d.contours()["y"], cb["y"])


def test_correction_in_the_mixed_step(trained):
    """The network attached to a mixed-precision flow (FP32 grids, captured
    step): the step runs, the nodes stay finite and move by about dt times
    the correction against the same mixed step without it."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    net, sc = trained["net"], trained["sc"]

    def mixed():
        h = Hydra(NG, node_capacity=1 << 22, precision=1)
        h.set_field(random_pv(NG, seed=21))
        h.recontour()
        h.evolution_init()
        return h

    a, b = mixed(), mixed()
    a.set_correction(net, sc.su, sc.sq, sc.sd)
    for _ in range(3):
        a.advance()
        b.advance()
    ca, cb = a.contours(), b.contours()
    assert len(ca["x"]) == len(cb["x"])
    d = np.hypot(ca["x"] - cb["x"], ca["y"] - cb["y"])
    assert np.all(np.isfinite(d))
    assert 0.0 < np.sqrt(np.mean(d ** 2)) < 1e-3


def test_correction_rejects_other_grid():
    from hydra_amd import HydraError, VelocityCNN
    h = _flow(31)
    net = VelocityCNN(256)
    with pytest.raises(HydraError):
        h.set_correction(net, 1.0, 1.0)
    net.close()
