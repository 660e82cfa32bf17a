"""Pins of the CPU oracle of the learned velocity correction (no reference
counterpart): BF16 rounding against known values, the exact network's
gradients against central finite differences (a wrong tap sign, a
transposed weight or a dropped GELU factor fails them), and the sampling
against a closed form."""
import numpy as np
import pytest

from oracle import cnn_oracle as K


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -3.140625, 0.0],
                 np.float32)
    r = K.bf16(x)
    # 2^-8 above 1 is a tie: to even (1.0); 3*2^-8: tie to even upward (1 + 2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == 1.0
    assert r[4] == -3.140625 and r[5] == 0.0


def _problem(n=8, m=40, seed=0):
    rng = np.random.default_rng(seed)
    in3 = rng.normal(size=(n * n, 3)).astype(np.float32)
    L = 2 * np.pi
    xn = rng.uniform(-np.pi, np.pi, m)
    yn = rng.uniform(-np.pi, np.pi, m)
    tu = rng.normal(size=m) * 0.1
    tv = rng.normal(size=m) * 0.1
    return in3, xn, yn, tu, tv, n, L


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_exact_gradients_match_finite_differences(layer):
    in3, xn, yn, tu, tv, n, L = _problem()
    w = K.init_weights(3, scale_last=1.0).astype(np.float64)
    sl, nw = K.layer_slices()
    _, g = K.loss_and_grad(w, in3, xn, yn, tu, tv, n, L, bf=False)
    rng = np.random.default_rng(layer)
    ws, bs = sl[layer]
    picks = list(rng.integers(ws.start, ws.stop, 6)) + [bs.start, bs.stop - 1]
    for k in picks:
        if layer == 0 and ((k - ws.start) // 9) % 4 == 3 and k < ws.stop:
            continue  # the zero-padding input channel
        h = 1e-5 * max(1.0, abs(w[k]))
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        lp, _ = K.loss_and_grad(wp, in3, xn, yn, tu, tv, n, L, bf=False)
        lm, _ = K.loss_and_grad(wm, in3, xn, yn, tu, tv, n, L, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (layer, k, fd, g[k])


def test_sampling_is_bilinear_and_periodic():
    """An output that is linear in ix inside a cell is sampled exactly; a
    node beyond the last grid point interpolates towards ix = 0."""
    n, L = 8, 2 * np.pi
    out = np.zeros((n, n))
    out[3, 5], out[4, 5], out[3, 6], out[4, 6] = 1.0, 3.0, 5.0, 7.0
    h = L / n
    x, y = -np.pi + 3.25 * h, -np.pi + 5.5 * h
    idx, w = K._bilinear(np.array([x]), np.array([y]), n, L)
    val = sum(w[j] * out[idx[j][0], idx[j][1]] for j in range(4))
    assert val[0] == pytest.approx(((1 * 0.75 + 3 * 0.25) + (5 * 0.75 + 7 * 0.25)) / 2)
    idx, w = K._bilinear(np.array([-np.pi + 7.5 * h]), np.array([-np.pi]), n, L)
    assert idx[1][0][0] == 0 and w[1][0] == pytest.approx(0.5)
