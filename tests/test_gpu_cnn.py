"""The learned velocity correction on the GPU (hyd_cnn_*; no reference
counterpart) against its CPU oracle: the matrix-core operand layout, the
forward pass, the gradients of every weight (BF16 rounding at the same
points, so they agree to BF16 accuracy), and a training run showing what
the trained network gains over plain bilinear velint."""
import numpy as np
import pytest

from oracle import cnn_oracle as K

pytestmark = pytest.mark.gpu

L = 2 * np.pi


def _bf16_bits(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def test_mfma_operand_layout():
    from hydra_amd import mfma_probe
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 32)).astype(np.float32)
    B = rng.normal(size=(32, 16)).astype(np.float32)  # B[k][n]
    Ab, Bb = _bf16_bits(A), _bf16_bits(B)
    D = mfma_probe(Ab, Bb.T.copy())  # the probe takes B column by column
    ref = K.bf16(A) @ K.bf16(B)
    assert np.abs(D - ref).max() < 1e-4 * np.abs(ref).max()


def _problem(n, m, seed):
    rng = np.random.default_rng(seed)
    in3 = rng.normal(size=(n * n, 3)).astype(np.float32)
    xn = rng.uniform(-np.pi, np.pi, m)
    yn = rng.uniform(-np.pi, np.pi, m)
    tu = (rng.normal(size=m) * 0.1).astype(np.float32)
    tv = (rng.normal(size=m) * 0.1).astype(np.float32)
    return in3, xn, yn, tu, tv


def test_forward_and_gradients_match_oracle():
    from hydra_amd import VelocityCNN
    n, m = 16, 300
    in3, xn, yn, tu, tv = _problem(n, m, 1)
    net = VelocityCNN(n, seed=3)
    w = K.init_weights(5, scale_last=1.0)
    net.set_weights(w)
    assert np.array_equal(net.weights(), w)
    du, dv = net.forward(in3, xn, yn)
    du0, dv0, _ = K.forward(w, in3, xn, yn, n)
    scale = max(np.abs(du0).max(), np.abs(dv0).max())
    assert np.abs(du - du0).max() < 2e-2 * scale
    assert np.abs(dv - dv0).max() < 2e-2 * scale
    g, loss = net.gradients(in3, xn, yn, tu, tv)
    loss0, g0 = K.loss_and_grad(w, in3, xn, yn, tu, tv, n)
    assert loss == pytest.approx(loss0, rel=2e-2)
    sl, _ = K.layer_slices()
    for i in range(4):
        for s in sl[i]:
            a, b = g[s].astype(np.float64), g0[s]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            assert rel < 5e-2 and cos > 0.998, (i, s, rel, cos)
    net.close()


def _flow_sample(n, m, seed):
    """Coarse spectral flow on an n-grid: inputs (u, v, q) normalised, and at
    m random nodes the difference between the exact (spectral) velocity and
    velint's bilinear one, which the network learns."""
    from hydra_amd.ic import random_pv
    q = random_pv(n, seed=seed, k0=3.0, rms=4 * np.pi)   # [ix, iy]
    k = np.fft.fftfreq(n, d=1.0 / n)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    k2 = KX ** 2 + KY ** 2
    qh = np.fft.fft2(q)
    ph = np.where(k2 > 0, -qh / np.where(k2 > 0, k2, 1), 0)
    uh, vh = -1j * KY * ph, 1j * KX * ph
    u, v = np.fft.ifft2(uh).real, np.fft.ifft2(vh).real
    rng = np.random.default_rng(seed + 1000)
    xn = rng.uniform(-np.pi, np.pi, m)
    yn = rng.uniform(-np.pi, np.pi, m)
    ph_n = np.exp(1j * (np.outer(xn + np.pi, k)[:, :, None] + np.outer(yn + np.pi, k)[:, None, :]))
    ue = (ph_n * uh[None]).sum(axis=(1, 2)).real / n ** 2
    ve = (ph_n * vh[None]).sum(axis=(1, 2)).real / n ** 2
    idx, wt = K._bilinear(xn, yn, n, L)
    ub = sum(wt[j] * u[idx[j][0], idx[j][1]] for j in range(4))
    vb = sum(wt[j] * v[idx[j][0], idx[j][1]] for j in range(4))
    su, sq = np.sqrt((u ** 2 + v ** 2).mean()), np.sqrt((q ** 2).mean())
    in3 = np.stack([u / su, v / su, q / sq], axis=-1).reshape(n * n, 3)
    return in3, xn, yn, ((ue - ub) / su).astype(np.float32), ((ve - vb) / su).astype(np.float32)


def test_training_reduces_velint_error():
    """Trained on the method's own pairs (coarse field versus the exact
    spectral velocity at the nodes), the network removes a good part of
    bilinear velint's error on flows it has not seen."""
    from hydra_amd import VelocityCNN
    n, m = 32, 4000
    train = [_flow_sample(n, m, s) for s in range(12)]
    test = [_flow_sample(n, m, 100 + s) for s in range(3)]
    net = VelocityCNN(n, seed=7)
    for it in range(600):
        in3, xn, yn, tu, tv = train[it % len(train)]
        net.train_step(in3, xn, yn, tu, tv, lr=2e-3 if it < 400 else 5e-4)
    before = after = 0.0
    for in3, xn, yn, tu, tv in test:
        du, dv = net.forward(in3, xn, yn)
        before += np.sum(tu.astype(np.float64) ** 2 + tv.astype(np.float64) ** 2)
        after += np.sum((tu - du).astype(np.float64) ** 2 + (tv - dv).astype(np.float64) ** 2)
    ratio = np.sqrt(after / before)
    print(f"held-out velint error after correction / before: {ratio:.3f}")
    assert ratio < 0.7, ratio
    net.close()
