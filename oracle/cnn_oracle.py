"""CPU oracle of the learned velocity correction (test infrastructure only:
only tests/ may use it).  No reference counterpart: the reference has no
learned part.

The network: four 3x3 convolutions on an n x n doubly periodic grid
(pixel p = ix * n + iy), channels 3 (+1 zero pad) -> 32 -> 32 -> 32 -> 2,
GELU after the first three; the 2-channel output (du, dv) is sampled
bilinearly at node positions (grid point ix at xmin + ix * ell / n).  Loss =
mean over nodes of |(du, dv) - target|^2.

Weights, flattened as the library stores them: for each layer W[co][ci][tap]
(tap = 3 * (dx + 1) + (dy + 1), dx along ix, dy along iy) then bias[co].

bf=True rounds where the GPU kernels round to BF16 (inputs, weights,
activations, the pre-activation gradients passed between layers), in
float64 otherwise; bf=False is the exact network, whose gradients are
pinned by central finite differences in the tests.
"""
import numpy as np

CI = [4, 32, 32, 32]
CO = [32, 32, 32, 2]


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layer_slices():
    out, off = [], 0
    for ci, co in zip(CI, CO):
        nw = co * ci * 9
        out.append((slice(off, off + nw), slice(off + nw, off + nw + co)))
        off += nw + co
    return out, off


def num_weights():
    return layer_slices()[1]


def _shift(x, dx, dy):
    """x[(ix + dx) % n, (iy + dy) % n] for x of shape (n, n, C)."""
    return np.roll(np.roll(x, -dx, axis=0), -dy, axis=1)


def _conv(x, W, b):
    """x (n, n, ci), W (co, ci, 9), b (co) -> (n, n, co)."""
    out = np.zeros(x.shape[:2] + (W.shape[0],))
    for tap in range(9):
        dx, dy = tap // 3 - 1, tap % 3 - 1
        out += _shift(x, dx, dy) @ W[:, :, tap].T
    return out + b


def _conv_t(dz, W):
    """Transpose of _conv w.r.t. its input: dx[p] = sum_tap W^T dz[p - off]."""
    out = np.zeros(dz.shape[:2] + (W.shape[1],))
    for tap in range(9):
        dx, dy = tap // 3 - 1, tap % 3 - 1
        out += _shift(dz, -dx, -dy) @ W[:, :, tap]
    return out


def _conv_w(dz, x):
    """Weight gradient: dW[co][ci][tap] = sum_p dz[p][co] x[p + off][ci]."""
    co, ci = dz.shape[2], x.shape[2]
    dW = np.zeros((co, ci, 9))
    dzf = dz.reshape(-1, co)
    for tap in range(9):
        dx, dy = tap // 3 - 1, tap % 3 - 1
        dW[:, :, tap] = dzf.T @ _shift(x, dx, dy).reshape(-1, ci)
    return dW, dzf.sum(axis=0)


def _gelu(z):
    from scipy.special import erf
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def _gelu_grad(z):
    from scipy.special import erf
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def _bilinear(xn, yn, n, ell):
    gli = n / ell
    xx = gli * (np.asarray(xn) + 0.5 * ell)
    yy = gli * (np.asarray(yn) + 0.5 * ell)
    ix0 = np.floor(xx).astype(np.int64)
    iy0 = np.floor(yy).astype(np.int64)
    px, py = xx - ix0, yy - iy0
    ix0 %= n
    iy0 %= n
    ix1, iy1 = (ix0 + 1) % n, (iy0 + 1) % n
    idx = [(ix0, iy0), (ix1, iy0), (ix0, iy1), (ix1, iy1)]
    w = [(1 - px) * (1 - py), px * (1 - py), (1 - px) * py, px * py]
    return idx, w


def forward(w, in3, xn, yn, n, ell=2 * np.pi, bf=True):
    """Corrections (du, dv) at the nodes and the intermediate state."""
    R = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    sl, _ = layer_slices()
    x = np.zeros((n, n, 4))
    x[:, :, :3] = R(np.asarray(in3, np.float32).reshape(n, n, 3))
    xs, zs, Ws = [x], [], []
    for i in range(4):
        W = R(np.asarray(w[sl[i][0]], np.float64)).reshape(CO[i], CI[i], 9)
        b = np.asarray(w[sl[i][1]], np.float64)
        z = _conv(xs[-1], W, b)
        Ws.append(W)
        zs.append(z)
        if i < 3:
            xs.append(R(_gelu(z).astype(np.float32)) if bf else _gelu(z))
    idx, wt = _bilinear(xn, yn, n, ell)
    out = zs[3]
    du = sum(wt[j] * out[idx[j][0], idx[j][1], 0] for j in range(4))
    dv = sum(wt[j] * out[idx[j][0], idx[j][1], 1] for j in range(4))
    return du, dv, (xs, zs, Ws, idx, wt)


def loss_and_grad(w, in3, xn, yn, tu, tv, n, ell=2 * np.pi, bf=True):
    R = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    du, dv, (xs, zs, Ws, idx, wt) = forward(w, in3, xn, yn, n, ell, bf)
    m = len(xn)
    eu, ev = du - np.asarray(tu, np.float64), dv - np.asarray(tv, np.float64)
    loss = float(np.sum(eu * eu + ev * ev) / m)
    gout = np.zeros((n, n, 2))
    for j in range(4):
        np.add.at(gout[:, :, 0], (idx[j][0], idx[j][1]), 2.0 / m * eu * wt[j])
        np.add.at(gout[:, :, 1], (idx[j][0], idx[j][1]), 2.0 / m * ev * wt[j])
    sl, nw = layer_slices()
    grad = np.zeros(nw)
    dz = R(gout.astype(np.float32)) if bf else gout
    for i in range(3, -1, -1):
        dW, db = _conv_w(dz, xs[i])
        grad[sl[i][0]] = dW.ravel()
        grad[sl[i][1]] = db
        if i > 0:
            dx = _conv_t(dz, Ws[i])
            g = dx * _gelu_grad(zs[i - 1])
            dz = R(g.astype(np.float32)) if bf else g
    return loss, grad


def init_weights(seed, scale_last=0.1):
    """A He-style random initialisation (used by the tests; the library
    draws its own at creation)."""
    rng = np.random.default_rng(seed)
    sl, nw = layer_slices()
    w = np.zeros(nw)
    for i in range(4):
        fan = 9 * (3 if i == 0 else CI[i])
        W = rng.normal(0.0, np.sqrt(2.0 / fan), size=(CO[i], CI[i], 9))
        if i == 0:
            W[:, 3, :] = 0.0
        if i == 3:
            W *= scale_last
        w[sl[i][0]] = W.ravel()
        w[sl[i][1]] = rng.normal(0.0, 0.05, size=CO[i])
    return w.astype(np.float32)
