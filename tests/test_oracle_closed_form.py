"""Closed-form and invariant pins of the oracle, independent of the
reference build: each one fails for a plausible restatement mistake (a
dropped term, a wrong sign or index, a transposed operand)."""
import numpy as np
import pytest

L2 = 2 * np.pi


def hc_pack(x):
    """stafft's half-complex order of numpy's rfft, scaled 1/sqrt(n):
    index k holds Re X_k (k <= n/2), index n-k holds Im X_k."""
    n = x.shape[-1]
    X = np.fft.rfft(x, axis=-1) / np.sqrt(n)
    out = np.empty(x.shape)
    out[..., : n // 2 + 1] = X.real
    out[..., n // 2 + 1:] = X.imag[..., 1: n // 2][..., ::-1]
    return out


@pytest.mark.parametrize("n", [8, 16, 32, 128, 512])
def test_forfft_closed_form_and_inverse(n):
    from oracle.hyo import lib
    L = lib()
    f = np.zeros(5, dtype=np.int32)
    trig = np.zeros(2 * n)
    assert L.hyo_initfft(n, f, trig) == 0
    m = 3
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, m))          # x(m,n): vector index fastest
    y = np.ascontiguousarray(x).copy()
    L.hyo_forfft(m, n, y, trig, f)
    assert np.allclose(y.T, hc_pack(x.T), atol=1e-13, rtol=0)
    L.hyo_revfft(m, n, y, trig, f)
    assert np.allclose(y, x, atol=1e-13, rtol=0)


def test_ptospc_is_separable_closed_form():
    from oracle.hyo import Oracle
    ng = 32
    o = Oracle(ng)
    rng = np.random.default_rng(1)
    r = rng.standard_normal((ng, ng))       # r[ix, iy] = rvar(iy, ix)
    s = o.ptospc(r)                         # s[ky, kx] = svar(kx, ky)
    # explicit: transform along ix for every iy, then along iy for every kx
    a = hc_pack(r.T)          # rows iy, columns kx
    b = hc_pack(a.T)          # rows kx, columns ky
    assert np.allclose(s, b.T, atol=1e-12)


@pytest.mark.parametrize("kx,ky", [(1, 0), (0, 2), (3, 5), (7, 1)])
def test_main_invert_single_mode(kx, ky):
    """psi = -q/(k^2+kd^2); u = -psi_y, v = psi_x, derivative factors
    k*exp(-36 (k/kmax)^36) (spectral.f90:68-69, 136-147)."""
    from oracle.hyo import Oracle
    ng = 32
    kd = 0.5
    o = Oracle(ng, kd=kd)
    x = -np.pi + L2 * np.arange(ng) / ng
    X, Y = np.meshgrid(x, x, indexing="ij")
    q = np.cos(kx * X + ky * Y)
    uu, vv, pp = o.main_invert(o.ptospc(q))
    kmax = ng / 2
    fx = np.exp(-36 * (kx / kmax) ** 36)
    fy = np.exp(-36 * (ky / kmax) ** 36)
    k2 = kx * kx + ky * ky + kd * kd
    psi_amp = -1.0 / k2
    u_ref = -psi_amp * (-ky * fy) * np.sin(kx * X + ky * Y)
    v_ref = psi_amp * (-kx * fx) * np.sin(kx * X + ky * Y)
    assert np.allclose(uu, u_ref, atol=1e-12)
    assert np.allclose(vv, v_ref, atol=1e-12)


def test_velint_exact_for_linear_fields():
    """Bilinear interpolation reproduces a linear field exactly inside the
    grid (contours.f90:148-166); a swapped weight pair would not."""
    from oracle.hyo import Oracle
    ng = 32
    o = Oracle(ng)
    x = -np.pi + L2 * np.arange(ng) / ng
    X, Y = np.meshgrid(x, x, indexing="ij")
    uu = 0.3 * X - 1.7 * Y + 0.25
    vv = -2.0 * X + 0.5 * Y - 1.0
    rng = np.random.default_rng(2)
    n = 500
    xq = rng.uniform(-np.pi + 0.01, np.pi - L2 / ng - 0.01, n)
    yq = rng.uniform(-np.pi + 0.01, np.pi - L2 / ng - 0.01, n)
    o.set_contours(dict(x=xq, y=yq, nextq=np.arange(2, n + 2, dtype=np.int32) % n + 0,
                        i1q=np.array([1], np.int32), i2q=np.array([n], np.int32),
                        npq=np.array([n], np.int32), indq=np.array([1], np.int32), qjump=1.0))
    uq, vq = o.velint(uu, vv)
    assert np.allclose(uq, 0.3 * xq - 1.7 * yq + 0.25, atol=1e-12)
    assert np.allclose(vq, -2.0 * xq + 0.5 * yq - 1.0, atol=1e-12)


def brute_counts(c, ng, qjump):
    """Independent level count on the 4ng fine grid.  For every fine line
    x = x_f the signed crossings of the contour segments lying below each grid
    row are summed, plus the signed crossings of y = ymin left of x_f.  Which
    lines a segment crosses follows the reference's convention (the lines
    after the start node's fine cell up to the end node's cell,
    contours.f90:207,220-221); where it crosses is plain geometry."""
    nf = 4 * ng
    h = L2 / nf
    xf = -np.pi + h * np.arange(nf)
    x, y, nxt = c["x"], c["y"], c["nextq"] - 1
    dx = x[nxt] - x
    dx = dx - L2 * np.trunc(dx / np.pi)
    dy = y[nxt] - y
    dy = dy - L2 * np.trunc(dy / np.pi)
    cell = (nf * (x + np.pi) / L2).astype(np.int64)      # 0-based fine cell
    dc = cell[nxt] - cell
    dc = dc - nf * np.trunc(2 * dc / nf).astype(np.int64)
    counts = np.zeros((nf, nf), dtype=np.int64)   # [ix, iy]
    rows = -np.pi + h * np.arange(nf)
    for ix in range(nf):
        # line ix is crossed when it lies in (cell_i, cell_i + dc] (dc > 0)
        # or (cell_i + dc, cell_i] (dc < 0), modulo nf
        off = (ix - cell) % nf
        hit = ((dc > 0) & (off >= 1) & (off <= dc)) | ((dc < 0) & ((off <= 0) | (off > nf + dc)) & ((ix - cell - dc) % nf != 0) & ((cell - ix) % nf <= -dc - 1))
        d = xf[ix] - x[hit]
        d = d - L2 * np.round(d / L2)
        yc = y[hit] + (d / dx[hit]) * dy[hit]
        yc = yc - L2 * np.floor((yc + np.pi) / L2)
        sg = np.sign(dx[hit])
        counts[ix] = (sg[:, None] * (rows[None, :] > yc[:, None])).sum(axis=0)
    # bottom row: crossings of y = ymin, signed by -sign(dy), left of x_f
    cross = np.abs(y[nxt] - y) > np.pi
    p = -(y[cross] - np.pi * np.sign(dy[cross])) / dy[cross]
    xc = np.mod(x[cross] + p * dx[cross] + np.pi, L2) - np.pi
    sgb = -np.sign(dy[cross])
    qbot = np.array([(sgb * (xc < xf[ix])).sum() for ix in range(nf)])
    return counts + qbot[:, None]


def test_con2grid_counts_match_brute_force():
    from tests.helpers import oracle_state
    ng = 32
    o, qq, c = oracle_state(ng)
    qc, cnt = o.con2grid(with_counts=True)
    ref = brute_counts(c, ng, c["qjump"])
    got = cnt.reshape(4 * ng, 4 * ng)          # [ix, iy]
    mism = np.count_nonzero(got != ref)
    assert mism == 0, mism
    assert abs(qc.sum()) < 1e-9 * np.abs(qc).sum()


def test_renode_preserves_corners_and_spacing():
    """A triangle's vertices turn by more than 90 degrees, so renode keeps
    them as nodes (contours.f90:740-745); the new spacing respects the
    density bounds dm/2 .. ell."""
    from oracle.hyo import Oracle
    o = Oracle(64)
    verts = np.array([[-0.6, -0.4], [0.7, -0.3], [0.0, 0.6]])
    pts = []
    for a in range(3):
        p0, p1 = verts[a], verts[(a + 1) % 3]
        for t in np.arange(60) / 60.0:
            pts.append(p0 + t * (p1 - p0))
    pts = np.array(pts)
    xr, yr = o.renode(pts[:, 0], pts[:, 1])
    got = set(zip(np.round(xr, 12), np.round(yr, 12)))
    for vx, vy in verts:
        assert (round(vx, 12), round(vy, 12)) in got
    seg = np.hypot(np.diff(np.append(xr, xr[0])), np.diff(np.append(yr, yr[0])))
    assert seg.min() > 0.5 * o.getd("dm") - 1e-12
    assert seg.max() < o.getd("ell") * 1.01


def test_surgery_reconnects_close_contours_only():
    """Two facing loops closer than dm reconnect into one; farther apart they
    do not (contours.f90:554-563).  The gap is placed inside a surgery box:
    the search only visits segments listed in the node's own box."""
    from oracle.hyo import Oracle
    o = Oracle(64)
    dm = o.getd("dm")

    def run(gap):
        n = 200
        t = np.linspace(-1.0, 1.0, n)
        # two thin closed loops facing each other across the gap
        xa = np.concatenate([t, t[::-1]])
        ya = np.concatenate([np.full(n, -0.1), np.full(n, -gap / 2)])
        xb = np.concatenate([t, t[::-1]]) + 0.37 * (t[1] - t[0])   # nodes not aligned
        yb = np.concatenate([np.full(n, gap / 2), np.full(n, 0.1)])
        x = np.concatenate([xa, xb]) * (1 - 1e-12)
        y = (np.concatenate([ya, yb]) + 0.05) * (1 - 1e-12)
        m = 2 * n
        nxt = np.concatenate([np.arange(2, m + 1), [1], np.arange(m + 2, 2 * m + 1), [m + 1]]).astype(np.int32)
        o.set_contours(dict(x=x, y=y, nextq=nxt, i1q=np.array([1, m + 1], np.int32),
                            i2q=np.array([m, 2 * m], np.int32), npq=np.array([m, m], np.int32),
                            indq=np.array([1, 1], np.int32), qjump=1.0))
        o.surgery()
        return o.get("nq")

    assert run(0.5 * dm) == 1
    assert run(0.9 * dm) == 1
    assert run(1.5 * dm) == 2
    assert run(3.0 * dm) == 2
