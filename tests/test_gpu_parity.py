"""Bit-for-bit parity of the HIP path against the oracle (run on the GPU).

Every comparison here is exact (integer or float64 bit patterns): the plain
FP64 path follows the reference's operation order, so any difference is a
defect."""
import numpy as np
import pytest

from tests.helpers import bits_equal, oracle_state

pytestmark = pytest.mark.gpu

NGS = [32, 64]


@pytest.fixture(scope="module")
def hyd_cache():
    from hydra_amd import Hydra
    cache = {}

    def get(ng):
        if ng not in cache:
            cache[ng] = Hydra(ng)
        return cache[ng]
    return get


@pytest.mark.parametrize("ng", NGS)
def test_tables_match_oracle(ng, hyd_cache):
    from oracle.hyo import Oracle
    o = Oracle(ng)
    h = hyd_cache(ng)
    for name, n in [("green", None), ("flo", None), ("fhi", None), ("qdiss", None),
                    ("hrkx", ng), ("hrky", ng), ("rkx", ng), ("xtrig", 2 * ng)]:
        a = h.table(name)
        b = o.d(name, n).ravel() if n else o.d(name).ravel()
        assert bits_equal(a, b), name


@pytest.mark.parametrize("ng", NGS + [256])
def test_fft_bitwise(ng, hyd_cache):
    from oracle.hyo import Oracle
    o = Oracle(ng)
    h = hyd_cache(ng)
    rng = np.random.default_rng(ng)
    f = rng.standard_normal((ng, ng))
    assert bits_equal(h.ptospc(f), o.ptospc(f))
    s = rng.standard_normal((ng, ng))
    assert bits_equal(h.spctop(s), o.spctop(s))


@pytest.mark.parametrize("ng", NGS)
def test_invert_and_gradient_bitwise(ng, hyd_cache):
    from oracle.hyo import Oracle
    o = Oracle(ng)
    h = hyd_cache(ng)
    rng = np.random.default_rng(7 + ng)
    s = rng.standard_normal((ng, ng))
    for a, b in zip(h.main_invert(s), o.main_invert(s)):
        assert bits_equal(a, b)
    for a, b in zip(h.gradient(s), o.gradient(s)):
        assert bits_equal(a, b)


@pytest.mark.parametrize("ng", NGS)
def test_velint_bitwise(ng, hyd_cache):
    o, qq, c = oracle_state(ng)
    h = hyd_cache(ng)
    rng = np.random.default_rng(3)
    uu, vv = rng.standard_normal((2, ng, ng))
    a = h.velint(uu, vv, c["x"], c["y"])
    b = o.velint(uu, vv)
    assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])


@pytest.mark.parametrize("ng", NGS)
def test_con2grid_bitwise(ng, hyd_cache):
    o, qq, c = oracle_state(ng)
    h = hyd_cache(ng)
    qc_h, cnt_h = h.con2grid(c["x"], c["y"], c["nextq"], c["qjump"], with_counts=True)
    qc_o, cnt_o = o.con2grid(with_counts=True)
    assert np.array_equal(cnt_h.ravel(), cnt_o.ravel())
    assert bits_equal(qc_h, qc_o)


def test_sum_sequential_bitwise(hyd_cache):
    h = hyd_cache(64)
    rng = np.random.default_rng(11)
    cases = [rng.standard_normal(4096 * 4) * 10 + 3, rng.standard_normal(4096 * 4),
             np.round(rng.standard_normal(12345) * 64) / 64 + 2.0 ** -30,
             rng.integers(-5, 6, 9999) * 0.1, rng.standard_normal(777) ** 2,
             np.full(5000, 0.1), np.array([1.0, 1e-16, 1e-16, 1e-16, -1.0]),
             rng.standard_normal(5000) * 1e-300,
             np.array([1e16 if i % 2 == 0 else -1e16 + 1.5 for i in range(3001)]),
             np.r_[rng.standard_normal(8000) * 1e8, -rng.standard_normal(8000) * 1e8],
             np.r_[[2.0 ** 40], np.tile([0.5 ** 12 * 3, -0.5 ** 12], 6000)],
             np.r_[[-0.0, 0.0, -0.0], rng.standard_normal(600)], np.array([-0.0]),
             np.array([0.0, -0.0, -0.0]), np.zeros(0)]
    for a in cases:
        ref = 0.0
        for x in a:
            ref = ref + float(x)
        got = h.sum_sequential(a)
        assert np.float64(got).view(np.int64) == np.float64(ref).view(np.int64), (len(a), got, ref)


def test_sum_sequential_long_bitwise():
    """Long sums (up to the 4 ng^2 work space at ng=256) of zero-mean,
    tie-rich data shaped like the con2grid field: many binade changes,
    exact ties, chunk counts that are not multiples of 64."""
    from hydra_amd import Hydra
    h = Hydra(256)
    rng = np.random.default_rng(5)
    ng = 256
    x = np.linspace(-np.pi, np.pi, ng, endpoint=False)
    f = np.add.outer(np.sin(3 * x), np.cos(2 * x)) * 5 + rng.standard_normal((ng, ng)) * 0.1
    qj = 0.37
    lev = np.round(f / qj) * qj  # level steps: many exact ties
    cases = [lev.ravel(), (lev - lev.mean()).ravel(), f.ravel(),
             np.tile(lev.ravel(), 4)[: 4 * ng * ng - 77], rng.standard_normal(4 * ng * ng)]
    for a in cases:
        ref = 0.0
        for v in a.tolist():
            ref = ref + v
        got = h.sum_sequential(a)
        assert np.float64(got).view(np.int64) == np.float64(ref).view(np.int64), (len(a), got, ref)
    h.close()


def test_sum_sequential_binade_walks_bitwise():
    """Running sums that wander across powers of two inside the windows of
    1024 values (crossings up and down, several binades, sign changes, exact
    ties), the case the per-window class tables resolve; bit for bit."""
    from hydra_amd import Hydra
    h = Hydra(256)
    rng = np.random.default_rng(21)
    n = 4 * 256 * 256
    steps = rng.standard_normal(n)
    cases = []
    for level in (1024.0, 1.0, 2.0 ** -20, 3.0e5):
        a = steps * level * 0.02
        a[0] = level * 0.99  # start just below the power of two
        cases.append(a)
    walk = np.cumsum(rng.standard_normal(n) * 0.5)
    cases.append(np.diff(np.r_[0.0, np.sin(walk) * 700.0 + 1024.0 * np.cos(walk / 50.0)]))
    ties = np.round(rng.standard_normal(n) * 8.0) * 0.25
    ties[0] = 511.875
    cases.append(ties)
    sign = rng.standard_normal(n) * 3.0
    sign[::4096] += np.where(np.arange(len(sign[::4096])) % 2 == 0, 5e4, -5e4)
    cases.append(sign)
    for a in cases:
        ref = 0.0
        for v in a.tolist():
            ref = ref + v
        got = h.sum_sequential(a)
        assert np.float64(got).view(np.int64) == np.float64(ref).view(np.int64), (got, ref)
    h.close()


@pytest.mark.parametrize("ng", [32, 64])
def test_advance_bitwise(ng):
    """Several RK4 steps of the device-resident state, every field and node
    position bit for bit."""
    from hydra_amd import Hydra
    from oracle.hyo import Oracle
    from hydra_amd.ic import random_pv
    qq = random_pv(ng, seed=12345)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    c = o.contours()
    o.evolution_init()
    h = Hydra(ng)
    h.set_field(qq)
    assert h.scalar("qjump") == o.getd("qjump")
    h.set_contours(c)
    h.evolution_init()
    for name in ("qc", "qd", "qs"):
        assert bits_equal(h.field(name), o.d(name)), name
    for step in range(4):
        o.advance()
        h.advance()
        for name in ("qs", "qd", "qq", "uu", "vv", "pp"):
            assert bits_equal(h.field(name), o.d(name)), (step, name)
        cg = h.contours()
        n = o.get("nptq")
        assert bits_equal(cg["x"], o.d("xq", n)) and bits_equal(cg["y"], o.d("yq", n)), step
        assert h.scalar("t") == o.getd("t") and h.scalar("twist") == o.getd("twist")
    h.close()


@pytest.mark.parametrize("ng", NGS)
def test_renode_bitwise(ng, hyd_cache):
    o, qq, c = oracle_state(ng)
    h = hyd_cache(ng)
    rng = np.random.default_rng(ng)
    for j in list(range(4)) + list(rng.integers(0, len(c["i1q"]), 6)):
        i1, i2 = c["i1q"][j], c["i2q"][j]
        # perturb so that the contour is not already renoded
        xd = c["x"][i1 - 1:i2] + 1e-3 * np.sin(7 * c["y"][i1 - 1:i2])
        yd = c["y"][i1 - 1:i2]
        a = h.renode(xd, yd)
        b = o.renode(xd, yd)
        assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1]), j
    # a triangle: several corners
    verts = np.array([[-0.6, -0.4], [0.7, -0.3], [0.0, 0.6]])
    pts = np.concatenate([verts[a] + np.arange(40)[:, None] / 40.0 * (verts[(a + 1) % 3] - verts[a])
                          for a in range(3)])
    a = h.renode(pts[:, 0], pts[:, 1])
    b = o.renode(pts[:, 0], pts[:, 1])
    assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])


@pytest.mark.parametrize("ng", NGS)
def test_recontour_from_field_bitwise(ng):
    """congen from the initial field: contour count, node count, topology and
    positions identical to the oracle."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    qq = random_pv(ng, seed=12345)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    c = o.contours()
    h = Hydra(ng)
    h.set_field(qq)
    h.recontour()
    g = h.contours()
    assert len(g["x"]) == len(c["x"]) and len(g["i1q"]) == len(c["i1q"])
    for k in ("nextq", "i1q", "i2q", "npq", "indq"):
        assert np.array_equal(g[k], c[k]), k
    assert bits_equal(g["x"], c["x"]) and bits_equal(g["y"], c["y"])
    h.close()


def test_recontour_cycle_bitwise():
    """prepare + recontour with existing contours (the con2ugrid branch)
    after a few steps, then evolution_init: identical to the oracle."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    ng = 32
    qq = random_pv(ng, seed=7)
    o = Oracle(ng)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    c = o.contours()
    o.evolution_init()
    h = Hydra(ng)
    h.set_field(qq)
    h.set_contours(c)
    h.evolution_init()
    for _ in range(3):
        o.advance()
        h.advance()
    o.L.hyo_prepare(o.m)
    h.prepare()
    assert bits_equal(h.field("qr"), o.d("qr"))
    o.recontour(o.d("qr").copy())
    h.recontour()
    c2, g2 = o.contours(), h.contours()
    for k in ("nextq", "i1q", "i2q", "npq", "indq"):
        assert np.array_equal(g2[k], c2[k]), k
    assert bits_equal(g2["x"], c2["x"]) and bits_equal(g2["y"], c2["y"])
    o.evolution_init()
    h.evolution_init()
    for name in ("qc", "qd", "qs"):
        assert bits_equal(h.field(name), o.d(name)), name
    h.close()


def _hairpins(ng, nest=4):
    """Closed thin loops whose crossings of one fine x-gridline fall in one
    fine cell with both signs, `nest` of each: order-dependent cells
    (summed in node order from the buckets), inside the grid and on the
    bottom row (loops wrapping the periodic y boundary), plus one large
    circle so that the grid has ordinary cells too."""
    h = 2 * np.pi / (4 * ng)  # fine cell
    xs, ys, npq = [], [], []

    def loop(pts):
        xs.extend(p[0] for p in pts)
        ys.extend(p[1] for p in pts)
        npq.append(len(pts))
    # interior: rectangles across a few gridlines, both long sides in one fine row
    for cx, cy in ((0.3, -1.1), (-2.0, 0.7), (1.7, 2.2)):
        x0 = -np.pi + h * (round((cx + np.pi) / h) + 0.5)
        y0 = -np.pi + h * (round((cy + np.pi) / h) + 0.2)
        for k in range(nest):
            ya, yb = y0 + 0.05 * h * k, y0 + 0.55 * h - 0.05 * h * k
            loop([(x0 - 0.1 * k * h, ya), (x0 + 3 * h + 0.1 * k * h, ya),
                  (x0 + 3 * h + 0.1 * k * h, yb), (x0 - 0.1 * k * h, yb)])
    # bottom row: loops through y = +-pi, both vertical sides in one fine column
    for cx in (-0.9, 2.6):
        x0 = -np.pi + h * (round((cx + np.pi) / h) + 0.2)
        for k in range(nest):
            xa, xb = x0 + 0.05 * h * k, x0 + 0.55 * h - 0.05 * h * k
            d = 3 * h + 0.1 * h * k
            loop([(xa, np.pi - d), (xa, -np.pi + d), (xb, -np.pi + d), (xb, np.pi - d)])
    # one circle
    t = np.linspace(0.0, 2 * np.pi, 400, endpoint=False)
    loop(list(zip(1.3 * np.cos(t) - 0.2, 1.1 * np.sin(t) + 0.1)))
    n = len(xs)
    i1 = np.cumsum([0] + npq[:-1]) + 1
    nxt = np.empty(n, dtype=np.int32)
    for a, m in zip(i1, npq):
        idx = np.arange(a, a + m)
        nxt[idx - 1] = np.roll(idx, -1)
    nq = len(npq)
    return dict(x=np.array(xs), y=np.array(ys), nextq=nxt, i1q=i1.astype(np.int32),
                i2q=(i1 + np.array(npq) - 1).astype(np.int32), npq=np.array(npq, np.int32),
                indq=np.ones(nq, np.int32), qjump=0.1234567891234)


def _order_dependent(c, ng):
    """Fine cells and bottom-row counters with both signs and at least
    three of one (crossings located as contours.f90:206-269 does)."""
    import math
    nxf, ell, hl = 4 * ng, 2 * np.pi, np.pi
    glxfi = nxf / ell
    xg = -hl + (ell / nxf) * np.arange(nxf)
    cells, bottom = {}, {}
    x, y, nx = c["x"], c["y"], c["nextq"]
    for i in range(len(x)):
        xi, yi = x[i], y[i]
        xa, ya = x[nx[i] - 1], y[nx[i] - 1]
        ixci, ixca = 1 + int(glxfi * (xi + hl)), 1 + int(glxfi * (xa + hl))
        dx = (xa - xi) - ell * int((xa - xi) / hl)
        yy = ya - yi
        dy = yy - ell * int(yy / hl)
        if abs(yy) > hl:
            cc = math.copysign(1.0, dy)
            p = -(yi - hl * cc) / (dy + 1e-12)
            ix = 1 + int(glxfi * math.fmod(ell + hl + xi + p * dx, ell))
            bottom.setdefault(ix, []).append(-1 if cc > 0 else 1)
        ixdif = ixca - ixci
        nxc = ixdif - nxf * int(2 * ixdif / nxf)
        jump = 1 if nxc > 0 else -1
        for ncr in range(0, nxc, jump):
            ix = 1 + (ixci + (0 if jump == 1 else -1) + nxf + ncr) % nxf
            x2 = xg[ix - 1] - xi
            y2 = yi + (x2 - ell * int(x2 / hl)) / dx * dy
            iy = 2 + int(glxfi * (y2 - ell * int(y2 / hl) + hl))
            cells.setdefault((ix, iy), []).append(1 if dx > 0 else -1)

    def flagged(d):
        return sum(1 for v in d.values() if v.count(1) and v.count(-1)
                   and max(v.count(1), v.count(-1)) >= 3)
    return flagged(cells), flagged(bottom)


@pytest.mark.parametrize("ng", NGS)
def test_con2grid_order_dependent_cells_bitwise(ng, hyd_cache):
    """Fine cells and bottom-row counters holding both signs with several of
    each (summed in node order) give the oracle's grid, through the
    diagnostic sweep (with counts) and through the step's fused sweep."""
    from oracle.hyo import Oracle
    c = _hairpins(ng)
    o = Oracle(ng)
    o.set_contours(c)
    qc_o, cnt_o = o.con2grid(with_counts=True)
    cells, bottom = _order_dependent(c, ng)
    assert cells >= 6 and bottom >= 2  # the case is exercised
    h = hyd_cache(ng)
    for rep in range(2):  # the second sweep starts from the counters the first cleared
        qc_h, cnt_h = h.con2grid(c["x"], c["y"], c["nextq"], c["qjump"], with_counts=True)
        assert np.array_equal(cnt_h.ravel(), cnt_o.ravel())
        assert bits_equal(qc_h, qc_o)
        assert bits_equal(h.con2grid(c["x"], c["y"], c["nextq"], c["qjump"]), o.con2grid())
