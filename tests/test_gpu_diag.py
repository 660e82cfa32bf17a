"""The diagnostics on the GPU, bit for bit against the oracle (and the
reference build's golden vectors): spec1d (spectral.f90:370-424), l1norm /
l2norm (generic.f90:30-64), energy (evolution.f90:612-644)."""
import os

import numpy as np
import pytest

from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_ng32.npz")


@pytest.mark.parametrize("ng", [32, 64, 256])
def test_spec1d_and_norms_bitwise(ng):
    from hydra_amd import Hydra
    from oracle.hyo import Oracle
    h = Hydra(ng)
    o = Oracle(ng)
    rng = np.random.default_rng(ng)
    for trial in range(2):
        ss = rng.standard_normal((ng, ng)) * 10.0 ** rng.integers(-3, 4)
        for iopt in (0, 1):
            a, sa = h.spec1d(ss, iopt)
            b, sb = o.spec1d(ss, iopt)
            assert bits_equal(a, b), (ng, trial, iopt)
            assert bits_equal(sa, sb), (ng, trial, iopt)
        assert h.norms(ss) == o.norms(ss), (ng, trial)
    h.close()


def test_diag_golden():
    from hydra_amd import Hydra
    g = np.load(GOLDEN)
    h = Hydra(32)
    kmax = int(g["spec_kmax"])
    for iopt in (0, 1):
        spec, ss = h.spec1d(g["fft_s"], iopt)
        assert bits_equal(spec[:kmax + 1], g[f"spec{iopt}"][:kmax + 1]), iopt
        assert bits_equal(ss, g[f"spec{iopt}_ss"]), iopt
    assert h.norms(g["fft_f"]) == (float(g["norm_l1"]), float(g["norm_l2"]))
    h.close()


@pytest.mark.parametrize("ng,kd", [(32, 0.0), (64, 0.0), (64, 2.0)])
def test_energy_bitwise(ng, kd):
    """energy of the device-resident state after each of three steps."""
    from hydra_amd import Hydra
    from hydra_amd.ic import random_pv
    from oracle.hyo import Oracle
    qq = random_pv(ng, seed=777)
    o = Oracle(ng, kd=kd)
    o.init_from_field(qq)
    o.recontour(o.d("qr").copy())
    c = o.contours()
    o.evolution_init()
    h = Hydra(ng, kd=kd)
    h.set_field(qq)
    h.set_contours(c)
    h.evolution_init()
    for step in range(3):
        o.advance()
        h.advance()
        ek, ep = h.energy()
        eko, epo = o.energy()
        assert ek == eko and ep == epo, (step, ek, eko, ep, epo)
        assert ek > 0.0 and (ep > 0.0) == (kd > 0.0)
    h.close()


def test_energy_needs_plain_precision():
    from hydra_amd import Hydra, HydraError
    from hydra_amd.ic import random_pv
    h = Hydra(32, precision=1)
    h.set_field(random_pv(32, seed=1))
    with pytest.raises(HydraError):
        h.energy()
    h.close()
