"""initfft (stafft.f90:63-125) through the C ABI on the host, without a
context or device: factors and trig table bit-identical to the oracle's
(itself pinned against the reference build and the closed form), and the
lengths the reference would stop on are refused."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import bits_equal


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_initfft_matches_oracle(n):
    from hydra_amd import initfft
    from oracle.hyo import lib as olib
    f, t = initfft(n)
    fo = np.zeros(5, np.int32)
    to = np.zeros(2 * n)
    assert olib().hyo_initfft(n, fo, to) == 0
    assert np.array_equal(f, fo)
    assert bits_equal(t, to)


def test_initfft_against_reference_table():
    """The reference build's own xtrig at ng=32 (golden fixture)."""
    import os
    from hydra_amd import initfft
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_ng32.npz"))
    _, t = initfft(32)
    assert bits_equal(t, gold["tab_xtrig"])


@pytest.mark.parametrize("n", [4, 12, 48, 8192])
def test_initfft_refuses_other_lengths(n):
    from hydra_amd import HydraError, initfft
    with pytest.raises(HydraError, match="error -2"):
        initfft(n)
