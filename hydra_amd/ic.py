"""Seeded synthetic initial conditions for the contour-advection step.

The workloads are defined by a random-phase PV (vorticity) field on the
doubly periodic square [-L/2, L/2)^2 whose 1D spectrum is proportional to
k^3 exp(-2 (k/k0)^2), rescaled to a given rms.  (The reference's own
generator, ca/plane/qg/init/ranpv.f90:38-133, draws a k^5 spectrum with GNU
rand phases and rescales the maximum; the workloads here are defined by the
spectrum above, and the GPU path and the checker both start from the same
array, so parity does not depend on the generator.)

Layout: a physical field is returned as a float64 array of shape (nx, ny),
C order, i.e. the reference's qq(ny,nx) column-major with iy fastest.
"""
from __future__ import annotations

import numpy as np


def random_pv(ng: int, seed: int = 12345, k0: float = 6.0, rms: float = 4.0 * np.pi,
              ellx: float = 2.0 * np.pi) -> np.ndarray:
    """Random isotropic PV field with 1D spectrum ~ k^3 exp(-2 (k/k0)^2)."""
    rng = np.random.default_rng(seed)
    sc = 2.0 * np.pi / ellx
    kx = np.fft.fftfreq(ng, d=1.0 / ng) * sc
    ky = np.fft.rfftfreq(ng, d=1.0 / ng) * sc
    kk = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    # |q_hat|^2 ~ Q(k)/k so that the shell-summed spectrum is Q(k) ~ k^3 e^{-2(k/k0)^2}
    amp = kk * np.exp(-((kk / k0) ** 2))
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    spec = amp * np.exp(1j * phase)
    f = np.fft.irfft2(spec, s=(ng, ng))
    f *= rms / np.sqrt(np.mean(f * f))
    return np.ascontiguousarray(f, dtype=np.float64)


def circular_patch(ng: int, radius: float = 1.0, q0: float = 1.0, x0: float = 0.0,
                   y0: float = 0.0, ellx: float = 2.0 * np.pi) -> np.ndarray:
    """Smoothed top-hat patch (used by closed-form checks)."""
    x = -0.5 * ellx + ellx * np.arange(ng) / ng
    X, Y = np.meshgrid(x, x, indexing="ij")
    r = np.sqrt((X - x0) ** 2 + (Y - y0) ** 2)
    return np.ascontiguousarray(q0 * 0.5 * (1.0 - np.tanh((r - radius) / (2.0 * ellx / ng))))
