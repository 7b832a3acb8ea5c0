"""Initial states and terrains for the benchmark configurations.

All builders return global interior fields (h, hu, hv, b) of shape (ny, nx),
float64, evaluated at cell centres x = xlower + (i + 1/2) dx.
"""
from __future__ import annotations

import numpy as np


def radial_dam_break(n: int = 256, xlower: float = 0.0, xupper: float = 1.0,
                     radius: float = 0.2, h_in: float = 2.0, h_out: float = 1.0):
    """Plain-method radial dam break on [0,1]^2: h = h_in inside ``radius``
    about the centre, h_out elsewhere, u = v = 0, flat bottom b = 0."""
    dx = (xupper - xlower) / n
    xc = xlower + (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc)
    xm = 0.5 * (xlower + xupper)
    r = np.sqrt((X - xm) ** 2 + (Y - xm) ** 2)
    h = np.where(r < radius, h_in, h_out).astype(np.float64)
    z = np.zeros_like(h)
    return h, z.copy(), z.copy(), z.copy(), dx


def fractal_dem(n: int, rms: float = 3.0, spectral_slope: float = -2.0, seed: int = 0) -> np.ndarray:
    """Seeded spectral-synthesis terrain on an n x n grid: Gaussian white noise
    filtered so that the power spectrum falls as |k|^spectral_slope, zero mean,
    scaled to the given RMS (metres)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n, n))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    spec = np.fft.rfft2(noise) * k ** (spectral_slope / 2.0)
    z = np.fft.irfft2(spec, s=(n, n))
    z -= z.mean()
    z *= rms / z.std()
    return z


def uniform_dam_break(n: int = 4096, dx: float = 1.0, h0: float = 10.0, rms: float = 3.0,
                      regional_slope: float = 1.0 / 500.0, seed: int = 0):
    """Uniform large-batch dam break: fractal DEM (RMS ``rms``, spectral slope
    -2, ``seed``) plus a regional slope descending in +x; depth ``h0`` over
    the upper-left eighth of the domain (x < L/2, y >= 3L/4), dry elsewhere;
    u = v = 0.  Returns (h, hu, hv, b, dx)."""
    L = n * dx
    xc = (np.arange(n) + 0.5) * dx
    b = fractal_dem(n, rms=rms, seed=seed) - regional_slope * xc[None, :]
    h = np.zeros((n, n))
    h[int(0.75 * n):, : n // 2] = h0
    return h, np.zeros_like(h), np.zeros_like(h), b, dx
