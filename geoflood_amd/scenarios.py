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


def gaussian_sum_dem(n: int, length: float = 2000.0, valley_slope: float = 0.002, ngauss: int = 64,
                     amp=(0.0, 5.0), width=(20.0, 200.0), seed: int = 0) -> np.ndarray:
    """Overland-flood terrain on an n x n grid over [0, length]^2: a valley
    floor descending in +x with ``valley_slope`` plus ``ngauss`` Gaussian
    hills (amplitude U[amp], width U[width], centres uniform in the domain,
    drawn from ``seed`` in that order)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp[0], amp[1], ngauss)
    w = rng.uniform(width[0], width[1], ngauss)
    cx = rng.uniform(0.0, length, ngauss)
    cy = rng.uniform(0.0, length, ngauss)
    dx = length / n
    xc = (np.arange(n) + 0.5) * dx
    b = np.broadcast_to(valley_slope * (length - xc)[None, :], (n, n)).copy()
    tmp = np.empty((n, n))  # a[k] * outer(gy, gx), formed in place
    for k in range(ngauss):
        gx = np.exp(-((xc - cx[k]) ** 2) / (2.0 * w[k] ** 2))
        gy = np.exp(-((xc - cy[k]) ** 2) / (2.0 * w[k] ** 2))
        np.outer(gy, gx, out=tmp)
        np.multiply(tmp, a[k], out=tmp)
        b += tmp
    return b


def sine_hydrograph(peak_q: float, strip: float = 100.0, period: float = 3600.0, rows: int = 61):
    """Tabulated Q(t) = peak_q sin(pi t / period) m^3/s over [0, period],
    as discharge per unit width of a ``strip``-metre channel."""
    t = np.linspace(0.0, period, rows)
    return t, peak_q * np.sin(np.pi * t / period) / strip


def overland_flood(n: int, length: float = 2000.0, peak_q: float = 200.0, strip: float = 100.0,
                   seed: int = 0):
    """Initially dry overland flood: Gaussian-sum DEM (``seed``), hydrograph
    inflow through a ``strip``-metre channel centred on the west edge
    (bc2 sides: west inflow, east extrapolation, south and north walls).
    Returns a dict with the global fields h, hu, hv, b, the cell size dx,
    mthbc and the hydrograph (time, hu per unit width, b, x0, y0)."""
    b = gaussian_sum_dem(n, length, seed=seed)
    z = np.zeros((n, n))
    t, hu = sine_hydrograph(peak_q, strip)
    return dict(h=z.copy(), hu=z.copy(), hv=z.copy(), b=b, dx=length / n, mthbc=(0, 1, 3, 3),
                hydro=dict(time=t, hu=hu, b=strip, x0=0.0, y0=0.5 * length))


def reservoir_state(run, x1: float = 250.0, y0: float = 1750.0, h0: float = 30.0):
    """Initial state of a moving dam-break front for a resident-AMR run
    (geoflood_amd.amr.AmrRun): depth h0 over the reservoir x < x1, y >= y0
    (cell centres, ghost cells included), dry elsewhere, at rest.  Returns
    the state_fn that AmrRun.start / OracleAmr.start take."""
    def state(L, i, j, aux):
        X, Y = run.cell_centres(L, i, j)
        q = np.zeros(aux.shape[:3] + (3,))
        q[..., 0] = np.where((X < x1) & (Y >= y0), h0, 0.0)
        return q
    return state
