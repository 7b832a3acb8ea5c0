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
