"""Comparison of device and oracle results (test helper).

assert_same: the device and the oracle evaluate every FP64 value with the
same operations in the same order (no FMA contraction on either side, the
two powers correctly rounded on both when the oracle's pow mode is set), so
device runs are held to identical values: every dt, CFL and cell.

state_diff / assert_state_close: the per-cell measure used where two sides
may round differently -- the reference build's libm pow against correctly
rounded powers.  Such differences are carried by the flow, so long runs are
compared per cell:

* the wet/dry mask (h > dry_tolerance) must be identical;
* wet cells: |dh| / max(|h|) <= rtol, and for each momentum component
  |d(hu)| / max(|hu|, h sqrt(g h)) <= rtol -- relative to the larger of the
  momentum and the gravity-wave momentum scale of the cell, so a momentum at
  rounding level in still water is not compared against itself;
* dry cells (the reference discards their momentum at the next b4step2 and
  skips their Riemann problems): |dh|, |d(hu)|, |d(hv)| <= rtol * drytol.
"""
from __future__ import annotations

import numpy as np


def state_diff(got, ref, drytol=1e-3, grav=9.81):
    """Returns (max wet relative difference, max dry absolute difference /
    drytol, mask_equal) for states of shape (3, ...)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    wet_g = got[0] > drytol
    wet_r = ref[0] > drytol
    mask_equal = bool(np.array_equal(wet_g, wet_r))
    wet = wet_g & wet_r
    dry = ~wet_g & ~wet_r
    rel = 0.0
    if wet.any():
        h = np.maximum(np.abs(got[0][wet]), np.abs(ref[0][wet]))
        rel = float(np.max(np.abs(got[0][wet] - ref[0][wet]) / h))
        cscale = h * np.sqrt(grav * h)
        for m in (1, 2):
            a, b = got[m][wet], ref[m][wet]
            den = np.maximum(np.maximum(np.abs(a), np.abs(b)), cscale)
            rel = max(rel, float(np.max(np.abs(a - b) / den)))
    dry_abs = 0.0
    if dry.any():
        dry_abs = float(max(np.max(np.abs(got[m][dry] - ref[m][dry])) for m in range(3))) / drytol
    return rel, dry_abs, mask_equal


def assert_same(got, ref, what=""):
    """Identical values, element by element (NaN equal to NaN, +0 to -0)."""
    g = np.asarray(got, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    assert g.shape == r.shape, f"{what} shapes differ: {g.shape} vs {r.shape}"
    if not np.array_equal(g, r, equal_nan=True):
        bad = ~((g == r) | (np.isnan(g) & np.isnan(r)))
        i = np.unravel_index(np.argmax(bad), g.shape)
        raise AssertionError(f"{what} {int(bad.sum())} of {g.size} values differ; first at {tuple(map(int, i))}: "
                             f"{g[i]!r} vs {r[i]!r}")


def assert_state_close(got, ref, rtol, drytol=1e-3, grav=9.81):
    rel, dry_abs, mask_equal = state_diff(got, ref, drytol, grav)
    assert mask_equal, "wet/dry masks differ"
    assert rel <= rtol, f"wet-cell relative difference {rel:.3e} > {rtol:.1e}"
    assert dry_abs <= rtol, f"dry-cell difference {dry_abs:.3e} * drytol > {rtol:.1e} * drytol"
