"""Write tests/golden/amr_vectors.npz from the reference regrid averaging and
time interpolation (fclaw2d_source/fc2d_geoclaw_average_fort.f:
fc2d_geoclaw_fort_average2coarse; fc2d_geoclaw_timeinterp_fort.f) compiled by
oracle/Makefile (target ``ref``, library _ref/libgfref_amr.so):

* average: 8x8 coarse and fine patches (values on a 1/64 grid, wet, dry and
  exactly-dry_tolerance cells) and the coarse patch after averaging the fine
  patch into quadrant igrid = 0..3;
* timeinterp: two 8x8 states and the interpolated border layers.

    make -C oracle ref && python tests/golden/make_golden_amr.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefAmrLib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "amr_vectors.npz")
MX = 8


def q_random(rng, wet=0.7):
    sh = (MX + 4, MX + 4)
    q = np.round(rng.normal(0, 1, sh + (3,)) * 64) / 64
    q[..., 0] = np.abs(q[..., 0]) * (rng.random(sh) < wet)
    q[..., 0][rng.random(sh) < 0.05] = 1e-3
    return q


def main():
    ref = RefAmrLib()
    ref.set_dry_tolerance(1e-3)
    rng = np.random.default_rng(31)
    n = 48
    qc = np.stack([q_random(rng, 0.8) for _ in range(n)])
    qf = np.stack([q_random(rng, 0.6 if k % 5 else 0.0) for k in range(n)])
    ac = np.round(rng.normal(0, 2, (n, MX + 4, MX + 4, 1)) * 64) / 64
    af = np.round(rng.normal(0, 2, (n, MX + 4, MX + 4, 1)) * 64) / 64
    igrid = np.arange(n, dtype=np.int32) % 4
    out = qc.copy()
    for k in range(n):
        ref.average2coarse(out[k], qf[k].copy(), ac[k], af[k], int(igrid[k]))
    m = 16
    tc = np.round(rng.normal(0, 1, (m, MX + 4, MX + 4, 3)) * 64) / 64
    tl = np.round(rng.normal(0, 1, (m, MX + 4, MX + 4, 3)) * 64) / 64
    alpha = rng.uniform(0, 1, m)
    ti = np.zeros_like(tc)
    psize = 8 * (MX - 2) * 3
    for k in range(m):
        assert ref.timeinterp(tc[k], tl[k], ti[k], float(alpha[k]), psize) == 0
    np.savez_compressed(OUT, avg_qc=qc, avg_qf=qf, avg_auxc=ac, avg_auxf=af, avg_igrid=igrid, avg_out=out,
                        ti_curr=tc, ti_last=tl, ti_alpha=alpha, ti_out=ti, ti_psize=np.array([psize]))
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
