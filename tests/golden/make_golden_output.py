"""Write tests/golden/output_vectors.npz from the reference ASCII frame writer
(fclaw2d_source/fc2d_geoclaw_output_ascii_fort.f: fc2d_geoclaw_fort_write_header,
fc2d_geoclaw_fort_write_file) and the reference gauge sampler
(fortran_source/geoclaw_gauges_fort.f90: fc2d_geoclaw_update_gauge), both
compiled by oracle/Makefile (target ``ref``):

* frame: the fort.tNNNN and fort.qNNNN bytes the reference writes for a
  header and a sequence of patches of several shapes whose values include
  the format's edge cases (signed zeros, values at and below the 1e-99
  flush, three-digit exponents, rounding ties, Inf/NaN in the geometry);
* gauge: q and bathymetry sampled at random points of random 8x8 patches
  (interpolated, and the dry fallback).

    make -C oracle ref && python tests/golden/make_golden_output.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Params, RefLib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "output_vectors.npz")
MBC = 2
SPECIAL = [1.0, -1.0, 0.0, -0.0, 1e-100, -1e-100, 1e-99, 9.99e-100, 1.5e100, -2.5e-150,
           0.99999999999999999, 1234567890123456.5, 1234567890123457.5, 123.456, 5e-324,
           1.7976931348623157e308, -1e-99, 1e-98, 99999999999999999.0, 0.5, 2.5e-5, -7.25e12]
SHAPES = [(4, 3), (8, 8), (2, 5), (16, 16), (1, 1), (32, 32)]
FRAME = 123457          # file names keep the last four digits: fort.q3457
TIME = 3141.5926535897932


def frame_case(rng):
    patches, meta = [], []
    for k, (mx, my) in enumerate(SHAPES):
        q = rng.normal(0, 10, (my + 2 * MBC, mx + 2 * MBC, 3)) * 10.0 ** rng.integers(-120, 120, (my + 4, mx + 4, 3))
        aux = rng.normal(0, 5, (my + 2 * MBC, mx + 2 * MBC, 1))
        inner = q[MBC:MBC + my, MBC:MBC + mx].reshape(-1)
        idx = rng.choice(inner.size, min(len(SPECIAL), inner.size), replace=False)
        inner[idx] = SPECIAL[:len(idx)]
        q[MBC:MBC + my, MBC:MBC + mx] = inner.reshape(my, mx, 3)
        if k == 0:
            aux[MBC, MBC, 0] = -q[MBC, MBC, 0]          # eta exactly 0
            aux[MBC, MBC + 1, 0] = 1e-100 - q[MBC, MBC + 1, 0]
        geom = [float(rng.normal()), float(rng.normal()) * 1e3, float(rng.uniform(0.1, 3)),
                float(rng.uniform(1e-5, 1))]
        if k == 4:
            geom = [-0.0, float("inf"), float("nan"), 1e-300]
        ints = [k * 7 if k != 5 else 123456, k % 4 - (k == 2) * 5, k // 2, 0]
        patches.append((q, aux, geom, ints))
    return patches


def main():
    ref = RefLib()
    ref.set_params(Params())
    rng = np.random.default_rng(2024)
    patches = frame_case(rng)
    work = tempfile.mkdtemp()
    tb, qb = ref.write_frame(work, FRAME, TIME,
                             [(q.copy(), a.copy(), *g, i[0], i[1], i[2]) for q, a, g, i in patches],
                             ngrids=len(patches))
    rec = {}
    for k, (q, a, g, i) in enumerate(patches):
        rec[f"frame_q{k}"] = q
        rec[f"frame_aux{k}"] = a
    rec["frame_geom"] = np.array([p[2] for p in patches])
    rec["frame_ints"] = np.array([p[3] for p in patches], dtype=np.int32)
    rec["frame_hdr"] = np.array([FRAME, len(patches)], dtype=np.int64)
    rec["frame_time"] = np.array([TIME])
    rec["frame_t_bytes"] = np.frombuffer(tb, dtype=np.uint8)
    rec["frame_q_bytes"] = np.frombuffer(qb, dtype=np.uint8)
    # gauges: 8x8 patches, values on a 1/64 grid, some depths dry
    n, mx = 96, 8
    gq = np.round(rng.normal(0, 1, (n, mx + 4, mx + 4, 3)) * 64) / 64
    gq[..., 0] = np.abs(gq[..., 0])
    dry = rng.random((n, mx + 4, mx + 4)) < 0.08
    gq[..., 0][dry] = np.round(rng.uniform(0, 2e-4, dry.sum()) * 2 ** 20) / 2 ** 20
    ga = np.round(rng.normal(0, 4, (n, mx + 4, mx + 4, 1)) * 64) / 64
    ggeom = np.stack([np.round(rng.normal(0, 100, n)), np.round(rng.normal(0, 100, n)),
                      rng.choice([0.5, 1.0, 2.0, 0.25], n), rng.choice([0.5, 1.0, 2.0, 0.25], n)], 1)
    pts = rng.uniform(0, mx, (n, 2))
    pts[::7] = np.floor(pts[::7])          # points on cell edges
    gxy = np.stack([ggeom[:, 0] + pts[:, 0] * ggeom[:, 2], ggeom[:, 1] + pts[:, 1] * ggeom[:, 3]], 1)
    gout = np.zeros((n, 4))
    for k in range(n):
        qv, av = ref.update_gauge(gq[k].copy(), ga[k].copy(), *ggeom[k], *gxy[k])
        gout[k, :3] = qv
        gout[k, 3] = av[0]
    rec.update(gauge_q=gq, gauge_aux=ga, gauge_geom=ggeom, gauge_xy=gxy, gauge_out=gout)
    np.savez_compressed(OUT, **rec)
    print(OUT, os.path.getsize(OUT), "bytes;", len(tb), "+", len(qb), "frame bytes;", n, "gauge samples")


if __name__ == "__main__":
    main()
