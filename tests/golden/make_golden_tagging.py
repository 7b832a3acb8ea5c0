"""Write tests/golden/tagging_vectors.npz from the reference refinement
tagging (fclaw2d_source/fc2d_geoclaw_tag4refinement_fort.f90,
fc2d_geoclaw_tag4coarsening_fort.f90, fc2d_geoflood_flag2refine.f90 with
refinement_data_module and the regions common block) compiled by
oracle/Makefile (target ``ref``): random 8x8 patches (wet/dry, moving,
over random bathymetry) under random criteria, flowgrades and regions, with
the reference's refinement tag and the coarsening tag of four siblings.

    make -C oracle ref && python tests/golden/make_golden_tagging.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Params, RefLib, Refine  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tagging_vectors.npz")
MX, MBC = 8, 2


def random_refine(rng, k):
    fg = () if k % 3 == 0 else tuple(
        (float(rng.uniform(0, 3)), int(rng.integers(1, 4)), int(rng.integers(1, 5)))
        for _ in range(int(rng.integers(1, 3))))
    regs = () if k % 2 else tuple(
        (float(rng.uniform(-5, 5)), float(rng.uniform(-5, 5)), float(rng.uniform(5, 20)),
         float(rng.uniform(5, 20)), 0.0, float(rng.uniform(0, 10)), int(rng.integers(1, 5)), 6)
        for _ in range(int(rng.integers(0, 3))))
    return Refine(wave_tolerance=float(rng.uniform(0, 1)), deep_depth=float(rng.uniform(0, 5)),
                  max_level_deep=int(rng.integers(1, 5)),
                  max_velocity_depth_product=float(rng.uniform(0.1, 5)),
                  speed_tolerance=tuple(float(x) for x in rng.uniform(0, 3, int(rng.integers(1, 6)))),
                  flowgrades=fg, sea_level=float(rng.uniform(-1, 1)), regions=regs)


def random_patch(rng):
    # values on a 1/256 grid (plus exact dry_tolerance cells) keep the file small
    sh = (MX + 2 * MBC, MX + 2 * MBC)
    q = np.zeros(sh + (3,))
    q[..., 0] = np.where(rng.random(sh) < 0.8, 0.0, np.round(rng.uniform(0, 2, sh) * 256) / 256)
    q[..., 0][rng.random(sh) < 0.02] = 1e-3
    q[..., 1:] = np.round(rng.normal(0, 0.3, sh + (2,)) * 256) / 256
    return q, np.round(rng.normal(0, 3, sh + (1,)) * 256) / 256


def main():
    ref = RefLib()
    ref.set_params(Params())
    rng = np.random.default_rng(777)
    n = 150
    rec = dict(q=[], aux=[], geom=[], ints=[], tag_refine=[], tag_coarsen=[])
    criteria = []
    for k in range(n):
        rf = random_refine(rng, k)
        ref.set_refinement(rf)
        qs, auxs = zip(*[random_patch(rng) for _ in range(4)])
        level = int(rng.integers(1, 6))
        maxlevel = int(rng.integers(level, 7))
        init = int(rng.integers(0, 2))
        t = float(rng.uniform(0, 12))
        xl, yl, dx = float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3)), float(rng.uniform(0.1, 1))
        xls = [xl, xl + MX * dx, xl, xl + MX * dx]
        yls = [yl, yl, yl + MX * dx, yl + MX * dx]
        rec["tag_refine"].append(ref.tag4refinement(MX, MX, MBC, qs[0], auxs[0], xl, yl, dx, dx, t, level,
                                                    maxlevel, init))
        rec["tag_coarsen"].append(ref.tag4coarsening(MX, MX, MBC, qs, auxs, xls, yls, dx, dx, t, level,
                                                     maxlevel, init))
        rec["q"].append(np.stack(qs))
        rec["aux"].append(np.stack(auxs))
        rec["geom"].append([xl, yl, dx, t])
        rec["ints"].append([level, maxlevel, init])
        criteria.append(dict(rf.__dict__ if not hasattr(rf, "_sp") else
                             {k2: v for k2, v in rf.__dict__.items() if not k2.startswith("_")}))
    out = {k: np.asarray(v) for k, v in rec.items()}
    out["criteria_json"] = np.array(json.dumps(criteria))
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()}, "refine tags", int(out["tag_refine"].sum()),
          "coarsen tags", int(out["tag_coarsen"].sum()))


if __name__ == "__main__":
    main()
