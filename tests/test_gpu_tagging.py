"""Refinement tagging of resident patches on the device against the oracle's
tag4refinement / tag4coarsening (pinned to the reference build)."""
import numpy as np
import pytest

from oracle.oracle import Refine
from geoflood_amd.pool import BrickLayout, Pool

pytestmark = pytest.mark.gpu


def _criteria(rng, k):
    fg = () if k % 3 == 0 else ((float(rng.uniform(0, 2)), int(rng.integers(1, 4)), int(rng.integers(2, 5))),)
    regs = () if k % 2 else ((float(rng.uniform(0, 60)), float(rng.uniform(0, 60)), float(rng.uniform(60, 128)),
                             float(rng.uniform(60, 128)), 0.0, 5.0, int(rng.integers(2, 5)), 6),)
    return dict(wave_tolerance=float(rng.uniform(0, 1)), deep_depth=float(rng.uniform(0, 5)),
                max_level_deep=int(rng.integers(1, 5)), max_velocity_depth_product=float(rng.uniform(0.1, 5)),
                speed_tolerance=tuple(float(x) for x in rng.uniform(0, 3, int(rng.integers(1, 5)))),
                flowgrades=fg, sea_level=float(rng.uniform(-1, 1)), regions=regs)


@pytest.mark.parametrize("mx", [16, 32])
def test_tags_match_oracle(gf, oracle_c, mx):
    rng = np.random.default_rng(mx)
    n = 128
    m = n // mx
    h = np.where(rng.random((n, n)) < 0.9, 0.0, rng.uniform(0, 2, (n, n)))
    hu = rng.normal(0, 0.2, (n, n)) * (h > 0)
    hv = rng.normal(0, 0.2, (n, n)) * (h > 0)
    b = rng.normal(0, 3, (n, n))
    layout = BrickLayout(m, m, mx, mx, 1.0, 1.0, xlo=-3.0, ylo=2.0)
    pool = Pool(mx, mx, layout.npatch)
    pool.set_layout([layout])
    pool.upload_global(layout, h, hu, hv, b)
    q = pool.download_aos(0, layout.npatch)
    aux = pool.download_aux(0, layout.npatch)
    ox, oy = layout.origins()
    fams = [[(2 * fj) * m + 2 * fi, (2 * fj) * m + 2 * fi + 1, (2 * fj + 1) * m + 2 * fi,
             (2 * fj + 1) * m + 2 * fi + 1] for fj in range(m // 2) for fi in range(m // 2)]
    for k in range(12):
        crit = _criteria(rng, k)
        rf = Refine(**crit)
        level, maxlevel, t, init = int(rng.integers(1, 5)), 5, float(rng.uniform(0, 6)), k % 2
        print(f"case mx={mx} criteria {k}", flush=True)
        got = pool.tag_refine(np.arange(layout.npatch), t, level, maxlevel, init, **crit)
        ref = [oracle_c.tag4refinement(rf, mx, mx, 2, np.ascontiguousarray(q[p]), np.ascontiguousarray(aux[p]),
                                       ox[p], oy[p], 1.0, 1.0, t, level, maxlevel, init)
               for p in range(layout.npatch)]
        assert list(got) == ref
        gotc = pool.tag_coarsen(fams, t, level, maxlevel, init, **crit)
        refc = [oracle_c.tag4coarsening(rf, mx, mx, 2, [np.ascontiguousarray(q[p]) for p in f],
                                        [np.ascontiguousarray(aux[p]) for p in f], [ox[p] for p in f],
                                        [oy[p] for p in f], 1.0, 1.0, t, level, maxlevel, init)
                for f in fams]
        assert list(gotc) == refc
    pool.close()
