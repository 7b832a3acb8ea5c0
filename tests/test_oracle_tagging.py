"""Pins the oracle's refinement tagging (flag2refine, tag4refinement,
tag4coarsening) to the reference: golden tags written from the reference
build (tests/golden/make_golden_tagging.py), the reference build on fresh
random patches where present, and cases the criteria fix by themselves."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from oracle.oracle import Refine

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tagging_vectors.npz")
MX, MBC = 8, 2


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _refine(d):
    return Refine(**{k: (tuple(tuple(x) if isinstance(x, list) else x for x in v) if isinstance(v, list) else v)
                     for k, v in d.items()})


def test_golden_tags(oracle_c, gold):
    crit = json.loads(str(gold["criteria_json"]))
    for k in range(len(crit)):
        rf = _refine(crit[k])
        qs = [np.ascontiguousarray(a) for a in gold["q"][k]]
        auxs = [np.ascontiguousarray(a) for a in gold["aux"][k]]
        xl, yl, dx, t = gold["geom"][k]
        level, maxlevel, init = (int(v) for v in gold["ints"][k])
        assert oracle_c.tag4refinement(rf, MX, MX, MBC, qs[0], auxs[0], xl, yl, dx, dx, t, level, maxlevel,
                                       init) == gold["tag_refine"][k], k
        xls = [xl, xl + MX * dx, xl, xl + MX * dx]
        yls = [yl, yl, yl + MX * dx, yl + MX * dx]
        assert oracle_c.tag4coarsening(rf, MX, MX, MBC, qs, auxs, xls, yls, dx, dx, t, level, maxlevel,
                                       init) == gold["tag_coarsen"][k], k


def _patch(h=0.0, hu=0.0, hv=0.0, b=0.0):
    q = np.zeros((MX + 4, MX + 4, 3))
    q[..., 0], q[..., 1], q[..., 2] = h, hu, hv
    return q, np.full((MX + 4, MX + 4, 1), b)


def test_tagging_closed_cases(oracle_c):
    rf = Refine(wave_tolerance=0.1, deep_depth=100.0, max_level_deep=3, speed_tolerance=(1.0, 2.0))
    q, aux = _patch()
    # all dry, no regions: nothing to refine, and siblings may coarsen
    assert oracle_c.tag4refinement(rf, MX, MX, MBC, q, aux, 0, 0, 1, 1, 0.0, 1, 4) == 0
    assert oracle_c.tag4coarsening(rf, MX, MX, MBC, [q] * 4, [aux] * 4, [0, 8, 0, 8], [0, 0, 8, 8],
                                   1, 1, 0.0, 2, 4) == 1
    # one wet cell above the wave tolerance at a level below max_level_deep refines
    q2 = q.copy()
    q2[5, 5, 0] = 0.5
    assert oracle_c.tag4refinement(rf, MX, MX, MBC, q2, aux, 0, 0, 1, 1, 0.0, 1, 4) == 1
    # a region forcing level 3 refines a level-2 patch it overlaps, not one it misses
    rg = Refine(regions=((2.0, 2.0, 3.0, 3.0, 0.0, 10.0, 3, 5),))
    assert oracle_c.tag4refinement(rg, MX, MX, MBC, q, aux, 0, 0, 1, 1, 1.0, 2, 4) == 1
    assert oracle_c.tag4refinement(rg, MX, MX, MBC, q, aux, 10, 10, 1, 1, 1.0, 2, 4) == 0
    assert oracle_c.tag4refinement(rg, MX, MX, MBC, q, aux, 0, 0, 1, 1, 11.0, 2, 4) == 0
    # coarsening uses th_factor = 0 (an INTEGER 0.5 in the reference): any moving wet cell blocks it
    q3, _ = _patch(h=1.0, b=-200.0)
    q3[4, 4, 1] = 1e-9
    slow = Refine(wave_tolerance=1e9, deep_depth=100.0, speed_tolerance=(5.0,), max_velocity_depth_product=1e9)
    assert oracle_c.tag4refinement(slow, MX, MX, MBC, q3, _patch(b=-200.0)[1], 0, 0, 1, 1, 0.0, 1, 4) == 0
    assert oracle_c.tag4coarsening(slow, MX, MX, MBC, [q3] * 4, [_patch(b=-200.0)[1]] * 4, [0, 8, 0, 8],
                                   [0, 0, 8, 8], 1, 1, 0.0, 2, 4) == 0
    # no coarsening while the initial grids are built
    assert oracle_c.tag4coarsening(slow, MX, MX, MBC, [q3] * 4, [_patch(b=-200.0)[1]] * 4, [0, 8, 0, 8],
                                   [0, 0, 8, 8], 1, 1, 0.0, 2, 4, init_flag=1) == 1


def test_flowgrade_depth_threshold(oracle_c):
    """Depth flowgrade: refines below its min level exactly when some cell
    is deeper than the threshold."""
    for h, expect in ((0.49, 0), (0.51, 1)):
        q, aux = _patch()
        q[6, 3, 0] = h
        rf = Refine(flowgrades=((0.5, 1, 3),), speed_tolerance=(1e9,))
        assert oracle_c.tag4refinement(rf, MX, MX, MBC, q, aux, 0, 0, 1, 1, 0.0, 2, 4) == expect
        assert oracle_c.tag4refinement(rf, MX, MX, MBC, q, aux, 0, 0, 1, 1, 0.0, 3, 4) == 0


def test_vs_reference_random(oracle_c, oracle_ref):
    rng = np.random.default_rng(99)
    for k in range(400):
        rf = Refine(wave_tolerance=float(rng.uniform(0, 1)), deep_depth=float(rng.uniform(0, 5)),
                    max_level_deep=int(rng.integers(1, 5)), max_velocity_depth_product=float(rng.uniform(0.1, 5)),
                    speed_tolerance=tuple(rng.uniform(0, 3, int(rng.integers(1, 6)))),
                    flowgrades=() if k % 3 == 0 else ((float(rng.uniform(0, 3)), int(rng.integers(1, 4)), 3),),
                    sea_level=float(rng.uniform(-1, 1)),
                    regions=() if k % 2 else ((-2.0, -2.0, float(rng.uniform(0, 9)), 9.0, 0.0, 5.0, 4, 6),))
        oracle_ref.set_refinement(rf)
        sh = (MX + 4, MX + 4)
        q = np.zeros(sh + (3,))
        q[..., 0] = np.where(rng.random(sh) < 0.8, 0.0, rng.uniform(0, 2, sh))
        q[..., 1:] = rng.normal(0, 0.3, sh + (2,))
        aux = rng.normal(0, 3, sh + (1,))
        args = (MX, MX, MBC, q, aux, float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3)), 0.5, 0.5,
                float(rng.uniform(0, 8)), int(rng.integers(1, 5)), 5, int(rng.integers(0, 2)))
        assert oracle_c.tag4refinement(rf, *args) == oracle_ref.tag4refinement(*args)
