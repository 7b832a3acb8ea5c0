"""Pins the oracle's hydrograph inflow (interpolation, Riemann-invariant and
two-shock boundary depths, bc2 mthbc = 0) to the reference: golden vectors
written from the reference build (tests/golden/make_golden_inflow.py), the
reference build itself where present, and what the mathematics fixes
(linear interpolation at the knots, the Newton residual, the wall branch)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle.oracle import Hydrograph

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inflow_vectors.npz")
G = 9.81
TOL = float(np.float32(1.0e-6))
MX, MY, MBC = 16, 12, 2


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def hydro(gold, strip=(100.0, 0.0, 1000.0), sides=(1, 0, 0, 0)):
    return Hydrograph(gold["table_t"], gold["table_hu"], *strip, sides)


def test_golden_interpolation(oracle_c, gold):
    H = hydro(gold)
    got = np.array([oracle_c.hydro_interpolate(H, t) for t in gold["interp_t"]])
    assert np.array_equal(got, gold["interp_hu"])


def test_interpolation_knots_clamp_and_linearity(oracle_c, gold):
    t, hu = gold["table_t"], gold["table_hu"]
    H = hydro(gold)
    # t(1) is matched by the first interval's left end; later knots are the
    # right end of the previous interval, reached through the slope.
    assert oracle_c.hydro_interpolate(H, t[0]) == hu[0]
    for k in range(1, len(t)):
        assert abs(oracle_c.hydro_interpolate(H, t[k]) - hu[k]) < 1e-15
    assert oracle_c.hydro_interpolate(H, -10.0) == hu[0]
    assert oracle_c.hydro_interpolate(H, 1e5) == hu[-1]
    mid = 0.5 * (t[10] + t[11])
    assert abs(oracle_c.hydro_interpolate(H, mid) - 0.5 * (hu[10] + hu[11])) < 1e-14


def test_golden_newton(oracle_c, gold):
    ri = np.array([oracle_c.riemann_invariants(a, b, G)[0]
                   for a, b in zip(gold["newton_q0"], gold["newton_q1"])])
    ts = np.array([oracle_c.two_shock(a, b, G)[0]
                   for a, b in zip(gold["newton_q0"], gold["newton_q1"])])
    assert np.array_equal(ri, gold["ri_h0"], equal_nan=True)
    assert np.array_equal(ts, gold["ts_h0"], equal_nan=True)


def test_newton_residuals(oracle_c):
    """The returned depth solves the invariant / two-shock relation to the
    reference's tolerance whenever the iteration converges."""
    rng = np.random.default_rng(5)
    n_conv = 0
    for _ in range(300):
        hu0 = rng.uniform(0.05, 10.0)
        h1, hu1 = rng.uniform(0.05, 4.0), rng.uniform(-1.0, 5.0)
        u1 = hu1 / h1
        h0 = oracle_c.riemann_invariants([0.0, hu0], [h1, hu1], G)[0]
        if np.isfinite(h0) and h0 > 0:
            r = hu0 / h0 - 2 * np.sqrt(G * h0) - u1 + 2 * np.sqrt(G * h1)
            if abs(r) < 1e-3:
                assert abs(r) < TOL
                n_conv += 1
        h0 = oracle_c.two_shock([0.0, hu0], [h1, hu1], G)[0]
        if np.isfinite(h0) and h0 > 0:
            r = hu0 / h0 - u1 - (h0 - h1) * np.sqrt(G / 2 * (1 / h0 + 1 / h1))
            if abs(r) < 1e-3:
                assert abs(r) < TOL
                n_conv += 1
    assert n_conv > 300


def test_golden_bc2(oracle_c, gold):
    for k in range(len(gold["bc2_mthbc"])):
        s = gold["bc2_strip"][k]
        H = hydro(gold, strip=tuple(s), sides=tuple(gold["bc2_sides"][k]))
        q = gold["bc2_q_in"][k].copy()
        aux = gold["bc2_aux_in"][k].copy()
        rc = oracle_c.bc2_ex(MX, MY, MBC, q, aux, gold["bc2_mthbc"][k], *gold["bc2_geom"][k], H)
        assert rc == 0
        assert np.array_equal(q, gold["bc2_q_out"][k], equal_nan=True), k
        assert np.array_equal(aux, gold["bc2_aux_out"][k]), k


def test_bc2_inflow_without_hydrograph_is_an_error(oracle_c):
    q = np.zeros((MY + 4, MX + 4, 3))
    aux = np.zeros((MY + 4, MX + 4, 1))
    assert oracle_c.bc2_ex(MX, MY, MBC, q, aux, [0, 3, 3, 3], 0, 0, 1, 1, 0.0, None) == -1


def test_bc2_inflow_outside_strip_is_a_wall(oracle_c, gold):
    """Rows outside the channel strip get exactly the wall (mthbc = 3) state."""
    rng = np.random.default_rng(3)
    for side in range(4):
        sides = [0, 0, 0, 0]
        sides[side] = 1
        H = hydro(gold, strip=(1.0, -1e6, -1e6), sides=tuple(sides))
        q = rng.normal(0, 1, (MY + 4, MX + 4, 3))
        q[..., 0] = np.abs(q[..., 0])
        aux = rng.normal(0, 1, (MY + 4, MX + 4, 1))
        q2, a2 = q.copy(), aux.copy()
        m0 = [-1, -1, -1, -1]
        m3 = list(m0)
        m0[side], m3[side] = 0, 3
        assert oracle_c.bc2_ex(MX, MY, MBC, q, aux, m0, 0, 0, 1, 1, 900.0, H) == 0
        oracle_c.bc2(MX, MY, MBC, q2, a2, m3)
        assert np.array_equal(q, q2) and np.array_equal(aux, a2)


def test_bc2_inflow_dry_cell_gets_critical_depth(oracle_c, gold):
    """A dry boundary cell inside the strip gets h = max((hu/sqrt g)^(2/3),
    0.001), the discharge in q(2) and zero q(3)."""
    H = hydro(gold, strip=(1e6, 0.0, 0.0))
    q = np.zeros((MY + 4, MX + 4, 3))
    aux = np.zeros((MY + 4, MX + 4, 1))
    t = 900.0
    hu = oracle_c.hydro_interpolate(H, t)
    assert oracle_c.bc2_ex(MX, MY, MBC, q, aux, [0, -1, -1, -1], 0, 0, 1, 1, t, H) == 0
    h = max((hu / np.sqrt(G)) ** (2.0 / 3.0), 0.001)
    assert np.allclose(q[:, 0:2, 0], h, rtol=1e-15, atol=0)
    assert np.all(q[:, 0:2, 1] == hu) and np.all(q[:, 0:2, 2] == 0.0)


def test_vs_reference_bc2_random(oracle_c, oracle_ref, gold):
    rng = np.random.default_rng(11)
    t, hu = gold["table_t"], gold["table_hu"]
    for k in range(200):
        side = int(rng.integers(0, 4))
        sides = [0, 0, 0, 0]
        sides[side] = 1
        mthbc = [int(rng.choice([-1, 1, 3])) for _ in range(4)]
        mthbc[side] = 0
        H = Hydrograph(t, hu, rng.uniform(2, 30), rng.uniform(-5, 20), rng.uniform(-5, 20),
                       tuple(sides))
        oracle_ref.set_hydrograph(H)
        q = np.zeros((MY + 4, MX + 4, 3))
        q[..., 0] = np.where(rng.random(q.shape[:2]) < 0.4, 0.0, rng.uniform(0, 3, q.shape[:2]))
        q[..., 1:] = rng.normal(0, 1, q.shape[:2] + (2,))
        aux = rng.normal(0, 1, (MY + 4, MX + 4, 1))
        geom = (rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.5, 2), rng.uniform(0.5, 2),
                float(rng.choice([0.0, rng.uniform(0, 3600)])))
        q2, a2 = q.copy(), aux.copy()
        assert oracle_c.bc2_ex(MX, MY, MBC, q, aux, mthbc, *geom, H) == 0
        oracle_ref.bc2_ex(MX, MY, MBC, q2, a2, mthbc, *geom)
        assert np.array_equal(q, q2, equal_nan=True) and np.array_equal(aux, a2)
