"""Hydrograph inflow on the device (k_inflow through the C ABI) against the
CPU oracle's bc2 (mthbc = 0, pinned bitwise to the reference build) and the
oracle's brick driver.

Bar: identical values.  Ghost states, dt sequences and every cell of a run
equal the oracle's, evaluated with correctly rounded powers as the device
is.
"""
import numpy as np
import pytest

from oracle.oracle import Brick, Hydrograph, Params
from geoflood_amd.pool import BrickLayout, Pool
from geoflood_amd.scenarios import overland_flood
from _parity import assert_same

pytestmark = pytest.mark.gpu

MANNING = 0.035


@pytest.fixture
def flood_params(gf, oracle_c):
    oracle_c.set_params(Params(manning_n=MANNING))
    gf.set_params(manning_n=MANNING)
    yield
    oracle_c.set_params(Params())
    gf.set_params()


def _setup(n, mx, peak, y0=None, mthbc=(0, 1, 3, 3), seed=0, rows=None):
    sc = overland_flood(n, peak_q=peak, seed=seed)
    hy = sc["hydro"]
    if y0 is not None:
        hy["y0"] = y0
    if rows is not None:
        from geoflood_amd.scenarios import sine_hydrograph
        hy["time"], hy["hu"] = sine_hydrograph(peak, hy["b"], rows=rows)
    H = Hydrograph(hy["time"], hy["hu"], hy["b"], hy["x0"], hy["y0"], (1, 0, 0, 0))
    m = n // mx
    br = Brick.from_global(sc["h"], sc["hu"], sc["hv"], sc["b"], m, m, 2, sc["dx"], sc["dx"],
                           mthbc=mthbc, hydro=H)
    layout = BrickLayout(m, m, mx, mx, sc["dx"], sc["dx"], mthbc=mthbc)
    pool = Pool(mx, mx, layout.npatch)
    pool.set_layout([layout])
    pool.upload_global(layout, sc["h"], sc["hu"], sc["hv"], sc["b"])
    return sc, hy, br, layout, pool


@pytest.mark.parametrize("t,y0,rows", [(900.0, None, None), (0.0, None, None), (1800.0, 0.0, None),
                                       (64, None, 200), (130, None, 200), (2999.7, None, 200)])
def test_inflow_ghost_fill_forestclaw_order(gf, oracle_c, flood_params, t, y0, rows):
    """Random wet/dry states on a 4x4 brick of 16x16 patches: every cell of
    every patch (ghost ring, corners, aux) after the device ghost fill equals
    the oracle's ForestClaw-order fill with bc2's hydrograph side.  The
    61-row table's cases fall on knots (t = 0, 900, 1800 s); a 200-row table
    (the device searches 64 intervals at a time) is taken at its row 64 (a
    knot on the boundary of the first 64 intervals: the earlier interval
    holds it), at row 130 and between knots in the third 64."""
    rng = np.random.default_rng(int(t) + 7)
    n, mx = 64, 16
    sc, hy, br, layout, pool = _setup(n, mx, 200.0, y0=y0, rows=rows)
    if rows is not None and isinstance(t, int):
        t = float(hy["time"][t])                 # exactly on the table's knot
    print(f"case inflow ghost fill t={t!r} y0={y0} rows={rows}", flush=True)
    h = np.where(rng.random((n, n)) < 0.5, 0.0, rng.uniform(0, 2, (n, n)))
    h[rng.random((n, n)) < 0.05] = 1e-3
    hu = rng.normal(0, 0.5, (n, n)) * h
    hv = rng.normal(0, 0.5, (n, n)) * h
    br = Brick.from_global(h, hu, hv, sc["b"], 4, 4, 2, sc["dx"], sc["dx"], mthbc=br.mthbc, hydro=br.hydro)
    pool.upload_global(layout, h, hu, hv, sc["b"])
    sid = pool.add_scenario(layout, dt0=1.0, target_steps=1)
    pool.set_inflow(sid, hy["time"], hy["hu"], hy["b"], hy["x0"], hy["y0"])
    pool.ghost_fill(buf=0, t=t)
    assert oracle_c.brick_ghost_fill(br, t) == 0
    got = pool.download_aos(0, layout.npatch, buf=0)
    aux = pool.download_aux(0, layout.npatch)
    for k in range(layout.npatch):
        assert_same(got[k], br.q[k], k)
        assert np.array_equal(aux[k], br.aux[k]), k
    pool.close()


def _run_device(pool, layout, hy, steps, t0, dt0, mode):
    sid = pool.add_scenario(layout, dt0=dt0, target_steps=steps, t0=t0)
    pool.set_inflow(sid, hy["time"], hy["hu"], hy["b"], hy["x0"], hy["y0"])
    dts, prev = [], 0
    while True:
        active = pool.run_plain(1) if mode == "plain" else pool.run(1, graph=(mode == "graph"))
        info = pool.scenario(sid)
        if info["nsteps"] > prev:
            dts.append(info["last_dt"])
            prev = info["nsteps"]
        if not active:
            break
    return info, np.array(dts)


@pytest.mark.parametrize("mode,mx,peak,y0", [("graph", 16, 200.0, None), ("stream", 16, 2000.0, None),
                                             ("plain", 32, 200.0, None), ("graph", 16, 600.0, 40.0)])
def test_overland_flood_run(gf, oracle_c, flood_params, mode, mx, peak, y0):
    """Initially dry overland flood fed by the west-edge hydrograph, 100
    accepted steps from t = 600 s: states, dt sequence and wet/dry mask
    against the oracle."""
    n, steps = 128, 100
    print(f"case overland flood {n}^2 mx={mx} peak={peak} y0={y0} mode={mode}", flush=True)
    sc, hy, br, layout, pool = _setup(n, mx, peak, y0=y0)
    rec = oracle_c.brick_run(br, steps, 1.0, 0.9, 1.0, t0=600.0, nthreads=8)
    info, dts = _run_device(pool, layout, hy, steps, 600.0, 1.0, mode)
    assert info["nsteps"] == steps and info["retakes"] == rec["retakes"]
    assert_same(dts, rec["dt"])
    got = np.stack(pool.download_global(layout, buf=info["cur"]))
    ref = br.to_global()
    assert (ref[0] > 1e-3).sum() > 100  # the inflow has wetted part of the domain
    assert_same(got, ref)
    pool.close()


def test_inflow_without_hydrograph_is_rejected(gf, flood_params):
    sc, hy, br, layout, pool = _setup(64, 16, 200.0)
    pool.add_scenario(layout, dt0=1.0, target_steps=1)
    with pytest.raises(gf.GeoFloodError, match="no hydrograph"):
        pool.run(1)
    pool.close()
