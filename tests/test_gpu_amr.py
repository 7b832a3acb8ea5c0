"""Resident AMR on the device against the oracle's AMR cycle.

geoflood_amd.amr.AmrRun keeps every leaf of a quadtree forest in pool slots:
each attempt is one replay of the pool's graph (multi-level ghost update over
the stored coarse/fine links, one persistent step of all levels with one dt,
device dt control); every few steps the leaves are tagged on the device and
the forest is rebuilt in the other half of the pool by slot moves,
interpolate2fine and average2coarse.  oracle.amr_driver.OracleAmr runs the
same cycle with the oracle's routines over NaN-filled ghost cells.  The run
bar: identical forests and tags at every regrid, identical dt sequence and
identical values in every cell."""
import numpy as np
import pytest

from geoflood_amd.amr import AmrRun
from oracle.amr_driver import OracleAmr, eta_gradient_tags
from oracle.oracle import Params
from _parity import assert_same
from test_amr_forest import band_flags

pytestmark = pytest.mark.gpu


def _basin_dem(nx, ny, L):
    x = (np.arange(nx) + 0.5) / nx
    X, Y = np.meshgrid(x, x)
    return 0.35 * np.exp(-((X - 0.6) ** 2 + (Y - 0.4) ** 2) / 0.02) - 0.1 * X


def _hump(L, i, j, aux):
    """A hump of water over the basin: wet and dry cells at every level."""
    mx = aux.shape[1] - 4
    q = np.zeros(aux.shape[:3] + (3,))
    for k in range(len(L)):
        n = 2 * mx << L[k]
        c = (np.arange(-2, mx + 2) + 0.5) / n
        X, Y = np.meshgrid(c + i[k] * mx / n, c + j[k] * mx / n)
        q[k, ..., 0] = np.maximum(0.25 + 0.3 * np.exp(-((X - 0.3) ** 2 + (Y - 0.6) ** 2) / 0.01) - aux[k, ..., 0], 0.0)
    q[..., 0][q[..., 0] < 0.02] = 0.0
    return q


def _compare(dev: AmrRun, ora: OracleAmr, dts):
    assert dev.forest.same(ora.forest)
    info = dev.pool.scenario(dev.sid)
    assert info["retakes"] == ora.retakes and info["nsteps"] == ora.nsteps
    assert_same(dts, ora.dts)
    q, _ = dev.leaves_state()
    got = np.moveaxis(q[:, 2:-2, 2:-2], -1, 1)
    ref = ora.state()
    assert_same(np.moveaxis(got, 1, 0), np.moveaxis(ref, 1, 0))


def _run_both(dev, ora, steps):
    """Advance both by `steps` accepted steps; the device dt of each accepted step."""
    dts = []
    info = dev.pool.scenario(dev.sid)
    for _ in range(steps):
        n0 = info["nsteps"]
        dev.run(1)
        info = dev.pool.scenario(dev.sid)
        assert info["nsteps"] == n0 + 1
        dts.append(info["last_dt"])
    ora.run(steps)
    return dts


@pytest.mark.parametrize("graph", [True, False])
def test_moving_band_regrid_matches_oracle(gf, oracle_c, graph):
    """Levels 0-2 over a 2 x 2 brick of 16 x 16 patches, walls; a refinement
    band moving east refines and coarsens at every regrid (every 2 steps);
    20 steps through the graph (or the per-patch loop)."""
    print(f"case amr moving band graph={graph}", flush=True)
    gf.set_params(manning_n=0.03)
    oracle_c.set_params(Params(manning_n=0.03))
    log = []
    dev = AmrRun(16, 2, 2, 1.0, 1.0, 0, 2, _basin_dem, 64, eta_tol=1e9, regrid_interval=2, flags_fn=band_flags(),
                 log=log.append)
    ora = OracleAmr(oracle_c, 16, 2, 2, 1.0, 1.0, 0, 2, _basin_dem, eta_tol=1e9, regrid_interval=2,
                    flags_fn=band_flags())
    try:
        dev.start(0.0, 1e-4, state_fn=_hump, init_levels=False)
        ora.start(0.0, 1e-4, state_fn=_hump, init_levels=False)
        dts = []
        info = dev.pool.scenario(dev.sid)
        while info["nsteps"] < 20:
            n0 = info["nsteps"]
            dev.run(1, graph=graph)
            info = dev.pool.scenario(dev.sid)
            dts.append(info["last_dt"])
        ora.run(20)
        assert len(set(dev.stats["patches"])) > 2 and dev.stats["regrids"] == 10
        _compare(dev, ora, dts)
        # one log record per regrid, at the regrid's step, with the new forest
        assert [r["nsteps"] for r in log] == list(range(2, 21, 2))
        assert [r["leaves"] for r in log] == dev.stats["patches"][1:]
        assert any(r["changed"] for r in log) and all(r["maxcfl"] > 0 for r in log)
    finally:
        dev.close()
        gf.set_params()
        oracle_c.set_params(Params())


def test_tag_eta_gradient_matches_oracle(gf, oracle_c):
    """Device surface-gradient tags of every leaf of a two-level forest
    (random wet/dry states, ghost cells from the multi-level ghost update)
    equal the oracle's on the same forest, for several thresholds."""
    print("case amr eta-gradient tags", flush=True)
    rng = np.random.default_rng(5)

    def state(L, i, j, aux):
        q = np.zeros(aux.shape[:3] + (3,))
        q[..., 0] = np.where(rng.random(aux.shape[:3]) < 0.002, 0.005, 0.0)   # a few wet cells
        return q

    dev = AmrRun(16, 2, 2, 1.0, 1.0, 0, 2, _basin_dem, 64, regrid_interval=2, flags_fn=band_flags())
    ora = OracleAmr(oracle_c, 16, 2, 2, 1.0, 1.0, 0, 2, _basin_dem, regrid_interval=2, flags_fn=band_flags())
    try:
        dev.start(0.0, 1e-4, init_levels=False)
        ora.start(0.0, 1e-4, init_levels=False)
        for _ in range(2):                       # two regrids of the moving band
            dev.regrid(); ora.regrid()
        assert dev.forest.same(ora.forest) and len(dev.forest.counts()) >= 2
        q = state(dev.forest.L, dev.forest.i, dev.forest.j, np.zeros((len(dev.forest), 20, 20, 1)))
        base = dev.half * dev.capacity
        _, aux = dev.leaves_state()
        inv = np.empty(len(dev.perm), dtype=np.int64)
        inv[dev.perm] = np.arange(len(dev.perm))
        dev.pool.upload_aos(base, q[inv], aux[inv], buf=0)
        ora.q = [x.copy() for x in q]
        for tol in (0.15, 0.5, 1.0):   # the bed slopes 0.1 everywhere, up to ~2 on the bump
            dev.eta_tol = ora.eta_tol = tol
            got = dev._tags(dev.forest, 0, 0.0)
            want = ora.tags(0.0)
            assert np.array_equal(got.astype(bool), want), tol
            assert 0 < want.sum() < len(want)
    finally:
        dev.close()


def _overland(maxlevel):
    from geoflood_amd.scenarios import gaussian_sum_dem, sine_hydrograph
    t, hu = sine_hydrograph(200.0, 100.0)
    hydro = dict(time=t, hu=hu, b=100.0, x0=0.0, y0=1000.0)
    dem = lambda nx, ny, L: gaussian_sum_dem(nx, 2000.0, seed=0)  # noqa: E731
    kw = dict(mthbc=(0, 1, 3, 3), hydro=hydro, eta_tol=0.01, regrid_interval=4, dt_max=10.0)
    return (16, 4, 4, 2000.0, 2000.0, 1, maxlevel, dem), kw


def test_overland_flood_amr_matches_oracle(gf, oracle_c):
    """The AMR overland flood at reduced size: 2 km square, 16 x 16 patches
    at levels 1-3 (128^2 cells at level 1), Gaussian-sum DEM (seed 0),
    hydrograph inflow on the west edge (100 m strip, Q = 200 sin(pi t /
    3600) m^3/s), east extrapolation, north/south walls, Manning n 0.035,
    tagged on h > dry_tolerance or |grad eta| > 0.01, regrid every 4 steps,
    from a dry bed at t = 0 (dt bound 10 s): 48 steps."""
    print("case amr overland flood levels 1-3, 48 steps", flush=True)
    gf.set_params(manning_n=0.035)
    oracle_c.set_params(Params(manning_n=0.035))
    args, kw = _overland(3)
    dev = AmrRun(*args[:8], 1024, **kw)
    ora = OracleAmr(oracle_c, *args, **kw)
    try:
        dev.start(0.0, 1.0)
        ora.start(0.0, 1.0)
        assert dev.forest.same(ora.forest)
        dts = _run_both(dev, ora, 48)
        assert len(dev.tag_log) == len(ora.tag_log) == 12
        for a, b in zip(dev.tag_log, ora.tag_log):
            assert np.array_equal(a.astype(bool), b)
        _compare(dev, ora, dts)
        q, _ = dev.leaves_state()
        assert (q[:, 2:-2, 2:-2, 0] > 1e-3).sum() > 100      # water has entered
    finally:
        dev.close()
        gf.set_params()
        oracle_c.set_params(Params())
