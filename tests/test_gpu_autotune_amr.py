"""The autotuner under dynamic AMR (the moving dam-break front at reduced
size) against the oracle's AMR cycle.

A 2 km square of level-0 bricks (32/mx patches a side, levels 1-3, finest
cells 7.8 m), the overland-flood DEM, walls, a 30 m deep reservoir over
x < 250 m, y >= 1750 m released at t = 0; leaves tagged where wet, regrid
every 2 steps, so the front refines the forest as it spreads.  The device
switches step variant every regrid (every tile shape, thread count, number
of patches per workgroup and LDS split of the patch size in turn): forests,
dt sequence and states must still match the oracle's run (identical forests,
identical dt sequence and cell values).  The tuning cache persists its choices: a
second run over the same populations times nothing."""
import numpy as np
import pytest

from geoflood_amd.amr import AmrRun
from geoflood_amd.autotune import TuneCache
from geoflood_amd.scenarios import gaussian_sum_dem, reservoir_state
from oracle.amr_driver import OracleAmr
from oracle.oracle import Params
from _parity import assert_same

pytestmark = pytest.mark.gpu

_DEM = {}


def _dem(nx, ny, L):
    if nx not in _DEM:
        _DEM[nx] = gaussian_sum_dem(nx, 2000.0, seed=0)
    return _DEM[nx]


def _args(mx):
    nx0 = 32 // mx
    return (mx, nx0, nx0, 2000.0, 2000.0, 1, 3, _dem), dict(eta_tol=1e30, regrid_interval=2)


@pytest.mark.parametrize("mx", [8, 16])
def test_front_amr_with_variant_switches_matches_oracle(gf, oracle_c, mx):
    print(f"case autotune amr front mx={mx}", flush=True)
    gf.set_params(manning_n=0.035)
    oracle_c.set_params(Params(manning_n=0.035))
    args, kw = _args(mx)
    dev = AmrRun(*args, (args[1] << 3) ** 2, **kw)
    ora = OracleAmr(oracle_c, *args, **kw)
    try:
        dev.start(0.0, 0.01, state_fn=reservoir_state(dev))
        ora.start(0.0, 0.01, state_fn=reservoir_state(dev))
        assert dev.forest.same(ora.forest)
        nv = len(dev.pool.variants()[1])
        dts, counts = [], []
        for k in range(8):
            dev.pool.set_variant(k % nv)
            n0 = dev.pool.scenario(dev.sid)["nsteps"]
            dev.run(2)
            info = dev.pool.scenario(dev.sid)
            assert info["nsteps"] == n0 + 2
            dts.append(info["last_dt"])
            ora.run(2)
            assert dev.forest.same(ora.forest), k
            counts.append(len(dev.forest))
        assert counts[-1] > counts[0]            # the front has refined the forest
        assert_same(dts, ora.dts[1::2])
        assert dev.pool.scenario(dev.sid)["retakes"] == ora.retakes
        q, _ = dev.leaves_state()
        got = np.moveaxis(q[:, 2:-2, 2:-2], -1, 0)
        ref = np.moveaxis(ora.state(), 1, 0)
        assert_same(got, ref)
    finally:
        dev.close()
        gf.set_params()
        oracle_c.set_params(Params())


def test_tune_cache_persists_choices(gf, tmp_path):
    """A run with a tuning cache re-selects the variant as the population
    grows and writes its choices; a second run with the same file finds
    every population bucket it meets in the cache (after the first, which
    it has met before) and times nothing."""
    print("case tune cache persistence", flush=True)
    gf.set_params(manning_n=0.035)
    path = str(tmp_path / "variants.json")
    try:
        logs = []
        for rep in range(2):
            cache = TuneCache(path)
            args, kw = _args(16)
            run = AmrRun(*args, (args[1] << 3) ** 2, tune_cache=cache, **kw)
            try:
                run.start(0.0, 0.01, state_fn=reservoir_state(run))
                run.run(12)
            finally:
                run.close()
            logs.append(cache.log)
        assert logs[0] and any(not e["cached"] for e in logs[0])
        assert logs[1] and all(e["cached"] for e in logs[1])
        assert [e["variant"] for e in logs[1]] == [e["variant"] for e in logs[0]][:len(logs[1])]
    finally:
        gf.set_params()
