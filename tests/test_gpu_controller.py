"""The device time-step controller on failure and at rest.

* A NaN written into a scenario's state stops that scenario with status
  GF_SCEN_NONFINITE at the next attempt (no step accepted: t, nsteps and the
  accepted buffer stay), whether the NaN lies in a wet patch (the step
  kernel's new state is checked) or in a steady dry one (the dry stream
  checks it), in the batched graph path and in the per-patch loop.  A second
  scenario in the same pool runs on, bit for bit as it runs alone.  The
  reference has no such check: a NaN CFL number fails "maxcfl > max_cfl"
  (fc2d_geoclaw_run.c:529) and the step is accepted.
* A dry start whose hydrograph is at zero discharge (CFL number 0) advances:
  with a dt bound the next dt is dt_max instead of the infinite dt of the
  reference's rule, and the run equals the oracle's loop with the same bound.
"""
import numpy as np
import pytest

from geoflood_amd.pool import BrickLayout, Pool
from oracle.oracle import Brick, Hydrograph, Params
from _parity import assert_same

pytestmark = pytest.mark.gpu

GF_SCEN_OK, GF_SCEN_NONFINITE = 0, 1


def _two_scenarios(n=256):
    from geoflood_amd.scenarios import uniform_dam_break
    a = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=3)
    b = uniform_dam_break(n, dx=1.0, h0=6.0, rms=2.0, seed=4)
    la = BrickLayout(n // 32, n // 32, 32, 32, 1.0, 1.0)
    lb = BrickLayout(n // 32, n // 32, 32, 32, 1.0, 1.0)
    return (a, la), (b, lb)


def _stepper(pool, mode):
    return (lambda k: pool.run(k, graph=True)) if mode == "graph" else (lambda k: pool.run_plain(k))


@pytest.mark.parametrize("mode", ["graph", "plain"])
@pytest.mark.parametrize("where", ["wet", "steady-dry"])
def test_nan_state_stops_its_scenario_only(gf, mode, where):
    print(f"case NaN in a {where} patch, {mode}", flush=True)
    (sa, la), (sb, lb) = _two_scenarios()
    # scenario B alone, 8 accepted steps
    solo = Pool(32, 32, lb.npatch)
    lb1 = BrickLayout(lb.mi, lb.mj, 32, 32, 1.0, 1.0)
    solo.set_layout([lb1])
    solo.upload_global(lb1, *sb[:4])
    sid = solo.add_scenario(lb1, dt0=0.01, target_steps=8)
    run = _stepper(solo, mode)
    while run(4):
        pass
    ib = solo.scenario(sid)
    ref_b = np.stack(solo.download_global(lb1, buf=ib["cur"]))
    solo.close()

    pool = Pool(32, 32, la.npatch + lb.npatch)
    try:
        pool.set_layout([la, lb])
        pool.upload_global(la, *sa[:4])
        pool.upload_global(lb, *sb[:4])
        A = pool.add_scenario(la, dt0=0.01, target_steps=100)
        B = pool.add_scenario(lb, dt0=0.01, target_steps=8)
        run = _stepper(pool, mode)
        while pool.scenario(A)["nsteps"] < 4:
            run(1)
        before = pool.scenario(A)
        assert before["status"] == GF_SCEN_OK
        st = np.stack(pool.download_global(la, buf=before["cur"]))
        if where == "wet":
            j, i = np.argwhere(st[0] > 1.0)[0]
        else:
            # far from the reservoir (upper-left eighth): dry and steady by now
            j, i = la.mj * 32 - 10, la.mi * 32 - 10
            assert st[0][j, i] < 1e-3
            if mode == "graph":
                assert pool.stats()["steady_patches"] > 0
        st[0][j, i] = np.nan
        pool.upload_global(la, st[0], st[1], st[2], sa[3], buf=before["cur"])
        while run(2):
            pass
        ia, ib2 = pool.scenario(A), pool.scenario(B)
        assert ia["status"] == GF_SCEN_NONFINITE and not ia["active"]
        assert (ia["nsteps"], ia["t"], ia["cur"]) == (before["nsteps"], before["t"], before["cur"])
        assert np.isnan(ia["maxcfl"])
        assert ib2["status"] == GF_SCEN_OK and ib2["nsteps"] == 8
        assert (ib2["t"], ib2["retakes"]) == (ib["t"], ib["retakes"])
        got_b = np.stack(pool.download_global(lb, buf=ib2["cur"]))
        assert np.array_equal(got_b, ref_b)
        # the accepted state of A is the one before the failed attempt
        kept = np.stack(pool.download_global(la, buf=ia["cur"]))
        assert np.isnan(kept[0][j, i]) and np.isfinite(np.delete(kept[0].ravel(), j * la.mi * 32 + i)).all()
    finally:
        pool.close()


@pytest.mark.parametrize("mode", ["graph", "plain"])
def test_dry_start_zero_discharge_advances(gf, oracle_c, mode):
    """128^2 overland flood (Gaussian-sum DEM, west hydrograph inflow at
    Q(t) = 400 sin(pi t / 3600) m^3/s) from a dry bed at t = 0: the first
    attempt has no wave (CFL 0) and is accepted with dt0, the next dt is
    dt_max, and 30 accepted steps equal the oracle's loop with the same
    bound (identical dt sequence and cell values)."""
    from geoflood_amd.scenarios import overland_flood
    print(f"case dry start at t=0, {mode}", flush=True)
    n, steps, dt0, dt_max = 128, 30, 1.0, 5.0
    sc = overland_flood(n, peak_q=400.0, seed=8)
    hy = sc["hydro"]
    layout = BrickLayout(n // 32, n // 32, 32, 32, sc["dx"], sc["dx"], mthbc=sc["mthbc"])
    pool = Pool(32, 32, layout.npatch)
    gf.set_params(manning_n=0.035)
    oracle_c.set_params(Params(manning_n=0.035))
    try:
        pool.set_layout([layout])
        pool.upload_global(layout, sc["h"], sc["hu"], sc["hv"], sc["b"])
        sid = pool.add_scenario(layout, dt0=dt0, target_steps=steps)
        pool.set_inflow(sid, hy["time"], hy["hu"], hy["b"], hy["x0"], hy["y0"])
        pool.set_dt_max(sid, dt_max)
        run = _stepper(pool, mode)
        run(1)
        first = pool.scenario(sid)
        assert first["maxcfl"] == 0.0 and first["nsteps"] == 1
        assert first["t"] == dt0 and first["dt"] == dt_max
        dts = [first["last_dt"]]
        prev = 1
        while True:
            active = run(1)
            info = pool.scenario(sid)
            if info["nsteps"] > prev:
                dts.append(info["last_dt"]); prev = info["nsteps"]
            if not active:
                break
        assert info["status"] == GF_SCEN_OK and info["nsteps"] == steps
        assert max(dts) <= dt_max and np.isfinite(info["dt"])
        br = Brick.from_global(sc["h"], sc["hu"], sc["hv"], sc["b"], n // 32, n // 32, 2, sc["dx"], sc["dx"],
                               mthbc=sc["mthbc"],
                               hydro=Hydrograph(hy["time"], hy["hu"], hy["b"], hy["x0"], hy["y0"], (1, 0, 0, 0)))
        rec = oracle_c.brick_run(br, steps, dt0, 0.9, 1.0, nthreads=8, dt_max=dt_max)
        assert rec["retakes"] == info["retakes"]
        assert_same(dts, rec["dt"])
        got = np.stack(pool.download_global(layout, buf=info["cur"]))
        ref = br.to_global()
        assert (ref[0] > 1e-3).sum() > 10          # water has entered
        assert_same(got, ref)
    finally:
        gf.set_params()
        oracle_c.set_params(Params())
        pool.close()


def test_advance_equals_stop_and_replay(gf):
    """gf_pool_advance to a step count equals setting the stop there and
    replaying attempts until the scenario stops (retakes included: the first
    dt is five times too large), in every counter, the dt and the state; it
    refuses a bound that the missing steps exceed and does nothing at a
    target already reached."""
    from geoflood_amd import GeoFloodError
    from geoflood_amd.scenarios import uniform_dam_break
    print("case advance vs stop-and-replay", flush=True)
    n = 256
    sc = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=5)
    out = []
    for mode in ("loop", "advance"):
        lay = BrickLayout(n // 32, n // 32, 32, 32, 1.0, 1.0)
        pool = Pool(32, 32, lay.npatch)
        try:
            pool.set_layout([lay])
            pool.upload_global(lay, *sc[:4])
            sid = pool.add_scenario(lay, dt0=0.5, target_steps=0)
            recs = []
            for target in (3, 7, 12):
                if mode == "loop":
                    pool.set_stop(sid, target)
                    k = target - pool.scenario(sid)["nsteps"]
                    while pool.run(k):
                        k = 1
                    recs.append(pool.scenario(sid))
                else:
                    recs.append(pool.advance(sid, target))
            if mode == "advance":
                again = pool.advance(sid, 12)
                assert again == recs[-1]                      # target reached: no attempt
                with pytest.raises(GeoFloodError, match="max_attempts"):
                    pool.advance(sid, 20, max_attempts=3)
                assert pool.scenario(sid)["nsteps"] == 12     # refused before any attempt
            out.append((recs, np.stack(pool.download_global(lay, buf=recs[-1]["cur"]))))
        finally:
            pool.close()
    (ra, sa), (rb, sb) = out
    assert ra[0]["retakes"] > 0
    for a, b in zip(ra, rb):
        assert a == b
    assert_same(sb, sa)
