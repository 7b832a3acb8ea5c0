"""Continuous batching of flood scenarios on one device pool: each result is
the same alone and in the mix, admission follows the HBM footprint, and a
served scenario matches the oracle's run of the same forest."""
import numpy as np
import pytest

from oracle.oracle import Brick, Hydrograph, Params
from geoflood_amd.serving import DT0, DT_MAX, LENGTH, MX, STRIP, T_START, Engine, Request
from geoflood_amd.scenarios import gaussian_sum_dem, sine_hydrograph
from _parity import assert_same

pytestmark = pytest.mark.gpu

T_SHORT = 250.0


def _reqs():
    return [Request(0, 128, 0, 0.03, 400.0, 0.0, T_SHORT),
            Request(1, 256, 1, 0.05, 1500.0, 8.0, T_SHORT),
            Request(2, 128, 2, 0.02, 60.0, 16.0, T_SHORT),
            Request(3, 256, 3, 0.04, 900.0, 24.0, T_SHORT)]


def test_alone_equals_in_mix(gf):
    """Every scenario's final state, step count and retakes are identical
    when it runs alone and when it runs with others admitted mid-flight
    (arrivals in attempt units: deterministic schedule)."""
    alone = {}
    for r in _reqs():
        print(f"case alone rid={r.rid} n={r.n}", flush=True)
        eng = Engine(capacity=64, substeps=4, keep_states=True)
        res = eng.serve([Request(r.rid, r.n, r.seed, r.manning, r.peak_q, 0.0, r.t_end)],
                        clock=lambda a: float(a))
        alone[r.rid] = (res[0], eng.states[r.rid])
        eng.close()
    print("case mix of 4", flush=True)
    eng = Engine(capacity=200, substeps=4, keep_states=True)
    res = eng.serve(_reqs(), clock=lambda a: float(a))
    assert len(res) == 4
    for r in res:
        a, st = alone[r.rid]
        assert (r.steps, r.retakes, r.t) == (a.steps, a.retakes, a.t)
        assert np.array_equal(eng.states[r.rid], st), r.rid
        assert r.wet_cells > 0
    # the mix really overlapped: request 1 was admitted before request 0 finished
    by = {r.rid: r for r in res}
    assert by[1].admitted < by[0].finished
    eng.close()


def test_admission_by_footprint(gf):
    """A pool of 80 slots holds one 256^2 scenario (64 slots) and one 128^2
    (16 slots): the second 256^2 request waits for the first to finish,
    and the 128^2 request behind it is admitted at once (backfill)."""
    reqs = [Request(0, 256, 5, 0.03, 300.0, 0.0, 120.0),
            Request(1, 256, 6, 0.03, 300.0, 1.0, 120.0),
            Request(2, 128, 7, 0.03, 300.0, 2.0, 120.0)]
    eng = Engine(capacity=80, substeps=4)
    res = {r.rid: r for r in eng.serve(reqs, clock=lambda a: float(a))}
    assert res[1].admitted >= res[0].finished
    assert res[2].admitted < res[0].finished
    eng.close()


def test_served_scenario_matches_oracle(gf, oracle_c):
    """A served 128^2 scenario (its own Manning n and hydrograph) against
    the oracle's outstyle_3 run of the same forest for the same number of
    accepted steps."""
    r = Request(0, 128, 11, 0.045, 800.0, 0.0, T_SHORT)
    eng = Engine(capacity=16, substeps=4, keep_states=True)
    res = eng.serve([r], clock=lambda a: float(a))[0]
    got = eng.states[0]
    eng.close()
    assert 20 < res.steps <= 150
    b = gaussian_sum_dem(r.n, LENGTH, seed=r.seed)
    z = np.zeros_like(b)
    t, hu = sine_hydrograph(r.peak_q, STRIP)
    H = Hydrograph(t, hu, STRIP, 0.0, 0.5 * LENGTH, (1, 0, 0, 0))
    dx = LENGTH / r.n
    br = Brick.from_global(z, z, z, b, r.n // MX, r.n // MX, 2, dx, dx, mthbc=(0, 1, 3, 3), hydro=H)
    oracle_c.set_params(Params(manning_n=r.manning))
    try:
        rec = oracle_c.brick_run(br, res.steps, DT0, 0.9, 1.0, t0=T_START, nthreads=8, dt_max=DT_MAX)
    finally:
        oracle_c.set_params(Params())
    assert rec["retakes"] == res.retakes
    assert abs((T_START + rec["dt"].sum()) - res.t) <= 1e-9 * res.t
    assert_same(got, br.to_global())
