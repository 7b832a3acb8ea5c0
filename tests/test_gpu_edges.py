"""Edge cases of the C ABI with an initialised library on the device.

Refusals are status codes with a message, raised before any launch:
- empty and malformed batches;
- pools of unsupported shape or size;
- parameters outside the supported set, which leave the active ones unchanged;
- runs without a scenario;
- slots out of range;
- an autotune with nothing to time (every scenario stopped).

The smallest pool, one 16x16 patch with walls on every side, and a batch of
ragged cell sizes both step exactly as the oracle does (fc2d_cpucuda.cpp:409-495,
cudaclaw_step2.cu:57-324 for the batch; fc2d_geoclaw_run.c:417-590 for the run)."""
import ctypes as ct

import numpy as np
import pytest

import geoflood_amd as G
from oracle.oracle import Brick, Params
from geoflood_amd.pool import BrickLayout, Pool
from _parity import assert_same

pytestmark = pytest.mark.gpu

GF_EINVAL, GF_EUNSUPPORTED, GF_ESTATE = -1, -4, -5


def _ptrs(arrays):
    return (ct.POINTER(ct.c_double) * len(arrays))(*[a.ctypes.data_as(ct.POINTER(ct.c_double)) for a in arrays])


def test_refused_batches(gf):
    print("case edges: refused batches", flush=True)
    lib = gf.load()
    q = [np.zeros((20, 20, 3))]
    a = [np.zeros((20, 20, 1))]
    d = np.array([0.1])
    m = ct.c_double(-1.0)
    dp = d.ctypes.data_as(ct.POINTER(ct.c_double))
    for fn in (lib.gf_step2_batch, lib.gf_step2_batch_noalloc):
        assert fn(16, 16, 2, 1, 0, _ptrs(q), _ptrs(a), dp, dp, 0.0, 0.01, ct.byref(m)) == GF_EINVAL
        assert b"batch size" in lib.gf_last_error()
        assert fn(16, 16, 2, 1, 1, None, _ptrs(a), dp, dp, 0.0, 0.01, ct.byref(m)) == GF_EINVAL
        assert fn(16, 16, 2, 1, 1, _ptrs(q), _ptrs(a), dp, dp, 0.0, 0.0, ct.byref(m)) == GF_EINVAL
        assert fn(16, 16, 2, 0, 1, _ptrs(q), _ptrs(a), dp, dp, 0.0, 0.01, ct.byref(m)) == GF_EINVAL
        bad = np.array([0.0])
        bp = bad.ctypes.data_as(ct.POINTER(ct.c_double))
        assert fn(16, 16, 2, 1, 1, _ptrs(q), _ptrs(a), bp, dp, 0.0, 0.01, ct.byref(m)) == GF_EINVAL
        assert b"dx, dy" in lib.gf_last_error()
        assert fn(12, 12, 2, 1, 1, _ptrs(q), _ptrs(a), dp, dp, 0.0, 0.01, ct.byref(m)) == GF_EUNSUPPORTED
        assert fn(16, 16, 3, 1, 1, _ptrs(q), _ptrs(a), dp, dp, 0.0, 0.01, ct.byref(m)) == GF_EUNSUPPORTED
    assert m.value == -1.0  # no refused call wrote a result


def test_refused_pools_and_runs(gf):
    print("case edges: refused pools and runs", flush=True)
    lib = gf.load()
    assert lib.gf_pool_create(12, 12, 2, 4) is None
    assert b"not supported" in lib.gf_last_error()
    assert lib.gf_pool_create(32, 32, 3, 4) is None
    assert lib.gf_pool_create(32, 32, 2, 0) is None
    assert b"capacity" in lib.gf_last_error()
    pool = Pool(16, 16, 4)
    try:
        act = ct.c_int(-1)
        assert lib.gf_pool_run(pool.h, 1, 1, ct.byref(act)) == GF_ESTATE
        assert lib.gf_pool_run(pool.h, -1, 1, ct.byref(act)) == GF_EINVAL
        assert act.value == -1
        q = np.zeros((1, 20, 20, 3))
        qp = q.ctypes.data_as(ct.POINTER(ct.c_double))
        assert lib.gf_pool_upload(pool.h, 4, 1, 0, qp, None, 1) == GF_EINVAL       # slot past the end
        assert lib.gf_pool_upload(pool.h, -1, 1, 0, qp, None, 1) == GF_EINVAL
        assert lib.gf_pool_upload(pool.h, 0, 1, 2, qp, None, 1) == GF_EINVAL       # no buffer 2
        # a scenario over slots that were never described
        assert lib.gf_pool_add_scenario(pool.h, 0, 2, 0.0, 0.01, 0.9, 1.0, 10, 0.0) == GF_EINVAL
        layout = BrickLayout(1, 1, 16, 16, 0.1, 0.1)
        pool.set_layout([layout])
        assert lib.gf_pool_add_scenario(pool.h, 0, 1, 0.0, 0.0, 0.9, 1.0, 10, 0.0) == GF_EINVAL  # dt0 = 0
    finally:
        pool.close()


def test_refused_params_leave_the_active_ones(gf):
    print("case edges: refused params", flush=True)
    lib = gf.load()
    G.set_params(manning_n=0.041)
    before = G.GfParams()
    assert lib.gf_get_params(ct.byref(before)) == 0
    try:
        for kw, code in ((dict(order=3), GF_EINVAL), (dict(transverse=5), GF_EINVAL),
                         (dict(mthlim=(4, 7, 4)), GF_EINVAL), (dict(grav=0.0), GF_EINVAL),
                         (dict(dry_tolerance=-1.0), GF_EINVAL)):
            p = G.make_params(**kw)
            assert lib.gf_set_params(ct.byref(p)) == code, kw
        p = G.make_params()
        p.mcapa = 1
        assert lib.gf_set_params(ct.byref(p)) == GF_EUNSUPPORTED
        after = G.GfParams()
        assert lib.gf_get_params(ct.byref(after)) == 0
        assert bytes(after) == bytes(before)
    finally:
        G.set_params()


def test_one_patch_pool_and_ragged_batch_match_oracle(gf, oracle_c):
    """The smallest pool (one patch, walls on all four sides) for 30 steps
    through the graph, and a host batch whose patches have different cell
    sizes: identical to the oracle."""
    print("case edges: one-patch pool, ragged batch", flush=True)
    gf.set_params(manning_n=0.03)
    oracle_c.set_params(Params(manning_n=0.03))
    try:
        n, dx = 16, 0.5
        x = (np.arange(n) + 0.5) * dx
        X, Y = np.meshgrid(x, x)
        b = 0.2 * np.sin(X) * np.cos(Y)
        h = np.maximum(1.0 + 0.4 * np.exp(-((X - 3) ** 2 + (Y - 4) ** 2)) - b, 0.0)
        z = np.zeros_like(h)
        br = Brick.from_global(h, z, z, b, 1, 1, 2, dx, dx)
        rec = oracle_c.brick_run(br, 30, 0.01, 0.9, 1.0, nthreads=1)
        layout = BrickLayout(1, 1, n, n, dx, dx, (3, 3, 3, 3))
        pool = Pool(n, n, 1)
        try:
            pool.set_layout([layout])
            pool.upload_global(layout, h, z, z, b)
            sid = pool.add_scenario(layout, dt0=0.01, target_steps=30)
            dts, prev = [], 0
            while True:
                act = pool.run(1, graph=True)
                info = pool.scenario(sid)
                if info["nsteps"] > prev:
                    dts.append(info["last_dt"])
                    prev = info["nsteps"]
                if not act:
                    break
            assert info["nsteps"] == 30 and info["retakes"] == rec["retakes"]
            assert_same(dts, rec["dt"])
            assert_same(np.stack(pool.download_global(layout, buf=info["cur"])), br.to_global())
        finally:
            pool.close()

        rng = np.random.default_rng(17)
        qs, auxs, ref = [], [], []
        cells = [0.05, 0.1, 0.4, 2.0]
        c_ref = 0.0
        for k, d in enumerate(cells):
            hh = 0.2 + rng.random((20, 20))
            q = np.zeros((20, 20, 3))
            q[..., 0] = hh
            q[..., 1:] = rng.normal(0, 0.3, (20, 20, 2)) * hh[..., None]
            aux = np.ascontiguousarray(rng.normal(0, 0.1, (20, 20, 1)))
            qo = q.copy()
            c_ref = max(c_ref, oracle_c.patch_update(16, 16, 2, qo, aux, d, d, 0.0, 0.004))
            qs.append(q); auxs.append(aux); ref.append(qo)
        lib = gf.load()
        dxs = np.array(cells)
        m = ct.c_double(0.0)
        dp = dxs.ctypes.data_as(ct.POINTER(ct.c_double))
        assert lib.gf_step2_batch(16, 16, 2, 1, len(qs), _ptrs(qs), _ptrs(auxs), dp, dp, 0.0, 0.004,
                                  ct.byref(m)) == 0
        assert m.value == c_ref
        for q, qo in zip(qs, ref):
            assert_same(q[2:-2, 2:-2], qo[2:-2, 2:-2])
    finally:
        gf.set_params()
        oracle_c.set_params(Params())


def test_autotune_refuses_an_empty_population(gf):
    """With its only scenario stopped (at its step target) a pool lists no
    patch: the autotuner refuses instead of timing empty launches, and keeps
    the variant; once the scenario runs again it times and selects."""
    print("case edges: autotune on a stopped scenario", flush=True)
    from geoflood_amd import GeoFloodError
    from geoflood_amd.scenarios import uniform_dam_break
    n = 128
    sc = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=9)
    layout = BrickLayout(n // 32, n // 32, 32, 32, 1.0, 1.0)
    pool = Pool(32, 32, layout.npatch)
    try:
        pool.set_layout([layout])
        pool.upload_global(layout, *sc[:4])
        sid = pool.add_scenario(layout, dt0=0.01, target_steps=2)
        while pool.run(2, graph=True):
            pass
        assert not pool.scenario(sid)["active"]
        pool.set_variant(1)
        with pytest.raises(GeoFloodError, match="no active or dry patch"):
            pool.autotune(reps=1)
        assert pool.variants()[0] == 1
        info = pool.scenario(sid)
        pool.set_stop(sid, info["nsteps"] + 2)
        best, ms = pool.autotune(reps=1)
        assert pool.variants()[0] == best and min(ms) > 0
        assert pool.scenario(sid)["nsteps"] == info["nsteps"]      # timing advances nothing
    finally:
        pool.close()
