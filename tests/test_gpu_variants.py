"""Step variants (tile shape, threads per tile, patches per workgroup, LDS
split) and the autotuner: every variant produces bit-identical states and dt
sequences (a tile with its own 2-cell halo computes exactly what the
whole-patch step computes; patches stepped side by side in one workgroup,
or with their x-face increments in registers, compute what they compute
alone), 64x64 patches (tiled) match the oracle, and autotuning advances no
scenario."""
import numpy as np
import pytest

from oracle.oracle import Brick, Params
from geoflood_amd.pool import BrickLayout, Pool
from geoflood_amd.scenarios import uniform_dam_break
from _parity import assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture
def dam_params(gf, oracle_c):
    oracle_c.set_params(Params(manning_n=0.03))
    gf.set_params(manning_n=0.03)
    yield
    oracle_c.set_params(Params())
    gf.set_params()


def _pool(n, mx, seed=2):
    h, hu, hv, b, dx = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=seed)
    layout = BrickLayout(n // mx, n // mx, mx, mx, dx, dx)
    pool = Pool(mx, mx, layout.npatch)
    pool.set_layout([layout])
    pool.upload_global(layout, h, hu, hv, b)
    return pool, layout, (h, hu, hv, b, dx)


def _run(pool, layout, steps, graph=True):
    sid = pool.add_scenario(layout, dt0=0.01, target_steps=steps)
    dts, prev = [], 0
    while True:
        act = pool.run(1, graph=graph)
        info = pool.scenario(sid)
        if info["nsteps"] > prev:
            dts.append(info["last_dt"])
            prev = info["nsteps"]
        if not act:
            break
    return info, np.array(dts), np.stack(pool.download_global(layout, buf=info["cur"]))


@pytest.mark.parametrize("mx", [8, 16, 32, 64])
def test_variants_bit_identical(gf, dam_params, mx):
    n, steps = 256, 25
    pool, layout, _ = _pool(n, mx)
    _, variants = pool.variants()
    pool.close()
    out = []
    for v in range(len(variants)):
        print(f"case mx={mx} variant {v} {variants[v]}", flush=True)
        pool, layout, _ = _pool(n, mx)
        pool.set_variant(v)
        out.append(_run(pool, layout, steps, graph=(v % 2 == 0)))
        pool.close()
    i0, d0, s0 = out[0]
    for info, d, st in out[1:]:
        assert info["nsteps"] == i0["nsteps"] and info["retakes"] == i0["retakes"]
        assert np.array_equal(d, d0)
        assert np.array_equal(st, s0)


def test_plain_tiled_equals_batched(gf, dam_params):
    """64x64 patches through the per-patch launch loop (tiled k_step, host
    dt rule) and the batched graph path give identical results."""
    res = []
    for mode in ("plain", "graph"):
        pool, layout, _ = _pool(256, 64)
        sid = pool.add_scenario(layout, dt0=0.01, target_steps=10)
        while pool.run_plain(5) if mode == "plain" else pool.run(5):
            pass
        info = pool.scenario(sid)
        res.append((info["t"], np.stack(pool.download_global(layout, buf=info["cur"]))))
        pool.close()
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])


def test_64x64_patches_match_oracle(gf, oracle_c, dam_params):
    n, steps = 256, 40
    pool, layout, (h, hu, hv, b, dx) = _pool(n, 64, seed=4)
    info, dts, got = _run(pool, layout, steps)
    pool.close()
    br = Brick.from_global(h, hu, hv, b, n // 64, n // 64, 2, dx, dx)
    rec = oracle_c.brick_run(br, steps, 0.01, 0.9, 1.0, nthreads=8)
    assert_same(dts, rec["dt"])
    assert_same(got, br.to_global())


def test_autotune_does_not_advance(gf, dam_params):
    pool, layout, _ = _pool(512, 32)
    sid = pool.add_scenario(layout, dt0=0.01, target_steps=6)
    pool.run(3)
    before = pool.scenario(sid)
    st0 = np.stack(pool.download_global(layout, buf=before["cur"]))
    best, ms = pool.autotune(reps=2)
    after = pool.scenario(sid)
    assert before == after
    assert np.array_equal(st0, np.stack(pool.download_global(layout, buf=after["cur"])))
    assert pool.variants()[0] == best and len(ms) == len(pool.variants()[1]) and min(ms) > 0
    # and the run continues exactly as an untuned one
    while pool.run(2):
        pass
    t_tuned = pool.scenario(sid)["t"]
    pool.close()
    pool, layout, _ = _pool(512, 32)
    sid = pool.add_scenario(layout, dt0=0.01, target_steps=6)
    while pool.run(2):
        pass
    assert pool.scenario(sid)["t"] == t_tuned
    pool.close()
