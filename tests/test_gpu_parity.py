"""Parity of the HIP patch step (through the C ABI) with the CPU oracle.

Bar: identical values.  The kernels form every FP64 value with the
reference's operations in the reference's order (no FMA contraction) and
evaluate h**(7/3) correctly rounded, as the oracle does in its pow mode, so
single steps, CFL maxima, dt sequences and every cell of multi-step runs
equal the oracle's; closed-domain mass is conserved to 1e-12.
"""
import numpy as np
import pytest

from oracle.oracle import Brick, Params
from geoflood_amd.pool import BrickLayout, Pool
from _parity import assert_same
from geoflood_amd.scenarios import radial_dam_break

pytestmark = pytest.mark.gpu



def random_patch(rng, mx=32, my=32, mbc=2, wet_frac=0.7, topo=0.5, hscale=1.0):
    h = rng.random((my + 2 * mbc, mx + 2 * mbc)) * hscale
    h[rng.random(h.shape) > wet_frac] = 0.0
    q = np.zeros(h.shape + (3,))
    q[..., 0] = h
    q[..., 1] = rng.normal(0, 1, h.shape) * h
    q[..., 2] = rng.normal(0, 1, h.shape) * h
    aux = np.ascontiguousarray((rng.normal(0, topo, h.shape) if topo else np.zeros(h.shape))[..., None])
    return q, aux


@pytest.mark.parametrize("seed", range(6))
def test_single_patch_random_states(gf, oracle_c, seed):
    """Random wet/dry patches over random bathymetry, ghosts given: the
    drop-in batch step against the oracle's geoclaw_update."""
    rng = np.random.default_rng(seed)
    print(f"case seed={seed} 8 patches 32x32", flush=True)
    qs, auxs, ref = [], [], []
    cfl_ref = 0.0
    for k in range(8):
        q, aux = random_patch(rng, wet_frac=[1.0, 0.7, 0.3, 0.05][k % 4], topo=[0, 0.5, 2.0][k % 3],
                              hscale=[0.01, 1, 5][k % 3])
        qo = q.copy()
        cfl_ref = max(cfl_ref, oracle_c.patch_update(32, 32, 2, qo, aux, 0.1, 0.1, 0.0, 0.01))
        qs.append(q); auxs.append(aux); ref.append(qo)
    cfl = gf.step2_batch(qs, auxs, 0.1, 0.1, 0.0, 0.01)
    assert cfl == cfl_ref
    for q, qo in zip(qs, ref):
        a, b = q[2:-2, 2:-2], qo[2:-2, 2:-2]
        assert_same(a, b)
        assert np.array_equal(a[..., 0] > 1e-3, b[..., 0] > 1e-3)


def test_all_dry_patch(gf, oracle_c):
    """Dry tile (every cell h <= drytol): b4step2 + src2 only, CFL 0."""
    rng = np.random.default_rng(7)
    q, aux = random_patch(rng, wet_frac=1.0, hscale=1e-3)
    q[..., 0] = np.minimum(q[..., 0], 1e-3)
    q[3, 4, 0] = 1e-3  # exactly at the tolerance: dry for rpn2, untouched by b4step2
    q[3, 4, 1] = 0.2
    qo = q.copy()
    c_ref = oracle_c.patch_update(32, 32, 2, qo, aux, 0.1, 0.1, 0.0, 0.01)
    c = gf.step2_batch([q], [aux], 0.1, 0.1, 0.0, 0.01)
    assert c == 0.0 and c_ref == 0.0
    np.testing.assert_array_equal(q[2:-2, 2:-2], qo[2:-2, 2:-2])


@pytest.mark.parametrize("mx", [8, 16])
def test_small_patch_shapes(gf, oracle_c, mx):
    rng = np.random.default_rng(mx)
    q, aux = random_patch(rng, mx=mx, my=mx, wet_frac=0.6, topo=0.5)
    qo = q.copy()
    c_ref = oracle_c.patch_update(mx, mx, 2, qo, aux, 0.05, 0.07, 0.0, 0.004)
    c = gf.step2_batch([q], [aux], 0.05, 0.07, 0.0, 0.004)
    assert c == c_ref
    assert_same(q[2:-2, 2:-2], qo[2:-2, 2:-2])


@pytest.mark.parametrize("maux", [3, 5])
def test_single_patch_several_aux_fields(gf, oracle_c, maux):
    """One patch with maux > 1 (bathymetry first, the other fields unused by
    the step): the one-patch pool's staging grows to hold (3 + maux) fields
    for two patches, and the step equals the oracle's."""
    rng = np.random.default_rng(40 + maux)
    q, a1 = random_patch(rng, mx=16, my=16, wet_frac=0.8, topo=0.5)
    aux = np.ascontiguousarray(np.concatenate([a1, rng.random(a1.shape[:2] + (maux - 1,))], axis=-1))
    qo, ao = q.copy(), aux.copy()
    c_ref = oracle_c.patch_update(16, 16, 2, qo, ao, 0.05, 0.05, 0.0, 0.004, maux=maux)
    c = gf.step2_batch([q], [aux], 0.05, 0.05, 0.0, 0.004)
    assert c == c_ref
    assert_same(q[2:-2, 2:-2], qo[2:-2, 2:-2])
    np.testing.assert_array_equal(aux, ao)


def test_host_batch_chunked_staging(gf, oracle_c):
    """1,500 8x8 patches with maux = 3 pass through the pinned staging in
    several chunks on both halves of the double-buffered upload and download;
    two calls of one HostBatch equal two oracle steps per patch, and the
    extra aux fields are left as they were.  Between the calls the host puts
    the initial ghost rings back on both sides, as ForestClaw refills ghosts
    before each step (the reference's step also writes partial updates into
    the ghost ring, the batch step leaves it as it was)."""
    from geoflood_amd import HostBatch
    rng = np.random.default_rng(1500)
    n, mx = 1500, 8
    print(f"case {n} patches {mx}x{mx} maux=3, 2 calls", flush=True)
    qs, auxs, ref = [], [], []
    for k in range(n):
        q, a1 = random_patch(rng, mx=mx, my=mx, wet_frac=[1.0, 0.6, 0.0][k % 3], topo=0.5)
        aux = np.ascontiguousarray(np.concatenate([a1, rng.random(a1.shape[:2] + (2,))], axis=-1))
        qs.append(q); auxs.append(aux); ref.append((q.copy(), aux.copy()))
    ghosts = [q.copy() for q in qs]
    ring = np.ones(qs[0].shape[:2], dtype=bool)
    ring[2:-2, 2:-2] = False
    hb = HostBatch(qs, auxs, 0.05, 0.05)
    for dt in (0.004, 0.003):
        for q, (qo, _), q0 in zip(qs, ref, ghosts):
            q[ring] = q0[ring]
            qo[ring] = q0[ring]
        c = hb.step(0.0, dt)
        c_ref = 0.0
        for qo, ao in ref:
            c_ref = max(c_ref, oracle_c.patch_update(mx, mx, 2, qo, ao, 0.05, 0.05, 0.0, dt, maux=3))
        assert c == c_ref
    for k, (q, aux) in enumerate(zip(qs, auxs)):
        qo, ao = ref[k]
        # two steps, identical values
        assert_same(q[2:-2, 2:-2], qo[2:-2, 2:-2], k)
        np.testing.assert_array_equal(aux, ao)


def _config1(gf):
    h, hu, hv, b, dx = radial_dam_break(256)
    layout = BrickLayout(8, 8, 32, 32, dx, dx)
    pool = Pool(32, 32, 64)
    pool.set_layout([layout])
    pool.upload_global(layout, h, hu, hv, b)
    return pool, layout, (h, hu, hv, b, dx)


def test_config1_one_step_ghost_fill(gf, oracle_c):
    """Device ghost exchange + one batched step against the oracle brick
    (ForestClaw-order face/corner copies and bc2 walls on the host)."""
    pool, layout, (h, hu, hv, b, dx) = _config1(gf)
    br = Brick.from_global(h, hu, hv, b, 8, 8, 2, dx, dx)
    c_ref = oracle_c.brick_advance(br, 0.0, 5e-4)
    pool.ghost_fill(buf=0)
    c, _ = pool.step(5e-4, src=0, dst=1)
    got = np.stack(pool.download_global(layout, buf=1))
    assert c == c_ref
    assert_same(got, br.to_global())


def test_config1_ghost_cells_match_forestclaw_order(gf, oracle_c):
    """Every ghost cell the device fills equals the host ForestClaw-order fill."""
    rng = np.random.default_rng(3)
    h = 1 + rng.random((96, 64)); hu = rng.normal(size=h.shape); hv = rng.normal(size=h.shape)
    b = rng.normal(size=h.shape)
    for mthbc in [(3, 3, 3, 3), (1, 3, 1, 3), (1, 1, 1, 1)]:
        layout = BrickLayout(2, 3, 32, 32, 1.0, 1.0, mthbc)
        pool = Pool(32, 32, 6)
        pool.set_layout([layout])
        pool.upload_global(layout, h, hu, hv, b)
        pool.ghost_fill(buf=0)
        got = pool.download_aos(0, 6, 0)
        br = Brick.from_global(h, hu, hv, b, 2, 3, 2, 1.0, 1.0, mthbc)
        oracle_c.brick_ghost_fill(br)
        for k in range(6):
            np.testing.assert_array_equal(got[k], br.q[k])
        pool.close()


def _run_oracle_config1(oracle_c, nsteps=100):
    h, hu, hv, b, dx = radial_dam_break(256)
    br = Brick.from_global(h, hu, hv, b, 8, 8, 2, dx, dx)
    rec = oracle_c.brick_run(br, nsteps, 1e-4, desired_cfl=0.9, max_cfl=1.0, nthreads=4)
    return br, rec


@pytest.mark.parametrize("graph", [True, False])
def test_config1_100_steps(gf, oracle_c, graph):
    """Radial dam break, 8x8 patches of 32x32, 100 accepted steps with the
    device-side dt controller (retakes included)."""
    print(f"case config1 100 steps graph={graph}", flush=True)
    br, rec = _run_oracle_config1(oracle_c)
    pool, layout, _ = _config1(gf)
    sid = pool.add_scenario(layout, dt0=1e-4, desired_cfl=0.9, max_cfl=1.0, target_steps=100)
    dts, cfls = [], []
    prev = 0
    while True:
        active = pool.run(1, graph=graph)
        info = pool.scenario(sid)
        if info["nsteps"] > prev:
            dts.append(info["last_dt"]); cfls.append(info["maxcfl"]); prev = info["nsteps"]
        if not active:
            break
    assert info["nsteps"] == 100
    assert info["retakes"] == rec["retakes"]
    assert_same(dts, rec["dt"])
    assert_same(cfls, rec["cfl"])
    got = np.stack(pool.download_global(layout, buf=info["cur"]))
    ref = br.to_global()
    assert_same(got, ref)
    assert np.array_equal(got[0] > 1e-3, ref[0] > 1e-3)
    mass0 = radial_dam_break(256)[0].sum()
    assert abs(got[0].sum() - mass0) / mass0 <= 1e-12


def test_plain_equals_batched(gf):
    """The per-patch launch loop and the batched graph path give bitwise
    identical states and dt sequences."""
    out = []
    for mode in ("plain", "graph"):
        pool, layout, _ = _config1(gf)
        sid = pool.add_scenario(layout, dt0=1e-4, target_steps=12)
        while (pool.run_plain(4) if mode == "plain" else pool.run(4, graph=True)):
            pass
        info = pool.scenario(sid)
        out.append((info, np.stack(pool.download_global(layout, buf=info["cur"]))))
        pool.close()
    (i1, s1), (i2, s2) = out
    assert i1["nsteps"] == i2["nsteps"] == 12 and i1["t"] == i2["t"]
    assert np.array_equal(s1, s2)


def test_dry_bed_dam_break_batched_split(gf, oracle_c):
    """Dam break onto a dry fractal DEM (the 4096^2 workload at 512^2): the
    batched path classifies patches into active and dry lists every attempt;
    its states, dt sequence and wet/dry mask must match the oracle."""
    from geoflood_amd.scenarios import uniform_dam_break
    n, steps = 512, 60
    h, hu, hv, b, dx = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=1)
    oracle_c.set_params(Params(manning_n=0.03))
    gf.set_params(manning_n=0.03)
    try:
        print(f"case dry-bed dam break {n}^2 {steps} steps", flush=True)
        br = Brick.from_global(h, hu, hv, b, n // 32, n // 32, 2, dx, dx)
        rec = oracle_c.brick_run(br, steps, 0.01, 0.9, 1.0, nthreads=8)
        layout = BrickLayout(n // 32, n // 32, 32, 32, dx, dx)
        pool = Pool(32, 32, layout.npatch)
        pool.set_layout([layout])
        pool.upload_global(layout, h, hu, hv, b)
        sid = pool.add_scenario(layout, dt0=0.01, target_steps=steps)
        dts, prev = [], 0
        while True:
            active = pool.run(1, graph=True)
            info = pool.scenario(sid)
            if info["nsteps"] > prev:
                dts.append(info["last_dt"]); prev = info["nsteps"]
            if not active:
                break
        st = pool.stats()
        assert st["dry_patches"] > 0 and st["active_patches"] > 0
        assert 0 < st["steady_patches"] < st["dry_patches"]
        assert_same(dts, rec["dt"])
        got = np.stack(pool.download_global(layout, buf=info["cur"]))
        ref = br.to_global()
        assert_same(got, ref)
    finally:
        oracle_c.set_params(Params())
        gf.set_params()


def test_steady_dry_patches_equal_plain(gf):
    """Dry patches that stay below dry_tolerance for two steps are left
    untouched by the batched path.  A thin film (h < dry_tolerance) with
    momentum, written into such patches between steps, must still be
    stepped (momentum zeroed) and the run must equal the plain loop, which
    steps every patch, bit for bit."""
    from geoflood_amd.scenarios import uniform_dam_break
    n = 256
    h, hu, hv, b, dx = uniform_dam_break(n, dx=1.0, h0=10.0, rms=3.0, seed=2)
    film = np.zeros_like(h)
    film[: n // 2, 3 * n // 4:] = 2e-4
    out = []
    for mode in ("plain", "graph"):
        print(f"case steady dry {mode}", flush=True)
        layout = BrickLayout(n // 32, n // 32, 32, 32, dx, dx)
        pool = Pool(32, 32, layout.npatch)
        pool.set_layout([layout])
        pool.upload_global(layout, h, hu, hv, b)
        sid = pool.add_scenario(layout, dt0=0.01, target_steps=16)
        step = (lambda: pool.run_plain(1)) if mode == "plain" else (lambda: pool.run(1, graph=True))
        while pool.scenario(sid)["nsteps"] < 8:
            step()
        info = pool.scenario(sid)
        if mode == "graph":
            assert pool.stats()["steady_patches"] > 0
        cur = np.stack(pool.download_global(layout, buf=info["cur"]))
        cur[0] += film
        cur[1] += 5.0 * film
        cur[2] -= 3.0 * film
        pool.upload_global(layout, cur[0], cur[1], cur[2], b, buf=info["cur"])
        while step():
            pass
        info = pool.scenario(sid)
        assert info["nsteps"] == 16
        got = np.stack(pool.download_global(layout, buf=info["cur"]))
        out.append((info["t"], got))
        pool.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    g = out[1][1]
    wet_front = g[0] >= 1e-3
    assert not np.any(g[1][: n // 2, 3 * n // 4:][~wet_front[: n // 2, 3 * n // 4:]])


def test_plugin_single_step_update_protocol(gf, oracle_c):
    """The single_step_update slot over a host brick whose ghosts the host
    fills in ForestClaw order: flushes every buffer_len patches and at the
    last one, non-flushing calls return 0, and the running maximum equals
    the oracle's level CFL; the patches equal the oracle's geoclaw_update."""
    import ctypes as ct
    rng = np.random.default_rng(21)
    h = 1 + rng.random((96, 128)); hu = rng.normal(0, 0.5, h.shape) * h
    hv = rng.normal(0, 0.5, h.shape) * h; b = rng.normal(0, 0.2, h.shape)
    br = Brick.from_global(h, hu, hv, b, 4, 3, 2, 0.1, 0.1)
    ref = br.copy()
    cref = oracle_c.brick_advance(ref, 0.0, 0.004)
    oracle_c.brick_ghost_fill(br)
    lib = gf.load()
    plug = lib.gf_plugin_create(32, 32, 2, 1, 5)
    assert plug
    total, m, flushes = 12, 0.0, 0
    for it in range(total):
        c = ct.c_double(-1)
        rc = lib.gf_plugin_single_step_update(plug, gf.dptr(br.q[it]), gf.dptr(br.aux[it]), 0.1, 0.1,
                                              it, total, 0.0, 0.004, ct.byref(c))
        assert rc == 0
        if (it + 1) % 5 == 0 or it + 1 == total:
            flushes += 1
            assert c.value > 0
        else:
            assert c.value == 0.0
        m = max(m, c.value)
    lib.gf_plugin_destroy(plug)
    assert flushes == 3
    assert m == cref
    assert_same(br.to_global(), ref.to_global())


def test_64x64_batch_step_in_tiles(gf, oracle_c):
    """64x64 patches: the default variant steps each patch as four 32x32
    tiles, each reading its 2-cell halo from the neighbouring tiles' cells.
    The drop-in batch step reads one buffer and writes the other (never in
    place), so every tile sees the state before the step: the result equals
    the oracle's geoclaw_update."""
    rng = np.random.default_rng(64)
    print("case 64x64 step2_batch, 4 tiles per patch", flush=True)
    qs, auxs, ref = [], [], []
    c_ref = 0.0
    for k in range(4):
        q, aux = random_patch(rng, mx=64, my=64, wet_frac=[1.0, 0.7, 0.4, 0.9][k], topo=0.5,
                              hscale=[1.0, 2.0, 5.0, 0.5][k])
        qo = q.copy()
        c_ref = max(c_ref, oracle_c.patch_update(64, 64, 2, qo, aux, 0.1, 0.1, 0.0, 0.006))
        qs.append(q); auxs.append(aux); ref.append(qo)
    given = [q.copy() for q in qs]
    c = gf.step2_batch(qs, auxs, 0.1, 0.1, 0.0, 0.006)
    assert c == c_ref
    ring = np.ones(qs[0].shape[:2], dtype=bool)
    ring[2:-2, 2:-2] = False
    for q, qo, q0 in zip(qs, ref, given):
        assert_same(q[2:-2, 2:-2], qo[2:-2, 2:-2])
        # the ghost ring comes back as it was given
        np.testing.assert_array_equal(q[ring], q0[ring])


def test_in_place_step_only_untiled(gf, oracle_c):
    """gf_pool_step with src == dst: a variant with one workgroup per patch
    steps in place and equals the oracle; a variant that splits the patch
    into tiles is refused (GF_EUNSUPPORTED), since its tiles would read cells
    that other tiles have already overwritten; stepping into the other
    buffer with that variant equals the oracle."""
    from geoflood_amd import GeoFloodError
    rng = np.random.default_rng(11)
    n = 4
    qs, auxs = [], []
    for k in range(n):
        q, aux = random_patch(rng, wet_frac=0.8, topo=0.5, hscale=2.0)
        qs.append(q); auxs.append(aux)
    qs, auxs = np.array(qs), np.array(auxs)
    ref = [q.copy() for q in qs]
    c_ref = max(oracle_c.patch_update(32, 32, 2, r, a, 0.1, 0.1, 0.0, 0.01) for r, a in zip(ref, auxs))
    pool = Pool(32, 32, n)
    try:
        pool.set_patches(np.full(n, 0.1), np.full(n, 0.1), np.zeros(n), np.zeros(n))
        _, vs = pool.variants()
        whole = [k for k, v in enumerate(vs) if v[:2] == (32, 32)]
        tiled = [k for k, v in enumerate(vs) if v[:2] != (32, 32)]
        assert whole and tiled
        pool.set_variant(whole[0])
        pool.upload_aos(0, qs, auxs, buf=0)
        c, _ = pool.step(0.01, src=0, dst=0)
        got = pool.download_aos(0, n, buf=0)
        assert c == c_ref
        for g, r in zip(got, ref):
            assert_same(g[2:-2, 2:-2], r[2:-2, 2:-2])
        pool.set_variant(tiled[0])
        pool.upload_aos(0, qs, auxs, buf=0)
        with pytest.raises(GeoFloodError, match=r"\(-4\)"):
            pool.step(0.01, src=0, dst=0)
        with pytest.raises(GeoFloodError, match=r"\(-4\)"):
            pool.step(0.01, src=1, dst=1, plain=True)
        c, _ = pool.step(0.01, src=0, dst=1)
        got = pool.download_aos(0, n, buf=1)
        assert c == c_ref
        for g, r in zip(got, ref):
            assert_same(g[2:-2, 2:-2], r[2:-2, 2:-2])
    finally:
        pool.close()
