"""Regrid (average2coarse, interpolate2fine) and time interpolation of
resident patches against the reference's outputs (tests/golden/
amr_vectors.npz) and the oracle."""
import os

import numpy as np
import pytest

from oracle.oracle import Brick
from geoflood_amd.pool import BrickLayout, Pool

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "amr_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_average2coarse_matches_reference(gf, oracle_c, gold):
    """48 families of 8x8 patches: child g of family k holds the golden fine
    patch of case k (the other children copies of it); every quadrant of the
    coarse result equals the oracle's, and quadrant igrid the reference's."""
    qc, qf, ac, af, ig, out = (gold[k] for k in ("avg_qc", "avg_qf", "avg_auxc", "avg_auxf", "avg_igrid", "avg_out"))
    n = len(ig)
    pool = Pool(8, 8, 5 * n)
    try:
        pool.set_patches(np.ones(5 * n), np.ones(5 * n), np.zeros(5 * n), np.zeros(5 * n))
        q_all = np.concatenate([qc, np.repeat(qf, 4, axis=0)])
        a_all = np.concatenate([ac, np.repeat(af, 4, axis=0)])
        pool.upload_aos(0, q_all, a_all)
        pool.average2coarse(np.arange(n), n + np.arange(4 * n))
        got = pool.download_aos(0, n)
        inner = (slice(2, 10), slice(2, 10))
        for k in range(n):
            want = qc[k].copy()
            for g in range(4):
                oracle_c.average2coarse(want, qf[k], ac[k], af[k], g)
            assert np.array_equal(got[k][inner], want[inner]), k
            g = int(ig[k])
            sl = (slice(2 + 4 * (g // 2), 6 + 4 * (g // 2)), slice(2 + 4 * (g % 2), 6 + 4 * (g % 2)))
            assert np.array_equal(got[k][sl], out[k][sl]), k
    finally:
        pool.close()


def test_timeinterp_matches_reference(gf, gold):
    tc, tl, al, to = gold["ti_curr"], gold["ti_last"], gold["ti_alpha"], gold["ti_out"]
    m = len(al)
    pool = Pool(8, 8, 2 * m)
    try:
        pool.set_patches(np.ones(2 * m), np.ones(2 * m), np.zeros(2 * m), np.zeros(2 * m))
        pool.upload_aos(0, tl, None, buf=0)
        pool.upload_aos(0, tc, None, buf=1)
        pool.upload_aos(m, np.zeros_like(tc), None, buf=1)
        for k in range(m):   # one alpha per call
            pool.timeinterp([k], 0, 1, float(al[k]), [m + k], 1)
        got = pool.download_aos(m, m, buf=1)
        ring = np.zeros((12, 12), bool)
        ring[2:10, 2:10] = True
        ring[4:8, 4:8] = False
        for k in range(m):
            assert np.array_equal(got[k][ring], to[k][ring]), k
            assert np.all(got[k][4:8, 4:8] == 0.0)
    finally:
        pool.close()


def test_interpolate2fine_matches_oracle_on_a_brick(gf, oracle_c):
    """Children of coarse patches of a 4x4 brick (walls outside, same-level
    neighbours inside): the coarse halo is read through the neighbours, and
    the children equal the oracle's interpolation from the ghost-filled
    coarse patches, in the interior and the inner ghost layers it writes."""
    rng = np.random.default_rng(12)
    mx, n = 16, 64
    b = rng.normal(0, 0.3, (n, n))
    h = np.maximum(1.5 - b + rng.normal(0, 0.2, (n, n)), 0.0)
    h[rng.random((n, n)) < 0.1] = 0.0
    hu, hv = rng.normal(0, 0.5, (2, n, n)) * (h > 0)
    layout = BrickLayout(4, 4, mx, mx, 1.0, 1.0)
    fam = [0, 3, 5, 6, 10, 15]
    nk = 4 * len(fam)
    pool = Pool(mx, mx, layout.npatch + nk)
    try:
        pool.set_layout([layout])
        ox, oy = layout.origins()
        # children: their own slots after the brick, with fine bathymetry
        cdx = np.full(nk, 0.5)
        cxl = np.array([ox[c] + 0.5 * mx * (g % 2) for c in fam for g in range(4)])
        cyl = np.array([oy[c] + 0.5 * mx * (g // 2) for c in fam for g in range(4)])
        dx = np.concatenate([np.ones(layout.npatch), cdx])
        xl = np.concatenate([ox, cxl])
        yl = np.concatenate([oy, cyl])
        nb = np.concatenate([layout.neighbours(), np.full((nk, 8), -3, np.int32)])
        pool.set_patches(dx, dx, xl, yl, nb)
        pool.upload_global(layout, h, hu, hv, b)
        af = rng.normal(0, 0.3, (nk, mx + 4, mx + 4, 1))
        pool.upload_aos(layout.npatch, np.zeros((nk, mx + 4, mx + 4, 3)), af)
        pool.interpolate2fine(fam, layout.npatch + np.arange(nk))
        got = pool.download_aos(layout.npatch, nk)
        br = Brick.from_global(h, hu, hv, b, 4, 4, 2, 1.0, 1.0)
        oracle_c.brick_ghost_fill(br)
        for f, c in enumerate(fam):
            for g in range(4):
                want = np.zeros((mx + 4, mx + 4, 3))
                oracle_c.interpolate2fine(br.q[c], want, br.aux[c], af[4 * f + g], g)
                ig, jg = g % 2, g // 2
                rows = slice(2 - 2 * jg, mx + 2 + 2 * (1 - jg))
                cols = slice(2 - 2 * ig, mx + 2 + 2 * (1 - ig))
                assert np.array_equal(got[4 * f + g][rows, cols], want[rows, cols]), (c, g)
    finally:
        pool.close()


def test_regrid_rejects_overlapping_slots(gf):
    pool = Pool(8, 8, 10)
    try:
        pool.set_patches(np.ones(10), np.ones(10), np.zeros(10), np.zeros(10))
        with pytest.raises(RuntimeError):
            pool.average2coarse([0], [0, 1, 2, 3])
        with pytest.raises(RuntimeError):
            pool.interpolate2fine([0, 1], [2, 3, 4, 5, 6, 7, 8, 2])
    finally:
        pool.close()


CF_LINKS = [(s, g) for s in range(4) for g in range(2)] + [(s, 0) for s in range(4, 8)]


def _cf_pool(rng, mx, n_links):
    """Slots 0..n-1 coarse, n..2n-1 fine, random wet/dry states and bathymetry
    (ghost cells included), every side host-supplied (GF_NB_STORED)."""
    n = n_links
    pool = Pool(mx, mx, 2 * n)
    pool.set_patches(np.ones(2 * n), np.ones(2 * n), np.zeros(2 * n), np.zeros(2 * n))
    sh = (2 * n, mx + 4, mx + 4)
    q = rng.normal(0, 1, sh + (3,))
    q[..., 0] = np.abs(q[..., 0]) * (rng.random(sh) < 0.7)
    q[..., 0][rng.random(sh) < 0.05] = 1e-3  # exactly at the dry tolerance
    aux = rng.normal(0, 0.5, sh + (1,))
    pool.upload_aos(0, q, aux)
    return pool, q, aux


@pytest.mark.parametrize("mx", [8, 16, 32, 64])
def test_average_ghost_matches_oracle(gf, oracle_c, mx):
    """Every face half and corner in one call: the coarse patches' ghost
    cells equal the oracle's average_face / average_corner bit for bit, and
    nothing else of any patch changes."""
    rng = np.random.default_rng(200 + mx)
    n = len(CF_LINKS)
    print(f"case average_ghost mx={mx} links={n}", flush=True)
    pool, q, aux = _cf_pool(rng, mx, n)
    try:
        side = [s for s, _ in CF_LINKS]
        igrid = [g for _, g in CF_LINKS]
        pool.average_ghost(np.arange(n), n + np.arange(n), side, igrid)
        got = pool.download_aos(0, 2 * n)
        for k, (s, g) in enumerate(CF_LINKS):
            want = q[k].copy()
            oracle_c.average_ghost(want, q[n + k], aux[k], aux[n + k], s, g)
            assert np.array_equal(got[k], want), (k, s, g)
            assert np.array_equal(got[n + k], q[n + k])
    finally:
        pool.close()


@pytest.mark.parametrize("mx", [8, 16, 32, 64])
def test_interpolate_ghost_matches_oracle(gf, oracle_c, mx):
    """Every face half and corner in one call: the fine patches' ghost cells
    equal the oracle's interpolate_face / interpolate_corner bit for bit
    (coarse ghost cells read as stored), and nothing else changes."""
    rng = np.random.default_rng(300 + mx)
    n = len(CF_LINKS)
    print(f"case interpolate_ghost mx={mx} links={n}", flush=True)
    pool, q, aux = _cf_pool(rng, mx, n)
    try:
        side = [s for s, _ in CF_LINKS]
        igrid = [g for _, g in CF_LINKS]
        pool.interpolate_ghost(np.arange(n), n + np.arange(n), side, igrid)
        got = pool.download_aos(0, 2 * n)
        for k, (s, g) in enumerate(CF_LINKS):
            want = q[n + k].copy()
            oracle_c.interpolate_ghost(q[k], want, aux[k], aux[n + k], s, g)
            assert np.array_equal(got[n + k], want), (k, s, g)
            assert np.array_equal(got[k], q[k])
    finally:
        pool.close()


def test_cf_ghost_link_checks(gf):
    """Bad links are refused before any launch."""
    pool = Pool(8, 8, 4)
    try:
        pool.set_patches(np.ones(4), np.ones(4), np.zeros(4), np.zeros(4))
        bad = [
            ([0], [0], [1], [0]),            # coarse == fine
            ([0], [1], [8], [0]),            # side out of range
            ([0], [1], [1], [2]),            # igrid out of range
            ([0], [9], [1], [0]),            # slot out of range
            ([0, 0], [1, 2], [1, 1], [0, 0]),  # same coarse region twice
        ]
        for c, f, s, g in bad:
            with pytest.raises(gf.GeoFloodError):
                pool.average_ghost(c, f, s, g)
        with pytest.raises(gf.GeoFloodError):   # slot 1 fine in one link, coarse in the next
            pool.interpolate_ghost([0, 1], [1, 2], [1, 1], [0, 0])
        with pytest.raises(gf.GeoFloodError):   # same fine region twice
            pool.interpolate_ghost([0, 2], [1, 1], [1, 1], [0, 0])
        pool.average_ghost([0, 0], [1, 2], [1, 1], [0, 1])  # both halves of one face: fine
        pool.interpolate_ghost([], [], [], [])
    finally:
        pool.close()


def test_pool_copy_between_buffers(gf):
    """gf_pool_copy: the listed slots' states (ghost cells included) are
    copied exactly from one buffer to the other; other slots and the source
    are untouched; bad arguments are refused."""
    rng = np.random.default_rng(21)
    pool = Pool(16, 16, 6)
    try:
        pool.set_patches(np.ones(6), np.ones(6), np.zeros(6), np.zeros(6))
        a = rng.normal(size=(6, 20, 20, 3))
        b = rng.normal(size=(6, 20, 20, 3))
        pool.upload_aos(0, a, None, buf=0)
        pool.upload_aos(0, b, None, buf=1)
        pool.copy([4, 1], 0, 1)
        got0, got1 = pool.download_aos(0, 6, buf=0), pool.download_aos(0, 6, buf=1)
        assert np.array_equal(got0, a)
        want = b.copy()
        want[[1, 4]] = a[[1, 4]]
        assert np.array_equal(got1, want)
        for slots, s, d in (([0], 0, 0), ([0], 0, 2), ([6], 0, 1), ([-1], 1, 0)):
            with pytest.raises(gf.GeoFloodError):
                pool.copy(slots, s, d)
        pool.copy([], 1, 0)
    finally:
        pool.close()
