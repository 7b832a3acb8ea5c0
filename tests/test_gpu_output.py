"""ASCII frame output and gauge sampling of device-resident patches against
the reference writers' bytes and gauge samples (tests/golden/
output_vectors.npz, from fc2d_geoclaw_output_ascii_fort.f and
geoclaw_gauges_fort.f90) and against the oracle."""
import os
import re
import tempfile

import numpy as np
import pytest

from geoflood_amd.pool import BrickLayout, Pool, gauge_file_append, gauge_file_create

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "output_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _patch_blocks(qbytes: bytes):
    """Split fort.q bytes into the per-patch blocks (each opens with the
    grid_number line)."""
    starts = [m.start() for m in re.finditer(rb"(?m)^.{5}                 grid_number\n", qbytes)]
    return [qbytes[a:b] for a, b in zip(starts, starts[1:] + [len(qbytes)])]


@pytest.mark.parametrize("k", [1, 3, 5])
def test_frame_bytes_match_reference(gf, gold, k):
    """The square patches of the golden frame (8x8, 16x16, 32x32; their
    values include signed zeros, values at and below the 1e-99 flush,
    three-digit exponents and rounding ties) written from a pool."""
    q = gold[f"frame_q{k}"]
    aux = gold[f"frame_aux{k}"]
    xl, yl, dx, dy = (float(v) for v in gold["frame_geom"][k])
    num, level, blockno, _ = (int(v) for v in gold["frame_ints"][k])
    mx = q.shape[1] - 4
    frame, ngrids = int(gold["frame_hdr"][0]), int(gold["frame_hdr"][1])
    print(f"case frame patch {k} ({mx}x{mx})", flush=True)
    pool = Pool(mx, mx, 1)
    try:
        pool.set_patches([dx], [dy], [xl], [yl])
        pool.upload_aos(0, q[None], aux[None], buf=1)
        d = tempfile.mkdtemp()
        pool.write_frame(d, frame, float(gold["frame_time"][0]), first=0, n=1, buf=1, ngrids=ngrids,
                         patch_num=[num], level=[level], blockno=[blockno])
        tag = "%04d" % (frame % 10000)
        with open(os.path.join(d, "fort.t" + tag), "rb") as f:
            assert f.read() == gold["frame_t_bytes"].tobytes()
        with open(os.path.join(d, "fort.q" + tag), "rb") as f:
            got = f.read()
        assert got == _patch_blocks(gold["frame_q_bytes"].tobytes())[k]
        # the writer flushes |q| < 1e-99 to +0 in the pool's state, as the reference does to its patch
        after = pool.download_aos(0, 1, buf=1)[0]
        inner = (slice(2, -2), slice(2, -2))
        want = q[inner].copy()
        want[np.abs(want) < 1e-99] = 0.0
        assert np.array_equal(after[inner], want)
        assert not np.any(np.signbit(after[inner]) & (after[inner] == 0))
    finally:
        pool.close()


def test_frame_appends_patches_in_order(gf, oracle_c):
    """A brick written in two calls (header with the first range, then the
    rest appended) equals the oracle's file for the whole brick."""
    rng = np.random.default_rng(3)
    n, mx = 64, 16
    layout = BrickLayout(4, 4, mx, mx, 0.5, 0.25, xlo=-10.0, ylo=3.0)
    h = np.abs(rng.normal(1, 0.5, (n, n))) * (rng.random((n, n)) < 0.7)
    hu, hv, b = (rng.normal(0, 1, (n, n)) for _ in range(3))
    pool = Pool(mx, mx, layout.npatch)
    try:
        pool.set_layout([layout])
        pool.upload_global(layout, h, hu, hv, b)
        qa = pool.download_aos(0, layout.npatch)
        aa = pool.download_aux(0, layout.npatch)
        d = tempfile.mkdtemp()
        lv = np.arange(layout.npatch) % 3
        pool.write_frame(d, 5, 0.125, first=0, n=6, level=lv[:6], mpirank=2, ngrids=layout.npatch)
        pool.write_frame(d, 5, 0.125, first=6, n=layout.npatch - 6, level=lv[6:], mpirank=2, header=False)
        ox, oy = layout.origins()
        want = b"".join(oracle_c.frame_patch(qa[k], aa[k], float(ox[k]), float(oy[k]), 0.5, 0.25, k, int(lv[k]),
                                             0, 2) for k in range(layout.npatch))
        with open(os.path.join(d, "fort.q0005"), "rb") as f:
            assert f.read() == want
        with open(os.path.join(d, "fort.t0005"), "rb") as f:
            assert f.read() == oracle_c.frame_header(0.125, 3, 1, layout.npatch, 2)
    finally:
        pool.close()


def test_gauges_match_reference(gf, gold):
    """Golden gauge samples: 8x8 patches whose ghost cells are stored (the
    host supplies them, as ForestClaw does), points anywhere in the patch,
    including cell edges and dry neighbourhoods."""
    gq, ga, geom, xy = gold["gauge_q"], gold["gauge_aux"], gold["gauge_geom"], gold["gauge_xy"]
    n = len(gq)
    pool = Pool(8, 8, n)
    try:
        pool.set_patches(geom[:, 2], geom[:, 3], geom[:, 0], geom[:, 1])
        pool.upload_aos(0, gq, ga)
        out = pool.sample_gauges(np.arange(n), xy[:, 0], xy[:, 1])
        assert np.array_equal(out, gold["gauge_out"])
    finally:
        pool.close()


def test_gauges_at_patch_edges_read_the_neighbour(gf, oracle_c):
    """Gauges in the last row/column of a patch of a brick read the
    neighbour's interior (what a ForestClaw ghost fill puts there) or the
    wall mirror: equal to the oracle on the ghost-filled brick."""
    from oracle.oracle import Brick
    rng = np.random.default_rng(8)
    n, mx = 64, 16
    h = np.abs(rng.normal(1, 0.3, (n, n)))
    h[rng.random((n, n)) < 0.05] = 0.0
    hu, hv, b = (rng.normal(0, 1, (n, n)) for _ in range(3))
    layout = BrickLayout(4, 4, mx, mx, 1.0, 1.0)
    pool = Pool(mx, mx, layout.npatch)
    try:
        pool.set_layout([layout])
        pool.upload_global(layout, h, hu, hv, b)
        br = Brick.from_global(h, hu, hv, b, 4, 4, 2, 1.0, 1.0)
        oracle_c.brick_ghost_fill(br)
        ox, oy = layout.origins()
        slots, xs, ys, want = [], [], [], []
        for _ in range(400):
            k = int(rng.integers(layout.npatch))
            fx = rng.uniform(mx - 1.5, mx) if rng.random() < 0.7 else rng.uniform(0, mx)
            fy = rng.uniform(mx - 1.5, mx) if rng.random() < 0.7 else rng.uniform(0, mx)
            x, y = ox[k] + fx, oy[k] + fy
            qv, av = oracle_c.update_gauge(br.q[k], br.aux[k], float(ox[k]), float(oy[k]), 1.0, 1.0, x, y)
            slots.append(k); xs.append(x); ys.append(y); want.append(np.r_[qv, av])
        got = pool.sample_gauges(slots, xs, ys)
        assert np.array_equal(got, np.array(want))
    finally:
        pool.close()


def test_gauge_files(gf, oracle_c):
    d = tempfile.mkdtemp()
    gauge_file_create(d, 12, 86.001, 0.001)
    rec = np.array([[1.25, -0.5, 0.0, -1.25], [2.0, 3e-7, -4e120, 1e-100 - 2.0], [0.1, 0.2, 0.3, 0.4]])
    gauge_file_append(d, 12, [2, 1, 3], [0.5, 1e3, 7.0], rec)
    with open(os.path.join(d, "gauge00012.txt"), "rb") as f:
        got = f.read()
    want = oracle_c.gauge_header(12, 86.001, 0.001) + b"".join(
        oracle_c.gauge_line(lv, t, r[:3], r[3:]) for lv, t, r in zip([2, 1, 3], [0.5, 1e3, 7.0], rec))
    assert got == want
