"""gf_step2_batch_noalloc, the drop-in batch step that allocates nothing
(its device memory is a fixed per-GPU array), against gf_step2_batch on the
same host patches: bit-identical patches and CFL number on two calls in a
row, the second smaller than the first (the array keeps the first call's
contents), at 32x32 and at 64x64 (tiled, whose tile fold slots must start at
zero), and a batch beyond the array's size refused with GF_EINVAL."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _patches(rng, n, mx, maux=1):
    qs, auxs = [], []
    for k in range(n):
        h = rng.random((mx + 4, mx + 4)) * 2.0
        h[rng.random(h.shape) < 0.3] = 0.0
        q = np.zeros(h.shape + (3,))
        q[..., 0] = h
        q[..., 1] = rng.normal(size=h.shape) * h
        q[..., 2] = rng.normal(size=h.shape) * h
        qs.append(q)
        auxs.append(np.ascontiguousarray(rng.normal(0, 0.3, h.shape + (maux,))))
    return qs, auxs


def _call(gf, fn, qs, auxs, mx, dx, dt):
    n = len(qs)
    qp = (ct.POINTER(ct.c_double) * n)(*[gf.dptr(q) for q in qs])
    ap = (ct.POINTER(ct.c_double) * n)(*[gf.dptr(a) for a in auxs])
    dxa = np.full(n, dx)
    c = ct.c_double(-1.0)
    rc = fn(mx, mx, 2, auxs[0].shape[-1], n, qp, ap, gf.dptr(dxa), gf.dptr(dxa), 0.0, dt, ct.byref(c))
    return rc, c.value


@pytest.mark.parametrize("mx,sizes", [(32, (200, 37)), (64, (40, 3))])
def test_noalloc_equals_batch_step(gf, mx, sizes):
    lib = gf.load()
    rng = np.random.default_rng(mx)
    for n in sizes:
        print(f"case noalloc {n} patches {mx}x{mx}", flush=True)
        qs, auxs = _patches(rng, n, mx)
        ref = [q.copy() for q in qs]
        rc, c_ref = _call(gf, lib.gf_step2_batch, ref, auxs, mx, 0.1, 0.005)
        assert rc == 0
        rc, c = _call(gf, lib.gf_step2_batch_noalloc, qs, auxs, mx, 0.1, 0.005)
        assert rc == 0, lib.gf_last_error()
        assert c == c_ref and c > 0
        for a, b in zip(qs, ref):
            assert np.array_equal(a, b)


def test_noalloc_refuses_oversized_batch(gf):
    lib = gf.load()
    rng = np.random.default_rng(1)
    qs, auxs = _patches(rng, 300, 32)
    keep = [q.copy() for q in qs]
    rc, _ = _call(gf, lib.gf_step2_batch_noalloc, qs, auxs, 32, 0.1, 0.005)
    assert rc == -1 and b"scratch" in lib.gf_last_error()
    assert all(np.array_equal(a, b) for a, b in zip(qs, keep))
