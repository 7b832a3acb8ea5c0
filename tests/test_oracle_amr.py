"""The oracle's regrid and time interpolation between levels against the
reference (fc2d_geoclaw_average_fort.f average2coarse and
fc2d_geoclaw_timeinterp_fort.f, exact on tests/golden/amr_vectors.npz and on
fresh inputs through oracle/_ref/libgfref_amr.so when it is present), and
interpolate2fine -- whose slope function lives in ForestClaw, outside the
reference tree, so no reference output exists for it (parity unpinned) --
against what the interpolation fixes by itself."""
import os

import numpy as np
import pytest

from oracle.oracle import CLib, Params, ref_amr_lib

GOLD = os.path.join(os.path.dirname(__file__), "golden", "amr_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def clib():
    c = CLib()
    c.set_params(Params())
    return c


def test_average2coarse_matches_reference(gold, clib):
    for k in range(len(gold["avg_igrid"])):
        qc = gold["avg_qc"][k].copy()
        clib.average2coarse(qc, gold["avg_qf"][k], gold["avg_auxc"][k], gold["avg_auxf"][k], int(gold["avg_igrid"][k]))
        assert np.array_equal(qc, gold["avg_out"][k]), k


def test_timeinterp_matches_reference(gold, clib):
    psize = int(gold["ti_psize"][0])
    for k in range(len(gold["ti_alpha"])):
        out = np.zeros_like(gold["ti_curr"][k])
        assert clib.timeinterp(gold["ti_curr"][k], gold["ti_last"][k], out, float(gold["ti_alpha"][k]), psize) == 0
        assert np.array_equal(out, gold["ti_out"][k]), k
    out = np.zeros_like(gold["ti_curr"][0])
    assert clib.timeinterp(gold["ti_curr"][0], gold["ti_last"][0], out, 0.5, psize + 1) == 2


def test_average_and_timeinterp_match_reference_build_fresh(clib):
    ref = ref_amr_lib()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    ref.set_dry_tolerance(1e-3)
    rng = np.random.default_rng(5)
    for k in range(200):
        mx = (8, 16, 32)[k % 3]
        sh = (mx + 4, mx + 4)
        qc, qf = rng.normal(0, 1, (2,) + sh + (3,))
        qf[..., 0] = np.abs(qf[..., 0]) * (rng.random(sh) < 0.6)
        ac, af = rng.normal(0, 1, (2,) + sh + (1,))
        a, b = qc.copy(), qc.copy()
        ref.average2coarse(a, qf, ac, af, k % 4)
        clib.average2coarse(b, qf, ac, af, k % 4)
        assert np.array_equal(a, b)
        o1, o2 = np.zeros_like(qc), np.zeros_like(qc)
        ps = 8 * (mx - 2) * 3
        assert ref.timeinterp(qc, qf, o1, 0.3, ps) == clib.timeinterp(qc, qf, o2, 0.3, ps)
        assert np.array_equal(o1, o2)


def test_average2coarse_conserves_and_respects_dry(clib):
    """Wet fine cells over a coarse cell whose bathymetry is the mean of the
    fine ones: the coarse cell gets the mean depth and the summed momentum
    scaled to it; an all-dry family gives a dry coarse cell."""
    rng = np.random.default_rng(2)
    mx = 8
    qf = np.zeros((mx + 4, mx + 4, 3))
    qf[..., 0] = rng.uniform(0.5, 2, (mx + 4, mx + 4))
    qf[..., 1:] = rng.normal(0, 1, (mx + 4, mx + 4, 2))
    af = rng.normal(0, 0.1, (mx + 4, mx + 4, 1))
    for g in range(4):
        qc = np.zeros_like(qf)
        ac = np.zeros_like(af)
        ig, jg = g % 2, g // 2
        for j in range(mx // 2):
            for i in range(mx // 2):
                ac[2 + j + jg * mx // 2, 2 + i + ig * mx // 2, 0] = af[2 + 2 * j:4 + 2 * j, 2 + 2 * i:4 + 2 * i, 0].mean()
        clib.average2coarse(qc, qf, ac, af, g)
        for j in range(mx // 2):
            for i in range(mx // 2):
                blk = qf[2 + 2 * j:4 + 2 * j, 2 + 2 * i:4 + 2 * i]
                c = qc[2 + j + jg * mx // 2, 2 + i + ig * mx // 2]
                np.testing.assert_allclose(c[0], blk[..., 0].mean(), rtol=1e-13)
                np.testing.assert_allclose(c[1:], blk[..., 1:].sum((0, 1)) / 4, rtol=1e-12, atol=1e-14)
    qf[..., 0] = 0.0
    qc = np.ones_like(qf)
    clib.average2coarse(qc, qf, af * 0, af, 0)
    assert np.all(qc[2:2 + mx // 2, 2:2 + mx // 2] == 0.0)


def test_compute_slopes_cases(clib):
    for s in (-3.5, 0.25, 7.0):
        assert clib.compute_slopes(s, s) == s          # smooth data keeps its slope
    assert clib.compute_slopes(1.0, -2.0) == 0.0       # extremum
    assert clib.compute_slopes(0.0, 3.0) == 0.0
    assert clib.compute_slopes(1.0, 10.0) == 2.0       # 2*min(|sl|, |sr|)
    assert clib.compute_slopes(-4.0, -5.0) == -4.5     # |centred| is the smallest
    rng = np.random.default_rng(0)
    for sl, sr in rng.normal(0, 1, (1000, 2)):
        v = clib.compute_slopes(sl, sr)
        assert abs(v) <= min(2 * abs(sl), 2 * abs(sr), abs(sl + sr) / 2) + 1e-15
        assert v == 0.0 or np.sign(v) == np.sign(sl) == np.sign(sr)


def test_interpolate2fine_linear_fields_and_round_trip(clib):
    """Wet linear eta, hu, hv: each child reproduces the linear fields at its
    cell centres (and h = eta - b), in the quadrant of its igrid; averaging
    the four children back returns the coarse interior."""
    rng = np.random.default_rng(4)
    mx = 8
    for trial in range(20):
        c = rng.normal(0, 0.2, (3, 3))
        jj, ii = np.meshgrid(np.arange(-2, mx + 2), np.arange(-2, mx + 2), indexing="ij")
        xc, yc = ii + 0.5, jj + 0.5               # coarse cell centres, dx = 1
        qcoarse = np.zeros((mx + 4, mx + 4, 3))
        aux_c = rng.normal(0, 0.1, (mx + 4, mx + 4, 1))
        eta = 10 + c[0, 0] * xc + c[0, 1] * yc
        qcoarse[..., 0] = eta - aux_c[..., 0]
        qcoarse[..., 1] = c[1, 2] + c[1, 0] * xc + c[1, 1] * yc
        qcoarse[..., 2] = c[2, 2] + c[2, 0] * xc + c[2, 1] * yc
        kids = []
        for g in range(4):
            ig, jg = g % 2, g // 2
            fj, fi = np.meshgrid(np.arange(-2, mx + 2), np.arange(-2, mx + 2), indexing="ij")
            xf = (ig * mx / 2) + (fi + 0.5) / 2         # fine centres in coarse units
            yf = (jg * mx / 2) + (fj + 0.5) / 2
            aux_f = rng.normal(0, 0.1, (mx + 4, mx + 4, 1))
            qf = np.full((mx + 4, mx + 4, 3), np.nan)
            clib.interpolate2fine(qcoarse, qf, aux_c, aux_f, g)
            inner = (slice(2, mx + 2), slice(2, mx + 2))
            np.testing.assert_allclose(qf[inner][..., 0] + aux_f[inner][..., 0],
                                       (10 + c[0, 0] * xf + c[0, 1] * yf)[inner], atol=1e-12)
            for m in (1, 2):
                np.testing.assert_allclose(qf[inner][..., m], (c[m, 2] + c[m, 0] * xf + c[m, 1] * yf)[inner],
                                           atol=1e-12)
            kids.append((qf, aux_f))
        # average back (coarse bathymetry = mean of the fine cells it covers)
        back = np.zeros_like(qcoarse)
        ac2 = aux_c.copy()
        for g, (qf, af) in enumerate(kids):
            ig, jg = g % 2, g // 2
            for j in range(mx // 2):
                for i in range(mx // 2):
                    ac2[2 + j + jg * mx // 2, 2 + i + ig * mx // 2, 0] = af[2 + 2 * j:4 + 2 * j, 2 + 2 * i:4 + 2 * i, 0].mean()
        for g, (qf, af) in enumerate(kids):
            clib.average2coarse(back, np.nan_to_num(qf), ac2, af, g)
        inner = (slice(2, mx + 2), slice(2, mx + 2))
        want_h = (10 + c[0, 0] * xc + c[0, 1] * yc)[inner] - ac2[inner][..., 0]
        np.testing.assert_allclose(back[inner][..., 0], want_h, atol=1e-11)
        np.testing.assert_allclose(back[inner][..., 1:], qcoarse[inner][..., 1:], atol=1e-11)


def test_interpolate2fine_extremum_is_flat(clib):
    """At a local extremum of a coarse field the slope vanishes: the four
    fine cells under that coarse cell carry its value."""
    mx = 8
    qc = np.zeros((mx + 4, mx + 4, 3))
    qc[..., 0] = 1.0
    qc[..., 1] = 0.0
    qc[5, 5, 1] = 3.0          # coarse (i, j) = (4, 4) in Fortran indices: a peak
    ac = np.zeros((mx + 4, mx + 4, 1))
    qf = np.zeros_like(qc)
    clib.interpolate2fine(qc, qf, ac, np.zeros_like(ac), 0)
    assert np.all(qf[2 + 6:2 + 8, 2 + 6:2 + 8, 1] == 3.0)
    assert np.all(qf[2:mx + 2, 2:mx + 2, 0] == 1.0)
