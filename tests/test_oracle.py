"""The CPU oracle (oracle/gf_oracle.c) pinned three ways:

1. against the golden vectors written by the reference GeoClaw Fortran step
   (tests/golden/geoclaw_step_vectors.npz) -- bit for bit;
2. against the reference build itself (oracle/_ref) on fresh random inputs,
   where that build is present -- bit for bit;
3. against closed forms and invariants that the mathematics fixes, each
   chosen so that a plausible slip (dropped term, wrong sign or index,
   swapped operand) breaks it.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle.oracle import Brick, Params

GOLD = os.path.join(os.path.dirname(__file__), "golden", "geoclaw_step_vectors.npz")
DRYTOL, G = 1e-3, 9.81


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.int64)


def same(a, b):
    return np.array_equal(bits(a), bits(b))


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


# ----------------------------------------------------------------------------
# 1. golden vectors
# ----------------------------------------------------------------------------
def test_golden_riemanntype(oracle_c, gold):
    for c, o, r in zip(gold["rt_in"], gold["rt_out"], gold["rt_rare"]):
        out, rare = oracle_c.riemanntype(*c, DRYTOL, G)
        assert same(out, o) and np.array_equal(rare, r), c


def test_golden_riemann_aug_jcp(oracle_c, gold):
    for vin, sw, fw in zip(gold["jcp_in"], gold["jcp_sw"], gold["jcp_fw"]):
        a, b = oracle_c.riemann_aug_jcp(vin, DRYTOL, G)
        assert same(a, sw) and same(b, fw)


def test_golden_rpn2_rpt2_flux2(oracle_c, gold):
    mx, mbc = 16, 2
    for k in range(len(gold["sl_q"])):
        q1d = np.ascontiguousarray(gold["sl_q"][k])
        aux = [np.ascontiguousarray(gold["sl_aux"][k][:, i:i + 1]) for i in range(3)]
        ixy = int(gold["sl_ixy"][k])
        qq = q1d.copy()
        fw, s, am, ap = oracle_c.rpn2(ixy, mx, mbc, qq, aux[1])
        assert same(fw, gold["rpn2_fw"][k]) and same(s, gold["rpn2_s"][k])
        assert same(am, gold["rpn2_amdq"][k]) and same(ap, gold["rpn2_apdq"][k])
        assert same(qq, gold["rpn2_q"][k])
        asdq = np.ascontiguousarray(gold["rpt2_asdq"][k])
        for imp in (1, 2):
            bm, bp = oracle_c.rpt2(ixy, imp, mx, mbc, q1d, aux[0], aux[1], aux[2], asdq)
            assert same(bm, gold["rpt2_bm"][k][imp - 1]) and same(bp, gold["rpt2_bp"][k][imp - 1])
        fm, fp, gm, gp, cfl = oracle_c.flux2(ixy, mx, mbc, q1d.copy(), np.full(20, 0.3), *aux)
        assert same(fm, gold["flux2_fm"][k]) and same(fp, gold["flux2_fp"][k])
        assert same(gm, gold["flux2_gm"][k]) and same(gp, gold["flux2_gp"][k])
        assert same(cfl, gold["flux2_cfl"][k])


def test_golden_limiter(oracle_c, gold):
    for k in range(len(gold["lim_w"])):
        w = np.ascontiguousarray(gold["lim_w"][k]).copy()
        oracle_c.limiter(16, 2, w, np.ascontiguousarray(gold["lim_s"][k]), (4, 4, 4) if k % 2 == 0 else (1, 2, 3))
        assert same(w, gold["lim_out"][k])
    for (a, b, m), o in zip(gold["philim_in"], gold["philim_out"]):
        assert same(oracle_c.philim(a, b, int(m)), o)


def test_golden_patch_update(oracle_c, gold):
    for q, aux, qo, c in zip(gold["patch_q"], gold["patch_aux"], gold["patch_qout"], gold["patch_cfl"]):
        q = np.ascontiguousarray(q).copy()
        cc = oracle_c.patch_update(16, 16, 2, q, np.ascontiguousarray(aux), 0.1, 0.15, 0.0, 0.004)
        assert same(q, qo) and same(cc, c)


def test_golden_brick_step(oracle_c, gold):
    h, hu, hv, b = gold["brick_in"]
    br = Brick.from_global(h, hu, hv, b, 2, 2, 2, 0.05, 0.05)
    c = oracle_c.brick_advance(br, 0.0, 0.002)
    assert same(br.to_global(), gold["brick_out"]) and same(c, gold["brick_cfl"])


def test_golden_radial_dam_break(oracle_c, gold):
    from geoflood_amd.scenarios import radial_dam_break
    h, hu, hv, b, dx = radial_dam_break(256)
    br = Brick.from_global(h, hu, hv, b, 8, 8, 2, dx, dx)
    rec = oracle_c.brick_run(br, 100, 1e-4, 0.9, 1.0, nthreads=4)
    fin = br.to_global()
    assert same(rec["dt"], gold["dam_dt"]) and same(rec["cfl"], gold["dam_cfl"])
    assert same(rec["mass"], gold["dam_mass"]) and rec["retakes"] == int(gold["dam_retakes"])
    assert same(fin[:, ::16, ::16], gold["dam_final_sub"])
    assert hashlib.sha256(np.ascontiguousarray(fin).tobytes()).hexdigest() == str(gold["dam_final_sha256"])


# ----------------------------------------------------------------------------
# 2. against the reference build on fresh inputs
# ----------------------------------------------------------------------------
def test_vs_reference_random_patches(oracle_c, oracle_ref):
    rng = np.random.default_rng(11)
    for k in range(40):
        n = [8, 16, 32][k % 3]
        h = rng.random((n + 4, n + 4)) * rng.choice([0.005, 1.0, 6.0])
        h[rng.random(h.shape) < rng.random()] = 0.0
        q = np.zeros(h.shape + (3,))
        q[..., 0] = h
        q[..., 1:] = rng.normal(0, 2, h.shape + (2,)) * h[..., None]
        aux = np.ascontiguousarray(rng.normal(0, rng.choice([0, 0.3, 3.0]) + 1e-300, h.shape)[..., None])
        qa, qb = q.copy(), q.copy()
        ca = oracle_c.patch_update(n, n, 2, qa, aux, 0.2, 0.1, 0.0, 0.01)
        cb = oracle_ref.patch_update(n, n, 2, qb, aux, 0.2, 0.1, 0.0, 0.01)
        assert same(qa, qb) and same(ca, cb)


@pytest.mark.parametrize("order,transverse,mthlim", [(1, 0, (4, 4, 4)), (2, 1, (1, 1, 1)), (2, 2, (2, 3, 5)),
                                                      (2, 0, (4, 0, 4))])
def test_vs_reference_method_variants(oracle_c, oracle_ref, order, transverse, mthlim):
    p = Params(order=order, transverse=transverse, mthlim=mthlim, friction_forcing=0)
    oracle_c.set_params(p)
    oracle_ref.set_params(p)
    try:
        rng = np.random.default_rng(order * 10 + transverse)
        for k in range(8):
            h = 0.5 + rng.random((20, 20))
            h[rng.random(h.shape) < 0.2] = 0.0
            q = np.zeros(h.shape + (3,))
            q[..., 0] = h
            q[..., 1:] = rng.normal(0, 1, h.shape + (2,)) * h[..., None]
            aux = np.ascontiguousarray(rng.normal(0, 0.4, h.shape)[..., None])
            qa, qb = q.copy(), q.copy()
            assert same(oracle_c.patch_update(16, 16, 2, qa, aux, 0.1, 0.1, 0, 0.01),
                        oracle_ref.patch_update(16, 16, 2, qb, aux, 0.1, 0.1, 0, 0.01))
            assert same(qa, qb)
    finally:
        oracle_c.set_params(Params())
        oracle_ref.set_params(Params())


# ----------------------------------------------------------------------------
# 3. closed forms and invariants
# ----------------------------------------------------------------------------
def test_riemanntype_closed_forms(oracle_c):
    # dry branch: s1m = s2m = uR + uL - 2 sqrt(g hR) + 2 sqrt(g hL), hm = 0
    hL, uL = 2.0, 0.7
    out, rare = oracle_c.riemanntype(hL, 0.0, uL, 0.0, DRYTOL, G)
    assert out[0] == 0.0 and out[1] == out[2] == pytest.approx(uL + 2 * np.sqrt(G * hL))
    assert tuple(rare) == (1, 0)
    # two rarefactions: hm = (uL - uR + 2(cL + cR))^2 / (16 g) when the bracket is positive
    hL, hR, uL, uR = 1.0, 1.5, -2.0, 2.5
    out, rare = oracle_c.riemanntype(hL, hR, uL, uR, DRYTOL, G)
    cL, cR = np.sqrt(G * hL), np.sqrt(G * hR)
    assert tuple(rare) == (1, 1)
    assert out[0] == pytest.approx((uL - uR + 2 * (cL + cR)) ** 2 / (16 * G), rel=1e-14)
    assert out[1] == pytest.approx(uL + 2 * cL - 3 * np.sqrt(G * out[0]), rel=1e-13)


def _slice(h, hu, hv, b):
    q = np.ascontiguousarray(np.stack([h, hu, hv], axis=1).astype(np.float64))
    return q, np.ascontiguousarray(np.asarray(b, dtype=np.float64)[:, None])


def test_rpn2_uniform_state_and_dry(oracle_c):
    n = 20
    q, b = _slice(np.full(n, 1.3), np.full(n, 0.4), np.full(n, -0.2), np.zeros(n))
    fw, s, am, ap = oracle_c.rpn2(1, 16, 2, q.copy(), b)
    assert np.all(fw == 0.0) and np.all(am == 0.0) and np.all(ap == 0.0)
    q, b = _slice(np.full(n, 5e-4), np.full(n, 1e-4), np.zeros(n), np.zeros(n))
    fw, s, am, ap = oracle_c.rpn2(2, 16, 2, q.copy(), b)
    assert np.all(fw == 0.0) and np.all(s == 0.0)


def test_rpn2_lake_at_rest_well_balanced(oracle_c):
    rng = np.random.default_rng(5)
    n = 20
    b = rng.normal(0, 0.8, n)
    eta = 3.0
    q, bb = _slice(eta - b, np.zeros(n), np.zeros(n), b)
    fw, s, am, ap = oracle_c.rpn2(1, 16, 2, q.copy(), bb)
    assert np.max(np.abs(am)) < 1e-12 and np.max(np.abs(ap)) < 1e-12


def test_rpn2_fwave_mass_identity_and_mirror(oracle_c):
    """sum_w fwave(1,w) = huR - huL (Cramer row 2), and the mirror problem
    (left/right swapped, normal velocity negated) mirrors the waves."""
    rng = np.random.default_rng(6)
    n = 20
    h = 0.5 + rng.random(n) * 2
    hu = rng.normal(0, 1, n) * h
    hv = rng.normal(0, 1, n) * h
    b = rng.normal(0, 0.3, n)
    q, bb = _slice(h, hu, hv, b)
    fw, s, am, ap = oracle_c.rpn2(1, 16, 2, q.copy(), bb)
    for i in range(1, n):
        assert np.sum(fw[i, :, 0]) == pytest.approx(hu[i] - hu[i - 1], rel=1e-12, abs=1e-12)
    qm, bm = _slice(h[::-1].copy(), -hu[::-1], hv[::-1].copy(), b[::-1].copy())
    fwm, sm, amm, apm = oracle_c.rpn2(1, 16, 2, qm.copy(), bm)
    # Mirroring x -> -x maps q -> M q with M = diag(1, -1, 1) and the flux
    # difference to M (f(qR) - f(qL)); left-going waves become right-going,
    # so face i's amdq is M times the mirrored face's apdq.
    for i in range(2, n - 1):
        j = n - i
        np.testing.assert_allclose(am[i], apm[j] * np.array([1, -1, 1]), rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(ap[i], amm[j] * np.array([1, -1, 1]), rtol=1e-12, atol=1e-13)


def test_rpt2_reconstructs_jacobian_product(oracle_c):
    """With all waves admitted (eta above the neighbours' topography and
    no zero speed), bmasdq + bpasdq = R diag(s) R^-1 asdq, computed here by
    numpy's solver."""
    rng = np.random.default_rng(7)
    n = 20
    h = 1.0 + rng.random(n)
    hu = rng.normal(0, 0.5, n) * h
    hv = rng.normal(0, 0.5, n) * h
    q, aux2 = _slice(h, hu, hv, np.zeros(n))
    low = np.ascontiguousarray(np.full((n, 1), -5.0))
    asdq = rng.normal(0, 1, (n, 3))
    for imp in (1, 2):
        bm, bp = oracle_c.rpt2(1, imp, 16, 2, q, low, aux2, low, asdq)
        for i in range(1, n):
            c = i - 1 if imp == 1 else i
            u, v, c0 = hu[c] / h[c], hv[c] / h[c], np.sqrt(G * h[c])
            R = np.array([[1, 0, 1], [u, 1, u], [v - c0, 0, v + c0]])
            sv = np.array([v - c0, v, v + c0])
            want = R @ np.diag(sv) @ np.linalg.solve(R, asdq[i])
            np.testing.assert_allclose(bm[i] + bp[i], want, rtol=1e-10, atol=1e-12)
    # dry cell and cell below both neighbours' topography: no split
    q2 = q.copy()
    q2[:, 0] = 5e-4
    bm, bp = oracle_c.rpt2(1, 2, 16, 2, q2, low, aux2, low, asdq)
    assert np.all(bm == 0) and np.all(bp == 0)
    high = np.ascontiguousarray(np.full((n, 1), 50.0))
    bm, bp = oracle_c.rpt2(1, 1, 16, 2, q, high, aux2, high, asdq)
    assert np.all(bm == 0) and np.all(bp == 0)


def test_philim_mc_closed_form(oracle_c):
    for a, b in [(1.0, 0.5), (1.0, -3.0), (2.0, 8.0), (1.0, 1.0), (4.0, 0.1)]:
        r = b / a
        assert oracle_c.philim(a, b, 4) == max(0.0, min((1 + r) / 2, 2.0, 2 * r))
    assert oracle_c.philim(0.0, 1.0, 4) == 0.0


def test_src2_closed_form(oracle_c):
    rng = np.random.default_rng(8)
    q = np.zeros((20, 20, 3))
    h = 0.01 + rng.random((20, 20))
    q[..., 0] = h
    q[..., 1] = rng.normal(size=h.shape)
    q[..., 2] = rng.normal(size=h.shape)
    aux = np.zeros((20, 20, 1))
    q0 = q.copy()
    dt, n = 0.05, 0.025
    oracle_c.src2(16, 16, 2, q, aux, dt)
    hu, hv = q0[2:-2, 2:-2, 1], q0[2:-2, 2:-2, 2]
    gamma = np.sqrt(hu ** 2 + hv ** 2) * G * n ** 2 / h[2:-2, 2:-2] ** (7.0 / 3.0)
    np.testing.assert_allclose(q[2:-2, 2:-2, 1], hu / (1 + dt * gamma), rtol=1e-14)
    np.testing.assert_allclose(q[2:-2, 2:-2, 2], hv / (1 + dt * gamma), rtol=1e-14)
    assert np.array_equal(q[:2], q0[:2])  # ghost rows untouched


def _brick(h, hu, hv, b, mi, mj, dx=0.1, mthbc=(3, 3, 3, 3)):
    return Brick.from_global(h, hu, hv, b, mi, mj, 2, dx, dx, mthbc)


def test_step_constant_state_and_lake_at_rest(oracle_c):
    z = np.zeros((32, 32))
    br = _brick(np.full((32, 32), 1.7), z, z, z, 2, 2)
    oracle_c.brick_run(br, 5, 0.005)
    g = br.to_global()
    assert np.all(g[0] == 1.7) and np.all(g[1] == 0) and np.all(g[2] == 0)
    rng = np.random.default_rng(9)
    b = rng.normal(0, 0.3, (32, 32))
    br = _brick(2.0 - b, z, z, b, 2, 2)
    oracle_c.brick_run(br, 10, 0.005)
    g = br.to_global()
    assert np.max(np.abs(g[0] + b - 2.0)) < 1e-12 and np.max(np.abs(g[1:])) < 1e-12


def test_step_mass_conservation_and_transpose_symmetry(oracle_c):
    rng = np.random.default_rng(10)
    h = 1 + rng.random((32, 32))
    h[:, :10] = 0.0  # a dry band
    hu = rng.normal(0, 0.3, h.shape) * h
    hv = rng.normal(0, 0.3, h.shape) * h
    b = rng.normal(0, 0.2, h.shape)
    br = _brick(h, hu, hv, b, 2, 2)
    rec = oracle_c.brick_run(br, 20, 0.002)
    # b4step2 clips negative depths at the dry band (GeoClaw adds that mass);
    # without clipping, a walled domain conserves mass to roundoff
    hw = 1 + rng.random((32, 32))
    bw = _brick(hw, hu, hv, b, 2, 2)
    recw = oracle_c.brick_run(bw, 20, 0.002)
    m0 = hw.sum() * 0.01
    assert np.all(np.abs(recw["mass"] - m0) <= 1e-12 * m0)
    bt = _brick(h.T.copy(), hv.T.copy(), hu.T.copy(), b.T.copy(), 2, 2)
    rect = oracle_c.brick_run(bt, 20, 0.002)
    g, gt = br.to_global(), bt.to_global()
    np.testing.assert_allclose(gt[0], g[0].T, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(gt[1], g[2].T, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(gt[2], g[1].T, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(rect["dt"], rec["dt"], rtol=1e-13)


def test_step_y_uniform_matches_1d_reduction(oracle_c):
    """Data independent of y evolves independently of y, and one row of
    the 2-D step equals a 1-D first-order-in-y computation: every row is
    identical after the step."""
    rng = np.random.default_rng(12)
    x = 1 + rng.random(32)
    h = np.tile(x, (32, 1))
    hu = np.tile(rng.normal(0, 0.2, 32), (32, 1)) * h
    z = np.zeros_like(h)
    b = np.tile(rng.normal(0, 0.1, 32), (32, 1))
    br = _brick(h, hu, z, b, 2, 2)
    oracle_c.brick_run(br, 8, 0.002)
    g = br.to_global()
    for m in range(3):
        assert np.all(g[m] == g[m][0:1, :])
    assert np.all(g[2] == 0.0)


def test_ghost_fill_equals_global_reflection(oracle_c):
    """ForestClaw-order ghost exchange (face copies, corner copies, bc2 walls
    left/right/bottom/top) equals reflecting the global padded array."""
    rng = np.random.default_rng(13)
    h = 1 + rng.random((48, 32)); hu = rng.normal(size=h.shape); hv = rng.normal(size=h.shape)
    b = rng.normal(size=h.shape)
    br = _brick(h, hu, hv, b, 2, 3)
    oracle_c.brick_ghost_fill(br)
    H = np.pad(np.stack([h, hu, hv]), ((0, 0), (2, 2), (2, 2)), mode="symmetric")
    H[1, :, :2] *= -1; H[1, :, -2:] *= -1
    H[2, :2, :] *= -1; H[2, -2:, :] *= -1
    for pj in range(3):
        for pi in range(2):
            tile = H[:, pj * 16:pj * 16 + 20, pi * 16:pi * 16 + 20]
            assert np.array_equal(np.moveaxis(br.q[pj * 2 + pi], 2, 0), tile)


def test_brick_run_dt_bound(oracle_c):
    """The dt bound of the run loop (the library's gf_pool_set_scenario_dt_max;
    no counterpart in the reference).  dt_max = 0 is the reference's rule
    dt * desired_cfl / maxcfl, which test_golden_radial_dam_break pins to the
    reference build's dt sequence through the same loop.  With a bound, an
    all-dry brick (no face has a wave, CFL 0) steps dt0 then dt_max every
    step, and a dam break follows dt_k = min(dt_{k-1} * 0.9 / cfl_{k-1},
    dt_max) exactly, capped at first and below the cap later."""
    n = 64
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    b = 0.2 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.02)
    z = np.zeros_like(b)
    br = Brick.from_global(z, z, z, b, 2, 2, 2, 1.0 / n, 1.0 / n)
    rec = oracle_c.brick_run(br, 5, 1e-3, dt_max=0.25)
    assert np.all(rec["cfl"] == 0.0)
    assert rec["dt"].tolist() == [1e-3, 0.25, 0.25, 0.25, 0.25]
    h2 = np.where(X < 0.5, 2.0, 1.0)
    dt_max = 3e-3
    br = Brick.from_global(h2, z, z, z, 2, 2, 2, 1.0 / n, 1.0 / n)
    rec = oracle_c.brick_run(br, 12, 1e-3, dt_max=dt_max)
    assert rec["retakes"] == 0
    for k in range(1, 12):
        assert rec["dt"][k] == min(rec["dt"][k - 1] * 0.9 / rec["cfl"][k - 1], dt_max)
    assert (rec["dt"][1:] == dt_max).any() and (rec["dt"][1:] < dt_max).any()


def test_correctly_rounded_pow_mode(oracle_c):
    """The oracle's correctly rounded powers (gfo_set_pow_mode(1), the
    device's arithmetic): equal to 60-digit decimal arithmetic, and a patch
    update in that mode is bit for bit the default (libm) update whenever the
    libm's pow rounds correctly on the patch's depths."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    rng = np.random.default_rng(11)
    x = np.exp(rng.uniform(np.log(1e-3), np.log(20.0), 2000))
    try:
        oracle_c.set_pow_mode(True)
        for y in (7.0 / 3.0, 2.0 / 3.0):
            exact = np.array([float(Decimal(float(v)) ** Decimal(y)) for v in x])
            assert np.array_equal(oracle_c.pow(x, y), exact)
        q = np.zeros((36, 36, 3))
        h = 0.1 + rng.random((36, 36))
        q[..., 0] = h
        q[..., 1:] = rng.normal(0, 1, (36, 36, 2)) * h[..., None]
        aux = np.ascontiguousarray(rng.normal(0, 0.2, (36, 36, 1)))
        qa = q.copy()
        oracle_c.patch_update(32, 32, 2, qa, aux, 0.05, 0.05, 0.0, 0.002)
        hc = qa[..., 0].copy()
        cr = oracle_c.pow(hc, 7.0 / 3.0)
        oracle_c.set_pow_mode(False)
        lm = oracle_c.pow(hc, 7.0 / 3.0)
        qb = q.copy()
        oracle_c.patch_update(32, 32, 2, qb, aux, 0.05, 0.05, 0.0, 0.002)
        differ = (cr != lm).reshape(36, 36)[2:-2, 2:-2]
        same_cells = ~differ
        assert same(qa[2:-2, 2:-2][same_cells], qb[2:-2, 2:-2][same_cells])
    finally:
        oracle_c.set_pow_mode(False)
