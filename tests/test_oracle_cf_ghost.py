"""The oracle's coarse/fine ghost transfers (fc2d_geoclaw_average_fort.f
average_face / average_corner and fc2d_geoclaw_interpolate_fort.f
interpolate_face / interpolate_corner) for half-size neighbours of the same
orientation.

The reference's routines take their index map from ForestClaw
(fclaw2d_clawpatch_transform_face_half / _corner_half), which the reference
tree does not hold, so they cannot be built here.  Their per-cell arithmetic
is the regrid arithmetic: a coarse ghost cell averaged from four fine cells
is the cell that average2coarse (pinned against the reference build) gives
a coarse patch lying over the fine neighbour, and a fine ghost cell is the
value interpolate2fine gives the child of the coarse patch over the same
place, with the depth clipped at 0.  These tests hold the oracle to both
identities, to the geometry (which cells each side writes, the hanging
corners included) and to what the averaging and a limited linear
reconstruction fix exactly (flat-bed means, linear fields)."""
import numpy as np
import pytest

from oracle.oracle import CLib, Params, ref_amr_lib

MBC = 2
FACE_SIDES = [(s, g) for s in range(4) for g in range(2)]
CORNER_SIDES = [(s, 0) for s in range(4, 8)]


@pytest.fixture(scope="module")
def clib():
    c = CLib()
    c.set_params(Params())
    return c


def origin(side, igrid, mx):
    """The fine neighbour covers coarse cells ox+1..ox+mx/2 (same in y)."""
    h = mx // 2
    return {0: (-h, igrid * h), 1: (mx, igrid * h), 2: (igrid * h, -h), 3: (igrid * h, mx),
            4: (-h, -h), 5: (mx, -h), 6: (-h, mx), 7: (mx, mx)}[side]


def lattice(rng, mx, lo, n):
    """A field on coarse cell indices lo..lo+n-1 in both directions."""
    return lo, rng.normal(0, 0.3, (n, n))


def sample(field, i0, j0, mx):
    """The (mx+4)^2 AoS aux array of a patch whose cell (1,1) is lattice
    cell (i0+1, j0+1)."""
    lo, f = field
    idx = np.arange(1 - MBC, mx + MBC + 1)
    return np.ascontiguousarray(f[np.ix_(idx + j0 - lo, idx + i0 - lo)][..., None])


def random_q(rng, mx, wet=0.7):
    sh = (mx + 4, mx + 4)
    q = rng.normal(0, 1, sh + (3,))
    q[..., 0] = np.abs(q[..., 0]) * (rng.random(sh) < wet)
    return q


def ghost_strip(side, mx):
    """Coarse ghost cells (ic, jc) an average over `side` may write."""
    r = range(1, MBC + 1)
    if side == 0:
        return [(1 - b, j) for j in range(1, mx + 1) for b in r]
    if side == 1:
        return [(mx + b, j) for j in range(1, mx + 1) for b in r]
    if side == 2:
        return [(i, 1 - b) for b in r for i in range(1, mx + 1)]
    if side == 3:
        return [(i, mx + b) for b in r for i in range(1, mx + 1)]
    c = side - 4
    return [((mx + a) if c & 1 else (1 - a), (mx + b) if c & 2 else (1 - b)) for a in r for b in r]


def at(q, i, j):
    return q[j - 1 + MBC, i - 1 + MBC]


@pytest.mark.parametrize("mx", [8, 16, 32])
def test_average_ghost_is_average2coarse_over_the_neighbour(clib, mx):
    """Every coarse ghost cell the fine neighbour covers equals, bit for bit,
    the cell of quadrant 0 that average2coarse writes into a coarse patch V
    whose quadrant 0 is the fine neighbour (V's bathymetry is the coarse
    lattice's there); every other cell is left alone.  average2coarse is
    the reference build's when oracle/_ref is present."""
    ref = ref_amr_lib()
    if ref is not None:
        ref.set_dry_tolerance(1e-3)
    avg = ref if ref is not None else clib
    rng = np.random.default_rng(mx)
    bc = lattice(rng, mx, -mx - 4, 3 * mx + 8)
    for side, igrid in FACE_SIDES + CORNER_SIDES:
        ox, oy = origin(side, igrid, mx)
        qc, qf, qv = random_q(rng, mx), random_q(rng, mx), random_q(rng, mx)
        ac = sample(bc, 0, 0, mx)
        af = np.ascontiguousarray(rng.normal(0, 0.3, (mx + 4, mx + 4, 1)))
        av = sample(bc, ox, oy, mx)
        out = qc.copy()
        clib.average_ghost(out, qf, ac, af, side, igrid)
        avg.average2coarse(qv, qf, av, af, 0)
        written = {(i, j) for (i, j) in ghost_strip(side, mx)
                   if 1 <= i - ox <= mx // 2 and 1 <= j - oy <= mx // 2}
        assert len(written) == (MBC * mx // 2 if side < 4 else MBC * MBC)
        for i in range(1 - MBC, mx + MBC + 1):
            for j in range(1 - MBC, mx + MBC + 1):
                if (i, j) in written:
                    assert np.array_equal(at(out, i, j), at(qv, i - ox, j - oy)), (side, igrid, i, j)
                else:
                    assert np.array_equal(at(out, i, j), at(qc, i, j)), (side, igrid, i, j)


@pytest.mark.parametrize("mx", [8, 16, 32])
def test_interpolate_ghost_is_interpolate2fine_of_the_child(clib, mx):
    """Every fine ghost cell the coarse patch covers equals the same place
    in the child of the coarse patch that interpolate2fine fills (bathymetry
    from one fine lattice), with the depth clipped at 0; nothing else of the
    fine patch changes."""
    rng = np.random.default_rng(100 + mx)
    bf = lattice(rng, mx, -2 * mx - 8, 6 * mx + 16)
    for side, igrid in FACE_SIDES + CORNER_SIDES:
        ox, oy = origin(side, igrid, mx)
        qc = random_q(rng, mx)
        ac = np.ascontiguousarray(rng.normal(0, 0.3, (mx + 4, mx + 4, 1)))
        qf = random_q(rng, mx)
        af = sample(bf, 2 * ox, 2 * oy, mx)  # fine lattice: fine cell (1,1) of F at 2*ox+1
        out = qf.copy()
        clib.interpolate_ghost(qc, out, ac, af, side, igrid)
        kids = []
        for g in range(4):
            ig, jg = g % 2, g // 2
            qk = random_q(rng, mx)
            ak = sample(bf, ig * mx, jg * mx, mx)
            clib.interpolate2fine(qc, qk, ac, ak, g)
            kids.append(qk)
        changed = 0
        for i in range(1 - MBC, mx + MBC + 1):
            for j in range(1 - MBC, mx + MBC + 1):
                # global fine index of F's cell (i, j): 2*ox + i; its coarse cell:
                gi, gj = 2 * ox + i, 2 * oy + j
                ic, jc = (gi + 1) // 2, (gj + 1) // 2
                interior_f = 1 <= i <= mx and 1 <= j <= mx
                if side < 4:
                    along = jc if side < 2 else ic
                    across = ic if side < 2 else jc
                    edge = {0: 1, 1: mx, 2: 1, 3: mx}[side]
                    hit = (not interior_f and across == edge and 1 <= along <= mx)
                else:
                    c = side - 4
                    hit = (not interior_f and ic == (mx if c & 1 else 1) and jc == (mx if c & 2 else 1))
                if hit:
                    ig, jg = int(ic > mx // 2), int(jc > mx // 2)
                    ki, kj = gi - ig * mx, gj - jg * mx
                    want = at(kids[2 * jg + ig], ki, kj).copy()
                    want[0] = want[0] if want[0] > 0.0 else 0.0
                    assert np.array_equal(at(out, i, j), want), (side, igrid, i, j)
                    changed += 1
                else:
                    assert np.array_equal(at(out, i, j), at(qf, i, j)), (side, igrid, i, j)
        assert changed > 0


def test_interpolate_face_writes_the_hanging_corners(clib):
    """Face interpolation fills the fine neighbour's whole ghost strip along
    the face plus the 2x2 ghost corner on the side away from its sibling (the
    coarse loop runs over jc = 1..mx): lower half -> upper corner, upper
    half -> lower corner."""
    mx = 16
    rng = np.random.default_rng(9)
    qc = random_q(rng, mx, wet=1.0)
    ac = np.zeros((mx + 4, mx + 4, 1))
    for side in range(4):
        for igrid in range(2):
            qf = np.full((mx + 4, mx + 4, 3), np.nan)
            qf[2:-2, 2:-2] = 1.0
            clib.interpolate_ghost(qc, qf, ac, ac, side, igrid)
            got = {(i, j) for i in range(-1, mx + 3) for j in range(-1, mx + 3)
                   if not (1 <= i <= mx and 1 <= j <= mx) and np.isfinite(at(qf, i, j)).all()}
            strip = {0: (mx + 1, mx + 2), 1: (-1, 0), 2: (mx + 1, mx + 2), 3: (-1, 0)}[side]
            tang = range(1, mx + 3) if igrid == 0 else range(-1, mx + 1)
            if side < 2:
                want = {(i, j) for i in strip for j in tang}
            else:
                want = {(i, j) for j in strip for i in tang}
            assert got == want, (side, igrid)


def test_average_ghost_flat_bed_means(clib):
    """Wet fine cells over a flat bed: the coarse ghost depth is the mean of
    the four fine depths and the momentum their sum scaled by hc/hsum (to
    rounding: the mean momentum); dry fine cells (h <= drytol) give a dry
    coarse ghost cell."""
    mx = 16
    rng = np.random.default_rng(4)
    qf = np.zeros((mx + 4, mx + 4, 3))
    qf[..., 0] = rng.uniform(0.5, 2.0, qf.shape[:2])
    qf[..., 1:] = rng.normal(0, 1, qf.shape[:2] + (2,))
    z = np.zeros((mx + 4, mx + 4, 1))
    for side, igrid in FACE_SIDES + CORNER_SIDES:
        ox, oy = origin(side, igrid, mx)
        qc = np.full_like(qf, -7.0)
        clib.average_ghost(qc, qf, z, z, side, igrid)
        for (i, j) in ghost_strip(side, mx):
            if not (1 <= i - ox <= mx // 2 and 1 <= j - oy <= mx // 2):
                continue
            fi, fj = 2 * (i - ox) - 1, 2 * (j - oy) - 1
            four = np.array([at(qf, fi + s, fj + t) for t in (0, 1) for s in (0, 1)])
            np.testing.assert_allclose(at(qc, i, j), four.mean(axis=0), rtol=1e-14, atol=1e-15)
    dry = qf.copy()
    dry[..., 0] = 1e-3
    qc = np.full_like(qf, -7.0)
    clib.average_ghost(qc, dry, z, z, 1, 0)
    assert np.all(at(qc, mx + 1, 1) == 0.0)


def test_interpolate_ghost_reproduces_linear_fields(clib):
    """A linear surface over a flat bed and linear momenta (slopes equal on
    both sides, so the limiter passes them) are reproduced exactly at the
    fine ghost cell centres: this fixes the index map, the shifts and their
    signs."""
    mx = 16
    idx = np.arange(1 - MBC, mx + MBC + 1) - 0.5  # coarse cell centres, unit spacing
    X, Y = np.meshgrid(idx, idx)
    qc = np.zeros((mx + 4, mx + 4, 3))
    qc[..., 0] = 50.0 + 0.7 * X - 0.3 * Y
    qc[..., 1] = 2.0 - 0.25 * X + 0.5 * Y
    qc[..., 2] = -1.0 + 0.125 * X + 0.375 * Y
    z = np.zeros((mx + 4, mx + 4, 1))
    for side, igrid in FACE_SIDES + CORNER_SIDES:
        ox, oy = origin(side, igrid, mx)
        qf = np.full_like(qc, np.nan)
        clib.interpolate_ghost(qc, qf, z, z, side, igrid)
        n = 0
        for i in range(1 - MBC, mx + MBC + 1):
            for j in range(1 - MBC, mx + MBC + 1):
                v = at(qf, i, j)
                if not np.isfinite(v).all():
                    continue
                # fine cell centre in coarse units
                x = ox + (i - 0.5) / 2.0
                y = oy + (j - 0.5) / 2.0
                want = [50.0 + 0.7 * x - 0.3 * y, 2.0 - 0.25 * x + 0.5 * y, -1.0 + 0.125 * x + 0.375 * y]
                np.testing.assert_allclose(v, want, rtol=0, atol=1e-12)
                n += 1
        assert n == ((mx + 2) * MBC if side < 4 else MBC * MBC)


def test_bad_arguments(clib):
    q = np.zeros((12, 12, 3))
    a = np.zeros((12, 12, 1))
    with pytest.raises(ValueError):
        clib.average_ghost(q, q.copy(), a, a, 8, 0)
    with pytest.raises(ValueError):
        clib.interpolate_ghost(q, q.copy(), a, a, 0, 2)
    q2 = np.zeros((12, 20, 3))
    a2 = np.zeros((12, 20, 1))
    with pytest.raises(ValueError):
        clib.interpolate_ghost(q2, q2.copy(), a2, a2, 1, 0)
