"""The resident-AMR forest (geoflood_amd.amr.Forest) and the oracle's AMR
cycle on the CPU: a quadtree stand-in for ForestClaw's p4est forest.

Checked against what a forest must satisfy by construction: the leaves
partition the domain into aligned blocks; levels stay in [minlevel,
maxlevel]; adjacent leaves (faces and corners) differ by at most one level
(p4est's full 2:1 balance); a tagged leaf is refined, a family merges only
when all four siblings agree and the balance allows; the neighbour table is
symmetric; every finer leaf across a coarser leaf's side has a link; and in
the oracle's cycle every ghost cell of every leaf is filled by the ghost
update (a ghost left NaN fails), across refinement and coarsening."""
import numpy as np
import pytest

import _forest_ref as R
from geoflood_amd import GF_NB_STORED
from geoflood_amd.amr import DIRS, Forest, regrid_arrays, regrid_plan


def _check(f: Forest):
    M = f.maxlevel
    assert f.lev.min() >= f.minlevel and f.lev.max() <= M
    # partition: each finest patch in exactly one leaf, leaves aligned
    cover = np.zeros_like(f.lev)
    for L, oi, oj in zip(f.L, f.oi, f.oj):
        s = 1 << (M - L)
        assert oi % s == 0 and oj % s == 0
        cover[oj:oj + s, oi:oi + s] += 1
        assert (f.lev[oj:oj + s, oi:oi + s] == L).all()
    assert (cover == 1).all()
    # full balance
    p = np.pad(f.lev, 1, constant_values=-100)
    for dx, dy in DIRS:
        nb = p[1 + dy:1 + dy + f.ny, 1 + dx:1 + dx + f.nx]
        ok = nb >= 0
        assert (np.abs(nb - f.lev)[ok] <= 1).all()


def _random_forest(rng, steps=6):
    f = Forest(3, 2, 1, 4)
    for _ in range(steps):
        ref = rng.random(len(f)) < 0.3
        coa = rng.random(len(f)) < 0.6
        f = f.adapt(ref, coa)
        _check(f)
    return f


def test_forest_balance_and_partition():
    rng = np.random.default_rng(0)
    for _ in range(20):
        _random_forest(rng)


def test_adapt_refines_tags_and_coarsens_whole_families():
    f = Forest(2, 2, 0, 3)
    g = f.adapt(np.array([True, False, False, False]), np.zeros(4, bool))
    assert g.counts() == {0: 3, 1: 4}
    # refining one level-1 leaf of the corner family forces nothing at level 0 (2:1 holds)
    h = g.adapt(np.array([False] * 3 + [True, False, False, False]), np.zeros(7, bool))
    _check(h)
    assert h.counts() == {0: 3, 1: 3, 2: 4}
    # three siblings agreeing are not enough to merge
    k = np.flatnonzero(g.L == 1)
    coa = np.zeros(len(g), bool)
    coa[k[:3]] = True
    assert g.adapt(np.zeros(len(g), bool), coa).same(g)
    coa[k] = True
    assert g.adapt(np.zeros(len(g), bool), coa).same(f)
    # balance keeps a family from merging next to a finer leaf
    coa = np.zeros(len(h), bool)
    coa[h.L == 1] = True
    assert h.adapt(np.zeros(len(h), bool), coa).counts() == h.counts()


def test_tables_symmetric_and_links_complete():
    rng = np.random.default_rng(1)
    f = _random_forest(rng)
    nbr, (lc, lf, ls, lg) = f.tables((3, 1, 3, 3))
    opp = [1, 0, 3, 2, 7, 6, 5, 4]
    for k in range(len(f)):
        for d in range(8):
            t = nbr[k, d]
            if t >= 0:
                assert nbr[t, opp[d]] == k and f.L[t] == f.L[k]
    # a stored side of a leaf that has a coarser leaf across it is the fine end of a link
    fine_sides = {(int(b), opp[int(s)]) for b, s in zip(lf, ls)}
    for k in range(len(f)):
        for d, (dx, dy) in enumerate(DIRS[:4]):
            if nbr[k, d] != GF_NB_STORED:
                continue
            s = 1 << (f.maxlevel - f.L[k])
            ni, nj = f.oi[k] + dx * s, f.oj[k] + dy * s
            if f.lev[nj, ni] < f.L[k]:
                assert (k, d) in fine_sides
    assert (f.L[lf] == f.L[lc] + 1).all()


def test_regrid_plan_accounts_for_every_leaf():
    rng = np.random.default_rng(2)
    f = _random_forest(rng, 3)
    for _ in range(5):
        g = f.adapt(rng.random(len(f)) < 0.3, rng.random(len(f)) < 0.6)
        (kn, ko), refined, coarsened = regrid_plan(f, g)
        covered = set(kn.tolist()) | {k for _, ks in refined for k in ks} | {p for p, _ in coarsened}
        assert covered == set(range(len(g)))
        assert all(g.L[k] == f.L[p] + 1 for p, ks in refined for k in ks)
        assert all(f.L[k] == g.L[p] + 1 for p, ks in coarsened for k in ks)
        f = g


def test_regrid_arrays_equal_the_leaf_by_leaf_plan():
    """The vectorised regrid plan (AmrRun.regrid's) equals a leaf-by-leaf
    restatement over regrids that refine and coarsen, with families in
    ascending parent order and children in igrid order."""
    rng = np.random.default_rng(5)
    f = _random_forest(rng, 4)
    seen = [0, 0]
    for _ in range(8):
        g = f.adapt(rng.random(len(f)) < 0.25, rng.random(len(f)) < 0.7)
        (kn, ko), (rp, rk), (cp, ck) = regrid_arrays(f, g)
        (rkn, rko), rref, cref = R.ref_regrid_plan(f, g)
        np.testing.assert_array_equal(kn, rkn)
        np.testing.assert_array_equal(ko, rko)
        assert rp.tolist() == sorted(rref) and rk.tolist() == [rref[p] for p in sorted(rref)]
        assert cp.tolist() == sorted(cref) and ck.tolist() == [cref[p] for p in sorted(cref)]
        seen[0] += len(rp)
        seen[1] += len(cp)
        f = g
    assert seen[0] > 0 and seen[1] > 0


def band_flags(width=0.25, speed=0.15):
    """Refine the leaves whose centre lies in a vertical band moving east one
    step per regrid, coarsen the others: every regrid refines and coarsens."""
    def flags(forest, tags, k):
        x = (forest.i + 0.5) / (forest.nx0 << forest.L)
        lo = (0.05 + speed * k) % 1.0
        inb = (x >= lo) & (x < lo + width)
        return inb, ~inb
    return flags


def test_oracle_amr_cycle_fills_every_ghost(oracle_c):
    """A walled basin with a hump of water, levels 0-2 over a 2 x 2 brick of
    8 x 8 patches, a refinement band moving across it: 24 steps, a regrid
    every 2 (each refines and coarsens), every ghost update fully covered,
    depths non-negative throughout."""
    from oracle.amr_driver import OracleAmr

    def dem(nx, ny, L):
        x = (np.arange(nx) + 0.5) / nx
        X, Y = np.meshgrid(x, x)
        return 0.2 * np.exp(-((X - 0.6) ** 2 + (Y - 0.4) ** 2) / 0.02)

    def state(L, i, j, aux):
        q = np.zeros(aux.shape[:3] + (3,))
        q[..., 0] = np.maximum(0.5 - aux[..., 0], 0.0)
        return q

    r = OracleAmr(oracle_c, 8, 2, 2, 1.0, 1.0, 0, 2, dem, eta_tol=1e9, regrid_interval=2,
                  flags_fn=band_flags())
    r.start(0.0, 1e-3, state_fn=state, init_levels=False)
    r.run(24)
    assert r.nregrids == 12 and len(set(r.patches)) > 2
    assert (r.state()[:, 0] >= 0).all()


def test_eta_gradient_criterion_closed_forms():
    """The surface-gradient tags of the oracle (no counterpart in the
    reference): a linear surface eta = a x + c y flags exactly when
    sqrt(a^2 + c^2) exceeds the threshold (away from equality); any wet
    cell flags; a bump in one ghost cell reaches only the interior cell next
    to it; a flat dry bed never flags."""
    from oracle.amr_driver import eta_gradient_tags
    mx, dx = 8, 0.5
    c = (np.arange(-2, mx + 2) + 0.5) * dx
    X, Y = np.meshgrid(c, c)
    q = np.zeros((mx + 4, mx + 4, 3))
    for a, cc, tol in [(0.03, 0.04, 0.049), (0.03, 0.04, 0.051), (0.0, 0.02, 0.0199), (0.0, 0.02, 0.0201)]:
        aux = (a * X + cc * Y)[..., None]
        assert eta_gradient_tags(q, aux, mx, dx, dx, 1e-3, tol) == (np.hypot(a, cc) > tol)
    flat = np.zeros((mx + 4, mx + 4, 1))
    assert not eta_gradient_tags(q, flat, mx, dx, dx, 1e-3, 1e-6)
    w = q.copy()
    w[5, 6, 0] = 2e-3                       # one wet interior cell
    assert eta_gradient_tags(w, flat, mx, dx, dx, 1e-3, 1e9)
    g = q.copy()
    g[0, 5, 0] = 1.0                        # outer ghost layer: no interior difference reaches it
    assert not eta_gradient_tags(g, flat, mx, dx, dx, 1e-3, 0.1)
    g[1, 5, 0] = 1.0                        # inner ghost layer: (eta(i,1) - eta(i,-1)) / (2 dy) = -1
    assert eta_gradient_tags(g, flat, mx, dx, dx, 1e-3, 0.99)
    assert not eta_gradient_tags(g, flat, mx, dx, dx, 1e-3, 1.01)


def _random_forests(rng, trials):
    for trial in range(trials):
        f = Forest(int(rng.integers(1, 4)), int(rng.integers(1, 4)), int(rng.integers(0, 2)), int(rng.integers(1, 5)))
        for _ in range(int(rng.integers(0, 5))):
            t = rng.random(len(f)) < rng.uniform(0.05, 0.6)
            f = f.adapt(t, ~t & (rng.random(len(f)) < 0.5))
        yield trial, f


def test_native_forest_equals_numpy_reference():
    """The native forest (gf_forest.cpp) against the numpy reference
    (tests/_forest_ref.py) on random forests and flags: the same leaves in
    the same order, the same neighbour table and links in the same order
    (two sets of domain sides), and the same levels after adapt -- through
    the merge map and the closure, and the reference's shortcut where no
    leaf can change -- with the forest itself returned exactly when its
    levels stay.  Flags that refine only leaves at maxlevel and coarsen
    families kept by a finer neighbour exercise the unchanged case."""
    rng = np.random.default_rng(11)
    same = changed = 0
    for trial, f in _random_forests(rng, 300):
        ix = R.ref_index(f.lev, f.maxlevel)
        for k in ("L", "i", "j", "oi", "oj", "leaf_at"):
            assert np.array_equal(getattr(ix, k), getattr(f, k)), (trial, k)
        for mthbc in ((3, 3, 3, 3), (0, 1, 3, 1)):
            nb, links = f.tables(mthbc)
            rnb, rlinks = R.ref_tables(f, mthbc)
            assert np.array_equal(nb, rnb), trial
            for a, b in zip(links, rlinks):
                assert np.array_equal(a, b), trial
        for _ in range(5):
            refine = rng.random(len(f)) < rng.choice([0.0, 0.01, 0.3])
            if rng.random() < 0.7:
                refine &= f.L == f.maxlevel          # only leaves that cannot refine
            coarsen = ~refine & (rng.random(len(f)) < rng.choice([0.0, 0.5, 0.9, 1.0]))
            got = f.adapt(refine, coarsen)
            ref = R.ref_adapt(f, refine, coarsen)
            assert np.array_equal(got.lev, ref), trial
            assert np.array_equal(ref, R.ref_adapt_full(f, refine, ~refine & coarsen)), trial
            if np.array_equal(ref, f.lev):
                assert got is f
                same += 1
            else:
                assert got is not f
                changed += 1
    assert same > 0 and changed > 0


def _ref_tree(f, req):
    """The block-by-block form of the tree (each level's block maxima taken
    over the finest lattice): the reference for the pyramid form."""
    M = f.maxlevel
    lev = np.full_like(req, M)
    done = np.zeros(req.shape, dtype=bool)
    for L in range(M + 1):
        s = 1 << (M - L)
        b = req.reshape(req.shape[0] // s, s, req.shape[1] // s, s).max(axis=(1, 3))
        cond = np.repeat(np.repeat(b <= L, s, axis=0), s, axis=1)
        lev[cond & ~done] = L
        done |= cond
    return lev


def _ref_adapt_levels(f, refine, coarsen):
    """One regrid's levels written block by block, with a balance closure
    over _ref_tree."""
    M = f.maxlevel
    lev = f.lev.astype(np.int64)
    req = lev + (refine[f.leaf_at] & (lev < M))
    cc = coarsen[f.leaf_at]
    for L in range(f.minlevel + 1, M + 1):
        sp = 1 << (M - L + 1)
        b = lev.reshape(lev.shape[0] // sp, sp, lev.shape[1] // sp, sp)
        c = cc.reshape(cc.shape[0] // sp, sp, cc.shape[1] // sp, sp)
        fam = (b.min(axis=(1, 3)) == L) & (b.max(axis=(1, 3)) == L) & c.all(axis=(1, 3))
        req[np.repeat(np.repeat(fam, sp, axis=0), sp, axis=1)] = L - 1
    req = np.maximum(req, f.minlevel)
    while True:
        t = _ref_tree(f, req)
        p = np.pad(t, 1, constant_values=-1)
        nb = np.max([p[1 + dy:1 + dy + f.ny, 1 + dx:1 + dx + f.nx] for dx, dy in DIRS], axis=0)
        if (t >= nb - 1).all():
            return t
        req = np.maximum(req, nb - 1)


def test_pyramid_forest_equals_blockwise_reference():
    """The pyramid forms of the tree, the family merge and the balance
    closure (the numpy reference) give the same levels as the block-by-block
    forms on random forests, requests and flags; the native forest equals
    both."""
    rng = np.random.default_rng(21)
    for trial, f in _random_forests(rng, 120):
        req = rng.integers(0, f.maxlevel + 1, size=f.lev.shape)
        assert np.array_equal(R.ref_tree(f, req), _ref_tree(f, req)), trial
        refine = rng.random(len(f)) < rng.choice([0.0, 0.1, 0.4])
        coarsen = ~refine & (rng.random(len(f)) < rng.choice([0.3, 0.8, 1.0]))
        blockwise = _ref_adapt_levels(f, refine, coarsen)
        assert np.array_equal(R.ref_adapt_full(f, refine, coarsen), blockwise), trial
        assert np.array_equal(f.adapt(refine, coarsen).lev, blockwise), trial


def test_native_forest_refuses_non_forests():
    """Levels that are not aligned blocks, flags of the wrong length, and a
    2:1 violation in the neighbour tables are errors, not results."""
    lev = np.zeros((4, 4), dtype=np.int32)
    lev[0, 1] = 1                                   # a level-1 value inside a level-0 block
    with pytest.raises(ValueError):
        Forest(1, 1, 0, 2, lev)
    f = Forest(2, 2, 0, 2)
    with pytest.raises(ValueError):
        f.adapt(np.zeros(3, bool), np.zeros(4, bool))
    lev = np.zeros((8, 8), dtype=np.int32)
    lev[:4, :4] = 2                                 # level 2 next to level 0: not balanced
    g = Forest(2, 2, 0, 2, lev)
    with pytest.raises(ValueError, match="balanced"):
        g.tables()


def test_slot_table_equals_the_numpy_layout():
    """gff_slot_table (the neighbour table in pool slots, AmrRun._describe)
    equals the numpy gather and remap it replaces, for random leaf-to-slot
    permutations of adapted forests; a non-permutation is refused."""
    from geoflood_amd.amr import _forest_lib
    rng = np.random.default_rng(11)
    f = Forest(3, 2, 1, 4)
    for it in range(6):
        r = rng.random(len(f)) < 0.3
        f = f.adapt(r, ~r & (rng.random(len(f)) < 0.5))
        nbr, _ = f.tables((0, 1, 3, 3))
        n = len(f)
        perm = rng.permutation(n).astype(np.int64)
        base = int(rng.integers(0, 1000))
        inv_ref = np.empty(n, dtype=np.int64)
        inv_ref[perm] = np.arange(n)
        nb_ref = nbr[inv_ref]
        nb_ref = np.where(nb_ref >= 0, base + perm[np.maximum(nb_ref, 0)], nb_ref).astype(np.int32)
        inv = np.empty(n, dtype=np.int64)
        nb = np.empty((n, 8), dtype=np.int32)
        assert _forest_lib().gff_slot_table(n, nbr, perm, base, inv, nb) == 0
        np.testing.assert_array_equal(inv, inv_ref)
        np.testing.assert_array_equal(nb, nb_ref)
        assert (nb_ref < 0).any() and (nb_ref >= base).any()
    bad = perm.copy()
    bad[0] = bad[1]
    assert _forest_lib().gff_slot_table(n, nbr, bad, 0, inv, nb) == -1
    big = nbr.copy()
    big[0, 0] = n
    assert _forest_lib().gff_slot_table(n, big, perm, 0, inv, nb) == -1
