"""A two-level run on the device against the same pipeline on the oracle.

Layout: [0,1]^2, a 4x4 brick of 16x16 coarse patches (walls outside) whose
centre (or lower-left) 2x2 patches are refined into 16 fine patches (ratio
2), 28 leaves in all.  Every step, both sides fill the ghost cells the way
ForestClaw's ghost update orders them -- same-level face and corner copies,
fine -> coarse averaging, the coarse level's walls, coarse -> fine
interpolation, the fine level's walls -- and advance every leaf with one dt
(no subcycling).  The oracle fills freshly NaN ghost
cells, so a ghost region that no transfer covers would show.  The device
uses gf_pool_ghost_fill, gf_pool_average_ghost, gf_pool_interpolate_ghost
and gf_pool_step; the oracle its bc2, average_ghost, interpolate_ghost and
patch_update.  The order of a whole ghost update is ForestClaw's
(fclaw2d_ghost_update, not in the reference tree), restated here."""
import numpy as np
import pytest

from geoflood_amd import GF_NB_STORED, GF_NB_WALL
from geoflood_amd.pool import Pool
from _parity import assert_same

pytestmark = pytest.mark.gpu

MX, MBC = 16, 2
NC = 4                                   # coarse patches per side
CENTRE = {(1, 1), (2, 1), (1, 2), (2, 2)}
CORNER = {(0, 0), (1, 0), (0, 1), (1, 1)}  # fine patches on two walls
DXC, DXF = 1.0 / (NC * MX), 0.5 / (NC * MX)
DIRS = [(-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1)]


def leaves_multi(blocks):
    """Leaves of a nested layout: blocks[l] is the set of level-l lattice
    patches (i, j) that are refined (each block inside the previous one's
    children, 2:1 balanced).  Returns (level, i, j) leaves, coarse first."""
    out = []
    region = {(i, j) for j in range(NC) for i in range(NC)}
    for lev in range(len(blocks) + 1):
        ref = blocks[lev] if lev < len(blocks) else set()
        out += [(lev, i, j) for (i, j) in sorted(region, key=lambda t: (t[1], t[0])) if (i, j) not in ref]
        region = {(2 * i + ci, 2 * j + cj) for (i, j) in ref for cj in range(2) for ci in range(2)}
    return out


def build_links_multi(lv):
    """Neighbour table and coarse/fine links for leaves of any number of
    levels: a side whose same-level region is a leaf copies it, one covered by
    a coarser leaf or refined further is stored; links join each leaf to the
    finer leaves across its sides (coarse slot, fine slot, side, igrid)."""
    slot = {k: s for s, k in enumerate(lv)}
    nbr = np.full((len(lv), 8), GF_NB_WALL, dtype=np.int32)
    links = []
    for s, (lev, i, j) in enumerate(lv):
        n = NC << lev
        for d, (dx, dy) in enumerate(DIRS):
            a, b = i + dx, j + dy
            if not (0 <= a < n and 0 <= b < n):
                continue
            if (lev, a, b) in slot:
                nbr[s, d] = slot[(lev, a, b)]
                continue
            nbr[s, d] = GF_NB_STORED
            kids = [(ci, cj) for cj in range(2) for ci in range(2)
                    if (dx == 0 or ci == (0 if dx > 0 else 1)) and (dy == 0 or cj == (0 if dy > 0 else 1))]
            for ci, cj in kids:
                f = slot.get((lev + 1, 2 * a + ci, 2 * b + cj))
                if f is not None:
                    links.append((s, f, d, (cj if d < 2 else ci) if d < 4 else 0))
    return nbr, links


def leaves(refined):
    """(level, i, j) of every leaf with the coarse patches in `refined` split
    once: coarse (i, j) on the 4x4 lattice, fine (i, j) on the 8x8 lattice,
    coarse first."""
    return leaves_multi([refined] if refined else [])


def build_links(lv, refined):
    """Neighbour table (GF_NB_* or slot) and the coarse/fine links
    (coarse slot, fine slot, side, igrid) of a two-level layout."""
    return build_links_multi(lv)


def field(x, y):
    """A hump of water near the centre and a dry island over a bump whose
    edge crosses the refined block."""
    b = 0.5 * np.exp(-((x - 0.7) ** 2 + (y - 0.3) ** 2) / 0.02) - 0.1 * x
    h = np.maximum(0.3 + 0.4 * np.exp(-((x - 0.45) ** 2 + (y - 0.55) ** 2) / 0.005) - b, 0.0)
    h[h < 0.05] = 0.0
    return h, b


def initial(lv):
    """Patches of the analytic state with ghost cells (aux ghosts from a
    level-wide padded bathymetry, mirrored at the walls)."""
    qs, auxs = [], []
    for lev, i, j in lv:
        dx = DXC if lev == 0 else DXF
        n = NC * MX * (1 + lev)
        idx = np.arange(-MBC, n + MBC)
        # walls: ghost k outside mirrors interior cell k-1 inside (bc2's rule)
        mir = np.where(idx < 0, -idx - 1, np.where(idx >= n, 2 * n - idx - 1, idx))
        xc = (mir + 0.5) * dx
        X, Y = np.meshgrid(xc, xc)
        h, b = field(X, Y)
        sl = slice(i * MX, i * MX + MX + 2 * MBC), slice(j * MX, j * MX + MX + 2 * MBC)
        q = np.zeros((MX + 2 * MBC, MX + 2 * MBC, 3))
        q[..., 0] = h[sl[1], sl[0]]
        qs.append(q)
        auxs.append(np.ascontiguousarray(b[sl[1], sl[0]][..., None]))
    return np.array(qs), np.array(auxs)


def oracle_ghosts(oc, lv, nbr, links, qs, auxs, check=None):
    interior = (slice(MBC, MBC + MX), slice(MBC, MBC + MX))
    for q in qs:
        g = np.ones(q.shape[:2], bool)
        g[interior] = False
        q[g] = np.nan
    r = range(1, MBC + 1)
    for s in range(len(lv)):                              # same-level faces, then corners
        for d, (dx, dy) in enumerate(DIRS):
            t = nbr[s, d]
            if t < 0:
                continue
            # destination ghost rows/cols and source interior rows/cols (0-based array indices)
            di = list(range(MBC, MBC + MX)) if dx == 0 else ([MBC - k for k in r] if dx < 0 else [MBC + MX - 1 + k for k in r])
            dj = list(range(MBC, MBC + MX)) if dy == 0 else ([MBC - k for k in r] if dy < 0 else [MBC + MX - 1 + k for k in r])
            si = [x - dx * MX for x in di]
            sj = [y - dy * MX for y in dj]
            qs[s][np.ix_(dj, di)] = qs[t][np.ix_(sj, si)]
    def walls(level):
        for s, (lev, i, j) in enumerate(lv):
            if lev == level:
                mth = [3 if nbr[s, k] == GF_NB_WALL else -1 for k in range(4)]
                if max(mth) > 0:
                    oc.bc2(MX, MX, MBC, qs[s], auxs[s], mth)
    for c, f, side, ig in links:
        oc.average_ghost(qs[c], qs[f], auxs[c], auxs[f], side, ig)
    walls(0)                                              # the coarse level's physical sides
    for c, f, side, ig in links:
        oc.interpolate_ghost(qs[c], qs[f], auxs[c], auxs[f], side, ig)
    walls(1)                                              # then the fine level's
    for s in (range(len(qs)) if check is None else check):
        assert np.isfinite(qs[s]).all(), f"a ghost region of slot {s} is not filled"


@pytest.mark.parametrize("where", ["centre", "corner"])
def test_two_level_run_matches_oracle(gf, oracle_c, where):
    """The refined block in the middle of the domain, or in its lower-left
    corner, where fine patches meet the walls next to coarse neighbours (their
    ghost corners between a wall and a coarse/fine side are bc2's mirror of
    the interpolated strip)."""
    refined = CENTRE if where == "centre" else CORNER
    lv = leaves(refined)
    nbr, links = build_links(lv, refined)
    # centre: 8 coarse faces x 2 halves; corners: 4 diagonal coarse patches + 8 face
    # neighbours' inner corners.  corner: 4 coarse faces x 2 halves, 1 + 4 corners
    assert len(lv) == 28 and len(links) == (8 * 2 + 12 if where == "centre" else 4 * 2 + 5)
    n = len(lv)
    qs, auxs = initial(lv)
    dxs = np.array([DXC if l[0] == 0 else DXF for l in lv])
    xlo = np.array([l[1] * MX * d for l, d in zip(lv, dxs)])
    ylo = np.array([l[2] * MX * d for l, d in zip(lv, dxs)])
    pool = Pool(MX, MX, n)
    try:
        pool.set_patches(dxs, dxs, xlo, ylo, nbr)
        pool.upload_aos(0, qs, auxs, buf=0)
        lc, lf, ls, lg = (np.array(x, dtype=np.int32) for x in zip(*links))
        oq = [q.copy() for q in qs]
        oa = [a.copy() for a in auxs]
        dt, t, cur = 4e-4, 0.0, 0
        for step in range(25):
            print(f"case two-level {where} step {step}", flush=True)
            pool.average_ghost(lc, lf, ls, lg, buf=cur)
            pool.interpolate_ghost(lc, lf, ls, lg, buf=cur)
            pool.ghost_fill(buf=cur)
            c_dev, _ = pool.step(dt, src=cur, dst=1 - cur)
            oracle_ghosts(oracle_c, lv, nbr, links, oq, oa)
            c_ref = 0.0
            for s in range(n):
                c_ref = max(c_ref, oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], dxs[s], dxs[s], t, dt))
            assert c_dev == c_ref, step
            assert 0.05 < c_ref < 1.0
            cur, t = 1 - cur, t + dt
        got = pool.download_aos(0, n, buf=cur)
        inner = (slice(MBC, MBC + MX), slice(MBC, MBC + MX))
        g = np.stack([np.moveaxis(got[s][inner], -1, 0) for s in range(n)], axis=1)
        r = np.stack([np.moveaxis(oq[s][inner], -1, 0) for s in range(n)], axis=1)
        assert_same(g, r)
        assert (r[0] > 1e-3).any() and (r[0] <= 1e-3).any()   # wet and dry cells both present
    finally:
        pool.close()


def test_two_level_subcycled_with_time_interpolation(gf, oracle_c):
    """Subcycling: per coarse step dt the fine level takes two steps of dt/2.
    The ghosts at t are filled as above; the coarse level steps into the other
    buffer, the fine level takes its first substep; for the second, the coarse
    level's border layers are interpolated in time (alpha = 1/2,
    gf_pool_timeinterp) into 12 more slots described like the coarse patches,
    whose ghost cells come from their same-level copies, the fine level's
    average at t + dt/2 and the walls, and the fine ghosts are interpolated
    from them; after the second substep gf_pool_copy brings the coarse state
    into the fine level's buffer.  The oracle does the same over freshly NaN
    arrays (the time-interpolated copies hold only their border layers), so a
    read beyond what the time interpolation writes would show."""
    lv = leaves(CENTRE)
    nbr, links = build_links(lv, CENTRE)
    nl = len(lv)
    coarse = [s for s, l in enumerate(lv) if l[0] == 0]
    fine = [s for s, l in enumerate(lv) if l[0] == 1]
    nc, nf = len(coarse), len(fine)
    assert coarse == list(range(nc)) and fine == list(range(nc, nl))
    ti = [nl + s for s in coarse]                        # time-interpolated copies
    nbr_ti = np.where(nbr[:nc] >= 0, nbr[:nc] + nl, nbr[:nc]).astype(np.int32)
    links_ti = [(nl + c, f, side, ig) for c, f, side, ig in links]
    qs, auxs = initial(lv)
    dxs = np.array([DXC if l[0] == 0 else DXF for l in lv])
    xlo = np.array([l[1] * MX * d for l, d in zip(lv, dxs)])
    ylo = np.array([l[2] * MX * d for l, d in zip(lv, dxs)])
    pool = Pool(MX, MX, nl + nc)
    try:
        pool.set_patches(np.concatenate([dxs, dxs[:nc]]), np.concatenate([dxs, dxs[:nc]]),
                         np.concatenate([xlo, xlo[:nc]]), np.concatenate([ylo, ylo[:nc]]),
                         np.concatenate([nbr, nbr_ti]))
        pool.upload_aos(0, qs, auxs, buf=0)
        pool.upload_aos(nl, qs[:nc], auxs[:nc], buf=0)
        pool.upload_aos(nl, qs[:nc], auxs[:nc], buf=1)
        L = [np.array(x, dtype=np.int32) for x in zip(*links)]
        LT = [np.array(x, dtype=np.int32) for x in zip(*links_ti)]
        oq = [q.copy() for q in qs]
        oa = [a.copy() for a in auxs]
        dtc, t, b = 8e-4, 0.0, 0
        dtf = dtc / 2
        psize = 8 * (MX - 2) * 3
        for step in range(10):
            print(f"case two-level subcycled coarse step {step}", flush=True)
            # --- device ---
            pool.average_ghost(*L, buf=b)
            pool.interpolate_ghost(*L, buf=b)
            pool.ghost_fill(0, nl, buf=b)
            cc, _ = pool.step(dtc, 0, nc, src=b, dst=1 - b)
            cf1, _ = pool.step(dtf, nc, nf, src=b, dst=1 - b)
            pool.timeinterp(coarse, b, 1 - b, 0.5, ti, 1 - b)
            pool.average_ghost(*LT, buf=1 - b)
            pool.interpolate_ghost(*LT, buf=1 - b)
            pool.ghost_fill(nc, nf, buf=1 - b)
            cf2, _ = pool.step(dtf, nc, nf, src=1 - b, dst=b)
            pool.copy(coarse, 1 - b, b)
            # --- oracle ---
            oracle_ghosts(oracle_c, lv, nbr, links, oq, oa)
            last = [oq[s].copy() for s in coarse]
            rc = max(oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], DXC, DXC, t, dtc) for s in coarse)
            rf1 = max(oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], DXF, DXF, t, dtf) for s in fine)
            mixed = []
            for k, s in enumerate(coarse):
                q = np.full_like(oq[s], np.nan)
                assert oracle_c.timeinterp(oq[s], last[k], q, 0.5, psize) == 0
                mixed.append(q)
            mixed += [oq[s] for s in fine]
            oracle_ghosts(oracle_c, lv, nbr, links, mixed, oa, check=fine)
            rf2 = max(oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], DXF, DXF, t + dtf, dtf) for s in fine)
            for got, want in ((cc, rc), (cf1, rf1), (cf2, rf2)):
                assert got == want, step
            t += dtc
        got = pool.download_aos(0, nl, buf=b)
        inner = (slice(MBC, MBC + MX), slice(MBC, MBC + MX))
        g = np.stack([np.moveaxis(got[s][inner], -1, 0) for s in range(nl)], axis=1)
        r = np.stack([np.moveaxis(oq[s][inner], -1, 0) for s in range(nl)], axis=1)
        assert_same(g, r)
    finally:
        pool.close()


def test_regrid_cycle_matches_oracle(gf, oracle_c):
    """A dynamic-AMR cycle on resident patches: 4 steps of a uniform 4x4
    brick, refinement of its centre 2x2 (gf_pool_interpolate2fine from the
    parents, whose halo is read through the one-level neighbour table), 6
    steps on two levels with the coarse/fine ghost transfers, coarsening
    back (gf_pool_average2coarse), 4 more steps on one level.  Slots are fixed
    (coarse patch (i, j) in slot 4j+i, the children in slots 16..31), a regrid
    only re-describes them; the oracle runs the same cycle with its own
    interpolate2fine / average2coarse."""
    parents = sorted(CENTRE, key=lambda t: (t[1], t[0]))

    def slot(lev, i, j):
        if lev == 0:
            return j * NC + i
        r = parents.index((i // 2, j // 2))
        return NC * NC + 4 * r + 2 * (j % 2) + (i % 2)

    nslot = NC * NC + 4 * len(parents)
    geo = {}
    for j in range(NC):
        for i in range(NC):
            geo[slot(0, i, j)] = (0, i, j)
    for pi, pj in parents:
        for cj in range(2):
            for ci in range(2):
                geo[slot(1, 2 * pi + ci, 2 * pj + cj)] = (1, 2 * pi + ci, 2 * pj + cj)
    order = [geo[s] for s in range(nslot)]
    qs, auxs = initial(order)                       # slot-ordered arrays; children's q is overwritten
    dxs = np.array([DXC if l[0] == 0 else DXF for l in order])
    xlo = np.array([l[1] * MX * d for l, d in zip(order, dxs)])
    ylo = np.array([l[2] * MX * d for l, d in zip(order, dxs)])

    def phase(refined):
        lvp = leaves(refined)
        nb, links = build_links(lvp, refined)
        sl = [slot(*l) for l in lvp]
        table = np.full((nslot, 8), GF_NB_STORED, dtype=np.int32)  # non-leaves: host-supplied, unused
        for k, s in enumerate(sl):
            table[s] = [sl[v] if v >= 0 else v for v in nb[k]]
        return lvp, nb, links, sl, table

    pool = Pool(MX, MX, nslot)
    try:
        oq = [q.copy() for q in qs]
        oa = [a.copy() for a in auxs]
        pool.set_patches(dxs, dxs, xlo, ylo, phase(set())[4])
        pool.upload_aos(0, qs, auxs, buf=0)
        dt, t, b = 4e-4, 0.0, 0
        kids = [[slot(1, 2 * pi + ci, 2 * pj + cj) for cj in range(2) for ci in range(2)] for pi, pj in parents]
        par = [slot(0, pi, pj) for pi, pj in parents]

        def run(refined, nsteps):
            nonlocal t, b
            lvp, nb, links, sl, table = phase(refined)
            L = [np.array([sl[x] for x in col], dtype=np.int32) if k < 2 else np.array(col, dtype=np.int32)
                 for k, col in enumerate(zip(*links))] if links else None
            for _ in range(nsteps):
                print(f"case regrid cycle t={t:.4f} levels={1 + bool(refined)}", flush=True)
                if L is not None:
                    pool.average_ghost(*L, buf=b)
                    pool.interpolate_ghost(*L, buf=b)
                pool.ghost_fill(0, nslot, buf=b)
                _, cfl = pool.step(dt, 0, nslot, src=b, dst=1 - b)
                views = [oq[s] for s in sl]
                oracle_ghosts(oracle_c, lvp, nb, links, views, [oa[s] for s in sl])
                ref = max(oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], dxs[s], dxs[s], t, dt) for s in sl)
                assert max(cfl[s] for s in sl) == ref
                b, t = 1 - b, t + dt
            return lvp, nb, links, sl

        run(set(), 4)
        # refine: the parents' halo as the one-level step sees it
        lvp, nb, links, sl = phase(set())[:4]
        oracle_ghosts(oracle_c, lvp, nb, links, [oq[s] for s in sl], [oa[s] for s in sl])
        pool.interpolate2fine(par, np.concatenate(kids), buf=b)
        for p, ks in zip(par, kids):
            for g, k in enumerate(ks):
                oracle_c.interpolate2fine(oq[p], oq[k], oa[p], oa[k], g)
        pool.set_patches(dxs, dxs, xlo, ylo, phase(CENTRE)[4])
        run(CENTRE, 6)
        # coarsen
        pool.average2coarse(par, np.concatenate(kids), buf=b)
        for p, ks in zip(par, kids):
            for g, k in enumerate(ks):
                oracle_c.average2coarse(oq[p], oq[k], oa[p], oa[k], g)
        pool.set_patches(dxs, dxs, xlo, ylo, phase(set())[4])
        run(set(), 4)
        got = pool.download_aos(0, NC * NC, buf=b)
        inner = (slice(MBC, MBC + MX), slice(MBC, MBC + MX))
        g = np.stack([np.moveaxis(got[s][inner], -1, 0) for s in range(NC * NC)], axis=1)
        r = np.stack([np.moveaxis(oq[s][inner], -1, 0) for s in range(NC * NC)], axis=1)
        assert_same(g, r)
    finally:
        pool.close()


def test_three_levels_match_oracle(gf, oracle_c):
    """Three levels (the centre 2x2 of the refined block refined again):
    averaging from both finer levels in one call (a middle-level patch is
    coarse for one link and fine for another), interpolation level by level
    from coarse to fine (the middle level's interpolated ghosts feed the
    finest level's), walls per level; 15 steps with one dt."""
    blocks = [CENTRE, {(3, 3), (4, 3), (3, 4), (4, 4)}]
    lv = leaves_multi(blocks)
    nbr, links = build_links_multi(lv)
    levels = np.array([l[0] for l in lv])
    assert sorted(set(levels)) == [0, 1, 2] and len(lv) == 12 + 12 + 16
    by_coarse = {0: [k for k in links if lv[k[0]][0] == 0], 1: [k for k in links if lv[k[0]][0] == 1]}
    n = len(lv)
    dxs = np.array([DXC / (1 << l[0]) for l in lv])
    qs, auxs = [], []
    for lev, i, j in lv:                               # the analytic state on each level's lattice
        dx = dxs[levels.tolist().index(lev)]
        nn = NC * MX << lev
        idx = np.arange(-MBC, nn + MBC)
        mir = np.where(idx < 0, -idx - 1, np.where(idx >= nn, 2 * nn - idx - 1, idx))
        X, Y = np.meshgrid((mir + 0.5) * dx, (mir + 0.5) * dx)
        h, b = field(X, Y)
        sy, sx = slice(j * MX, j * MX + MX + 2 * MBC), slice(i * MX, i * MX + MX + 2 * MBC)
        q = np.zeros((MX + 2 * MBC, MX + 2 * MBC, 3))
        q[..., 0] = h[sy, sx]
        qs.append(q)
        auxs.append(np.ascontiguousarray(b[sy, sx][..., None]))
    xlo = np.array([l[1] * MX * d for l, d in zip(lv, dxs)])
    ylo = np.array([l[2] * MX * d for l, d in zip(lv, dxs)])
    pool = Pool(MX, MX, n)
    try:
        pool.set_patches(dxs, dxs, xlo, ylo, nbr)
        pool.upload_aos(0, np.array(qs), np.array(auxs), buf=0)
        LA = [np.array(x, dtype=np.int32) for x in zip(*links)]
        LI = {c: [np.array(x, dtype=np.int32) for x in zip(*by_coarse[c])] for c in (0, 1)}
        oq = [q.copy() for q in qs]
        oa = [a.copy() for a in auxs]
        dt, t, cur = 2e-4, 0.0, 0
        interior = (slice(MBC, MBC + MX), slice(MBC, MBC + MX))
        for step in range(15):
            print(f"case three levels step {step}", flush=True)
            pool.average_ghost(*LA, buf=cur)
            pool.interpolate_ghost(*LI[0], buf=cur)
            pool.interpolate_ghost(*LI[1], buf=cur)
            pool.ghost_fill(buf=cur)
            c_dev, _ = pool.step(dt, src=cur, dst=1 - cur)
            # oracle: NaN ghosts, same-level copies, averaging, then per level walls and interpolation
            for q in oq:
                g = np.ones(q.shape[:2], bool)
                g[interior] = False
                q[g] = np.nan
            for s in range(n):
                for d, (dx, dy) in enumerate(DIRS):
                    tgt = nbr[s, d]
                    if tgt < 0:
                        continue
                    r = range(1, MBC + 1)
                    di = list(range(MBC, MBC + MX)) if dx == 0 else ([MBC - k for k in r] if dx < 0 else [MBC + MX - 1 + k for k in r])
                    dj = list(range(MBC, MBC + MX)) if dy == 0 else ([MBC - k for k in r] if dy < 0 else [MBC + MX - 1 + k for k in r])
                    oq[s][np.ix_(dj, di)] = oq[tgt][np.ix_([y - dy * MX for y in dj], [x - dx * MX for x in di])]
            for c, f, side, ig in links:
                oracle_c.average_ghost(oq[c], oq[f], oa[c], oa[f], side, ig)
            for lev in (0, 1, 2):
                for s in np.flatnonzero(levels == lev):
                    mth = [3 if nbr[s, k] == GF_NB_WALL else -1 for k in range(4)]
                    if max(mth) > 0:
                        oracle_c.bc2(MX, MX, MBC, oq[s], oa[s], mth)
                if lev < 2:
                    for c, f, side, ig in by_coarse[lev]:
                        oracle_c.interpolate_ghost(oq[c], oq[f], oa[c], oa[f], side, ig)
            assert all(np.isfinite(q).all() for q in oq), "a ghost region is not filled"
            c_ref = max(oracle_c.patch_update(MX, MX, MBC, oq[s], oa[s], dxs[s], dxs[s], t, dt) for s in range(n))
            assert c_dev == c_ref, step
            cur, t = 1 - cur, t + dt
        got = pool.download_aos(0, n, buf=cur)
        g = np.stack([np.moveaxis(got[s][interior], -1, 0) for s in range(n)], axis=1)
        r = np.stack([np.moveaxis(oq[s][interior], -1, 0) for s in range(n)], axis=1)
        assert_same(g, r)
    finally:
        pool.close()
