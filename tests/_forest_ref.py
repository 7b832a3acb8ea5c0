"""The numpy forest: the reference of the native one (geoflood_amd.amr.Forest
over gf_forest.cpp).  Test infrastructure only.

Functions of a forest f with levels f.lev on the finest lattice (f.nx x
f.ny, levels f.minlevel..f.maxlevel) and its leaves f.L, f.i, f.j, f.oi,
f.oj, f.leaf_at: the leaf index, the neighbour table and coarse/fine
links, and one regrid's levels -- pyramid block maxima for the tree, the
family merge over pyramids of minima, maxima and flags, the balance
closure, and a shortcut that keeps the forest's own levels when no leaf
can change.  tests/test_amr_forest.py checks these against block-by-block
forms and the native forest against these."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from geoflood_amd import GF_NB_STORED, GF_NB_WALL
from geoflood_amd.amr import DIRS
from geoflood_amd.pool import _KIND


def _block(a: np.ndarray, s: int) -> np.ndarray:
    """View of a 2-D array as (ny/s, s, nx/s, s) blocks."""
    return a.reshape(a.shape[0] // s, s, a.shape[1] // s, s)


def _pyramid(a: np.ndarray, levels: int, op) -> list:
    """[a, op over 2x2 blocks of a, op over 2x2 blocks of that, ...]:
    levels + 1 arrays, each half the size of the last in both directions."""
    out = [a]
    for _ in range(levels):
        b = out[-1]
        out.append(op(op(b[0::2, 0::2], b[1::2, 0::2]), op(b[0::2, 1::2], b[1::2, 1::2])))
    return out


def ref_index(lev: np.ndarray, M: int) -> SimpleNamespace:
    """Leaves of the levels lev (finest lattice, levels <= M) in the order
    (L, j, i): L, i, j, origins oi, oj, and leaf_at of every finest patch."""
    lev = np.asarray(lev, dtype=np.int64)
    ny, nx = lev.shape
    m = np.left_shift(1, M - lev) - 1                  # block size - 1
    I = np.arange(nx)[None, :]
    J = np.arange(ny)[:, None]
    jj, ii = np.nonzero(((I & m) == 0) & ((J & m) == 0))
    L = lev[jj, ii]
    i, j = ii >> (M - L), jj >> (M - L)
    order = np.argsort((L << 42) | (j << 21) | i, kind="stable")
    ids = np.full((ny, nx), -1, dtype=np.int64)
    ids[jj[order], ii[order]] = np.arange(len(order))
    return SimpleNamespace(L=L[order], i=i[order], j=j[order], oi=ii[order], oj=jj[order],
                           leaf_at=ids[J & ~m, I & ~m])


def ref_tables(f, mthbc):
    """Neighbour table (n, 8) and links (coarse, fine, side, igrid) in the
    native order: by direction, then the fine leaf's half, then coarse leaf."""
    M = f.maxlevel
    L, oi, oj = f.L, f.oi, f.oj
    n = len(L)
    s = np.left_shift(1, M - L)
    ids = np.arange(n)
    nbr = np.empty((n, 8), dtype=np.int32)
    kinds = [_KIND[mthbc[k]] for k in range(4)]
    lc, lf, ls, lg = [], [], [], []
    for d, (dx, dy) in enumerate(DIRS):
        ni, nj = oi + dx * s, oj + dy * s
        inside = (ni >= 0) & (ni < f.nx) & (nj >= 0) & (nj < f.ny)
        nic, njc = np.clip(ni, 0, f.nx - 1), np.clip(nj, 0, f.ny - 1)
        ln = f.lev[njc, nic]
        same = inside & (ln == L)
        out = kinds[d] if d < 4 else GF_NB_WALL
        nbr[:, d] = np.where(same, f.leaf_at[njc, nic], np.where(inside, GF_NB_STORED, out))
        finer = inside & (ln > L)
        if not finer.any():
            continue
        cis = [0] if dx > 0 else [1] if dx < 0 else [0, 1]
        cjs = [0] if dy > 0 else [1] if dy < 0 else [0, 1]
        h = s // 2
        for cj in cjs:
            for ci in cis:
                ki = np.where(finer, nic + ci * h, 0)
                kj = np.where(finer, njc + cj * h, 0)
                assert (f.lev[kj[finer], ki[finer]] == L[finer] + 1).all(), "forest is not 2:1 balanced"
                lc.append(ids[finer])
                lf.append(f.leaf_at[kj[finer], ki[finer]])
                ls.append(np.full(finer.sum(), d))
                lg.append(np.full(finer.sum(), (cj if d < 2 else ci) if d < 4 else 0))
    cat = (lambda v: np.concatenate(v).astype(np.int32)) if lc else (lambda v: np.zeros(0, np.int32))
    return nbr, (cat(lc), cat(lf), cat(ls), cat(lg))


def ref_tree(f, req: np.ndarray) -> np.ndarray:
    """The coarsest quadtree whose every finest patch is at least at level
    req there, from pyramid block maxima."""
    M = f.maxlevel
    mx = _pyramid(req, M, np.maximum)      # mx[M - L]: block maxima at level L
    lev = np.where(mx[M] <= 0, 0, -1)
    for L in range(1, M + 1):
        lev = np.repeat(np.repeat(lev, 2, axis=0), 2, axis=1)
        lev = np.where((lev < 0) & (mx[M - L] <= L), L, lev)
    return lev


def ref_balance(f, req: np.ndarray) -> np.ndarray:
    """The coarsest 2:1-balanced tree (faces and corners) at least as fine
    as req."""
    req = np.maximum(req, f.minlevel)
    while True:
        lev = ref_tree(f, req)
        p = np.pad(lev, 1, constant_values=-1)
        nb = np.max([p[1 + dy:1 + dy + f.ny, 1 + dx:1 + dx + f.nx] for dx, dy in DIRS], axis=0)
        need = nb - 1
        if (lev >= need).all():
            return lev
        req = np.maximum(req, need)


def ref_adapt_full(f, refine: np.ndarray, coarsen: np.ndarray) -> np.ndarray:
    """One regrid's levels, always through the merge map and the closure
    (coarsen already excludes refine)."""
    M = f.maxlevel
    lev = f.lev
    req = lev + (refine[f.leaf_at] & (lev < M))
    cc = coarsen[f.leaf_at]
    mn = _pyramid(lev, M, np.minimum)
    mxp = _pyramid(lev, M, np.maximum)
    al = _pyramid(cc, M, np.logical_and)
    target = np.full(mn[M].shape, -1, dtype=np.int64)   # at level 0 resolution
    for P in range(0, M):                                # parent level P = L - 1
        if P > 0:
            target = np.repeat(np.repeat(target, 2, axis=0), 2, axis=1)
        L = P + 1
        if L > f.minlevel:
            k = M - P
            target = np.where((mn[k] == L) & (mxp[k] == L) & al[k], P, target)
    target = np.repeat(np.repeat(target, 2, axis=0), 2, axis=1)
    req = np.where(target >= 0, target, req)
    return ref_balance(f, req)


def ref_changes(f, refine: np.ndarray, coarsen: np.ndarray) -> bool:
    """Whether a regrid can change the forest.  It cannot when no leaf below
    maxlevel is to refine and every family of four siblings (level L >
    minlevel) that is to coarsen has a face or corner neighbour of its
    parent's block finer than L: merging it would put level L-1 next to
    level L+1, which the closure undoes."""
    M = f.maxlevel
    if (refine & (f.L < M)).any():
        return True
    c = coarsen & (f.L > f.minlevel)
    if not c.any():
        return False
    key = (f.L[c] << 48) | ((f.j[c] >> 1) << 24) | (f.i[c] >> 1)
    keys, counts = np.unique(key, return_counts=True)
    fams = keys[counts == 4]
    if len(fams) == 0:
        return False
    p = np.pad(f.lev, 1, constant_values=-1)
    nb = np.max([p[1 + dy:1 + dy + f.ny, 1 + dx:1 + dx + f.nx] for dx, dy in DIRS], axis=0)
    FL, fj, fi = fams >> 48, (fams >> 24) & 0xFFFFFF, fams & 0xFFFFFF
    for L in np.unique(FL):
        sel = FL == L
        finest_near = _block(nb, 1 << (M - int(L) + 1)).max(axis=(1, 3))
        if (finest_near[fj[sel], fi[sel]] <= L).any():
            return True
    return False


def ref_adapt(f, refine, coarsen) -> np.ndarray:
    """One regrid's levels (refine and coarsen per leaf)."""
    refine = np.asarray(refine, dtype=bool)
    coarsen = np.asarray(coarsen, dtype=bool) & ~refine
    if not ref_changes(f, refine, coarsen):
        return np.asarray(f.lev, dtype=np.int64).copy()
    return ref_adapt_full(f, refine, coarsen)


def ref_regrid_plan(old, new):
    """Unchanged, refined and coarsened leaves of a regrid, leaf by leaf:
    (new leaves kept, their old leaves), {old parent: new children by
    igrid}, {new parent: old children by igrid}."""
    M = new.maxlevel
    kept_new, kept_old, refined, coarsened = [], [], {}, {}
    for k in range(len(new)):
        ko = int(old.leaf_at[new.oj[k], new.oi[k]])
        if old.L[ko] == new.L[k]:
            kept_new.append(k)
            kept_old.append(ko)
        elif old.L[ko] == new.L[k] - 1:
            refined.setdefault(ko, [None] * 4)[int(new.i[k] & 1) + 2 * int(new.j[k] & 1)] = k
        else:
            assert old.L[ko] == new.L[k] + 1
            h = 1 << (M - int(new.L[k]) - 1)
            coarsened[k] = [int(old.leaf_at[new.oj[k] + cj * h, new.oi[k] + ci * h]) for cj in range(2) for ci in range(2)]
    return (kept_new, kept_old), refined, coarsened
