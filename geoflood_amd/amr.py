"""Resident AMR on one device pool.

ForestClaw owns GeoFlood's forest: p4est refines and coarsens quadtree leaves
at every regrid (fc2d_geoclaw_run.c:569-575, fclaw2d_regrid), and the solver
plugin's hooks tag, interpolate and average patches (fc2d_cpucuda.cpp:
513-694).  For runs without ForestClaw -- the overland-flood and autotuner
workloads of this package, its tests and its benchmark -- this module keeps a
minimal stand-in for that forest and drives the library's device entry points
with it, so every leaf stays resident in the pool from the first step to the
last:

* ``Forest``: the leaves of a quadtree over a brick of level-0 patches,
  levels minlevel..maxlevel, 2:1 balanced across faces and corners (p4est's
  full balance).  It is stored as the level of the covering leaf of every
  patch of the finest lattice; the leaf index, the neighbour table and
  coarse/fine links, and refinement, coarsening of whole families and the
  balance closure are loops over that lattice in a small host library
  (gf_forest.cpp, libgfforest.so).
* ``AmrRun``: one scenario whose leaves occupy a contiguous range of pool
  slots.  Every attempt is one replay of the pool's graph (multi-level ghost
  update, one persistent step of every leaf with one dt -- ForestClaw's
  subcycle = F, the setting of every GeoFlood example -- and the device dt
  control).  Every ``regrid_interval`` accepted steps the leaves are tagged
  on the device (gf_pool_tag_eta_gradient: wet, or |grad eta| above a
  threshold), the forest adapts on the host, and the new forest is laid out
  in the other half of the pool: unchanged leaves are moved, refined ones
  interpolated from their parents, coarsened families averaged, all in the
  scenario's current buffer, before the scenario is re-targeted to the new
  slots.

The order of a whole ghost update (fine -> coarse averaging, then per level:
physical sides, interpolation to the next level) is ForestClaw's
fclaw2d_ghost_update, which is not in the reference tree; it is restated here
and in the library.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import GF_NB_STORED, GF_NB_WALL, last_timing
from .pool import _KIND, MBC, Pool, pad_bathymetry

DIRS = ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1))


_FOREST = None


def _forest_lib():
    """libgfforest.so (gf_forest.cpp): the forest's index, tables and adapt,
    and the neighbour table in pool slots."""
    global _FOREST
    if _FOREST is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgfforest.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (python -m geoflood_amd.build)")
        lib = ct.CDLL(path)
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        bp = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        ci = ct.c_int
        lib.gff_index.argtypes = [ci, ci, ci, ip, ip, ip, ip, ip, ip, ip]
        lib.gff_tables.argtypes = [ci, ci, ci, ip, ip, ci, ip, ip, ip, ip, ci, ci, ip, ip, ip, ip, ip]
        lib.gff_adapt.argtypes = [ci, ci, ci, ci, ip, ip, ci, ip, ip, ip, ip, ip, bp, bp, ip]
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        lib.gff_slot_table.argtypes = [ci, ip, lp, ci, lp, ip]
        for f in (lib.gff_index, lib.gff_tables, lib.gff_adapt, lib.gff_slot_table):
            f.restype = ci
        _FOREST = lib
    return _FOREST


class Forest:
    """Leaves of a quadtree over an nx0 x ny0 brick of level-0 patches.

    ``lev[J, I]`` is the level of the leaf covering patch (I, J) of the
    finest lattice (nx0 2^maxlevel by ny0 2^maxlevel patches).  A leaf
    (L, i, j) covers the finest patches [i s, (i+1) s) x [j s, (j+1) s),
    s = 2^(maxlevel - L).  Leaves are indexed in the order (level, j, i).
    The index, the tables and adapt are native (gf_forest.cpp)."""

    def __init__(self, nx0: int, ny0: int, minlevel: int, maxlevel: int, lev: np.ndarray | None = None):
        assert 0 <= minlevel <= maxlevel <= 15
        self.nx0, self.ny0 = nx0, ny0
        self.minlevel, self.maxlevel = minlevel, maxlevel
        self.nx, self.ny = nx0 << maxlevel, ny0 << maxlevel
        if lev is None:
            lev = np.full((self.ny, self.nx), minlevel, dtype=np.int32)
        self.lev = np.ascontiguousarray(lev, dtype=np.int32)
        assert self.lev.shape == (self.ny, self.nx)
        self._tables = {}
        self._index()

    def _index(self):
        cells = self.nx * self.ny
        L, i, j, oi, oj = (np.empty(cells, dtype=np.int32) for _ in range(5))
        at = np.empty((self.ny, self.nx), dtype=np.int32)
        n = _forest_lib().gff_index(self.nx, self.ny, self.maxlevel, self.lev, L, i, j, oi, oj, at)
        if n < 0:
            raise ValueError("the levels do not form a forest of aligned blocks")
        self._c = (at, L[:n].copy(), oi[:n].copy(), oj[:n].copy(), i[:n].copy(), j[:n].copy())  # int32, native calls
        self.L, self.i, self.j = L[:n].astype(np.int64), i[:n].astype(np.int64), j[:n].astype(np.int64)
        self.oi, self.oj = oi[:n].astype(np.int64), oj[:n].astype(np.int64)   # origins on the finest lattice
        self.leaf_at = at.astype(np.int64)                # covering leaf of every finest patch

    def __len__(self) -> int:
        return len(self.L)

    def counts(self) -> dict:
        return {int(L): int((self.L == L).sum()) for L in np.unique(self.L)}

    def same(self, other: "Forest") -> bool:
        return np.array_equal(self.lev, other.lev)

    # -- adaptation -------------------------------------------------------------------
    def adapt(self, refine: np.ndarray, coarsen: np.ndarray) -> "Forest":
        """The forest after one regrid: leaves with refine[k] split once
        (below maxlevel); families of four sibling leaves that all have
        coarsen[k] (and no refine) merge into their parent (above minlevel);
        then the coarsest 2:1-balanced closure (faces and corners), which may
        refine further leaves or keep a family from merging.  Returns this
        forest when nothing changes."""
        n = len(self.L)
        r = np.asarray(refine, dtype=bool)
        c = np.asarray(coarsen, dtype=bool) & ~r
        if r.shape != (n,) or c.shape != (n,):
            raise ValueError(f"refine and coarsen need one flag per leaf ({n})")
        out = np.empty_like(self.lev)
        at, L, oi, oj, i, j = self._c
        rc = _forest_lib().gff_adapt(self.nx, self.ny, self.minlevel, self.maxlevel, self.lev, at, n, L, i, j, oi, oj,
                                     np.ascontiguousarray(r, dtype=np.uint8), np.ascontiguousarray(c, dtype=np.uint8), out)
        if rc < 0:
            raise MemoryError("gff_adapt: work arrays")
        if rc == 0:
            return self
        return Forest(self.nx0, self.ny0, self.minlevel, self.maxlevel, out)

    # -- tables -------------------------------------------------------------------------
    def tables(self, mthbc=(3, 3, 3, 3)):
        """Neighbour table and coarse/fine links in leaf indices.

        nbr[k, d] for the directions left, right, bottom, top, left-bottom,
        right-bottom, left-top, right-top: the same-level leaf there, the bc2
        kind of the domain side (faces; GF_NB_WALL for corners outside), or
        GF_NB_STORED when the region is covered by a coarser leaf or refined
        further.  links: (coarse leaf, fine leaf, side, igrid) for every
        finer leaf across a side or corner of a coarser one (igrid: the
        fine leaf's half along a face, 0 the lower coordinate)."""
        key = tuple(mthbc)
        if key not in self._tables:
            self._tables[key] = self._build_tables(key)
        return self._tables[key]

    def _build_tables(self, mthbc):
        at, L, oi, oj = self._c[:4]
        n = len(L)
        nbr = np.empty((n, 8), dtype=np.int32)
        lc, lf, ls, lg = (np.empty(12 * max(n, 1), dtype=np.int32) for _ in range(4))
        kinds = np.array([_KIND[mthbc[k]] for k in range(4)], dtype=np.int32)
        nl = _forest_lib().gff_tables(self.nx, self.ny, self.maxlevel, self.lev, at, n, L, oi, oj, kinds,
                                      GF_NB_STORED, GF_NB_WALL, nbr, lc, lf, ls, lg)
        if nl == -2:
            raise MemoryError("gff_tables: work lists")
        if nl < 0:
            raise ValueError("forest is not 2:1 balanced")
        return nbr, (lc[:nl].copy(), lf[:nl].copy(), ls[:nl].copy(), lg[:nl].copy())


def regrid_arrays(old: Forest, new: Forest):
    """How the new forest's leaves arise from the old one's, as arrays:
    (new leaves whose leaf is unchanged, their old leaves); refined families
    (old parents in ascending order, (n, 4) their new children in igrid
    order, x fastest); coarsened families (new parents in ascending order,
    (n, 4) their old children in igrid order).  A regrid changes a leaf's
    level by at most one."""
    M = new.maxlevel
    ko = old.leaf_at[new.oj, new.oi]
    lo = old.L[ko]
    same = np.flatnonzero(lo == new.L)
    child = np.flatnonzero(lo == new.L - 1)
    par = np.flatnonzero(lo == new.L + 1)
    if len(same) + len(child) + len(par) != len(new):
        raise ValueError("a leaf changed level by more than one")
    g = (new.i[child] & 1) + 2 * (new.j[child] & 1)
    order = np.lexsort((g, ko[child]))          # by old parent, then igrid
    kids = child[order].reshape(-1, 4)
    parents = ko[child[order[::4]]]
    if not ((g[order].reshape(-1, 4) == np.arange(4)).all() and (ko[kids] == parents[:, None]).all()):
        raise ValueError("a refined leaf does not have four children")
    h = np.left_shift(1, M - new.L[par] - 1)[:, None]
    ci, cj = np.array([0, 1, 0, 1]), np.array([0, 0, 1, 1])
    ckids = old.leaf_at[new.oj[par][:, None] + cj * h, new.oi[par][:, None] + ci * h]
    return (same, ko[same]), (parents, kids), (par, ckids)


def regrid_plan(old: Forest, new: Forest):
    """regrid_arrays as lists: unchanged leaves (new, old), refined families
    [(old parent, [four new children])] and coarsened families [(new parent,
    [four old children])]."""
    moved, (rp, rk), (cp, ck) = regrid_arrays(old, new)
    refined = [(int(p), [int(k) for k in ks]) for p, ks in zip(rp, rk)]
    coarsened = [(int(p), [int(k) for k in ks]) for p, ks in zip(cp, ck)]
    return moved, refined, coarsened


class LevelDem:
    """Bathymetry of each level on its own lattice, with the ghost ring of
    every patch: ``dem_fn(nx_cells, ny_cells, level)`` gives the level's
    interior cell values, padded by bc2's rule at the domain sides
    (pad_bathymetry), so a patch's aux is a slice of its level's array."""

    def __init__(self, forest: Forest, mx: int, dem_fn, mthbc):
        self.mx = mx
        self.pad = {}
        for L in range(forest.maxlevel + 1):
            nx, ny = (forest.nx0 << L) * mx, (forest.ny0 << L) * mx
            self.pad[L] = pad_bathymetry(np.asarray(dem_fn(nx, ny, L), dtype=np.float64), mthbc)

    def aux(self, L, i, j) -> np.ndarray:
        """(n, mx+4, mx+4, 1) aux of leaves (L[k], i[k], j[k])."""
        m = self.mx
        out = np.empty((len(L), m + 2 * MBC, m + 2 * MBC, 1))
        for k, (l, a, b) in enumerate(zip(L, i, j)):
            out[k, ..., 0] = self.pad[int(l)][b * m:b * m + m + 2 * MBC, a * m:a * m + m + 2 * MBC]
        return out


class AmrRun:
    """One scenario of resident AMR in a device pool of 2 x capacity slots
    (two halves: the current forest in one, a regrid builds the next in the
    other).

    mx: patch size; (nx0, ny0): level-0 patches over [0, lx] x [0, ly];
    dem_fn(nx_cells, ny_cells, level) -> (ny, nx) bathymetry of a level;
    mthbc: bc2 kinds of the four sides (3 wall, 1 extrapolation, 0
    hydrograph, whose table ``hydro`` = dict(time, hu, b, x0, y0) is given);
    eta_tol: the |grad eta| refinement threshold."""

    def __init__(self, mx, nx0, ny0, lx, ly, minlevel, maxlevel, dem_fn, capacity, mthbc=(3, 3, 3, 3),
                 hydro=None, eta_tol=0.01, regrid_interval=4, manning_n=None, dt_max=0.0, flags_fn=None,
                 tune_cache=None, log=None):
        self.mx, self.lx, self.ly = mx, lx, ly
        # flags_fn(forest, tags, regrid_index) -> (refine, coarsen) per leaf;
        # default: refine the tagged leaves, coarsen families of untagged ones
        self.flags_fn = flags_fn
        # tune_cache (geoflood_amd.autotune.TuneCache): every fourth regrid,
        # when the active population per attempt has moved to another bucket,
        # re-select the step variant (cached choice, or an autotune)
        self.tune = tune_cache
        self._tune_bucket = None
        self._tune_stats = None
        self._retune_due = False
        self.mthbc, self.hydro, self.eta_tol = tuple(mthbc), hydro, eta_tol
        self.regrid_interval, self.manning_n, self.dt_max = regrid_interval, manning_n, dt_max
        self.capacity = capacity
        self.forest = Forest(nx0, ny0, minlevel, maxlevel)
        self.dem = LevelDem(self.forest, mx, dem_fn, mthbc)
        self.pool = Pool(mx, mx, 2 * capacity)
        self.half = 0            # the forest occupies slots [half * capacity, ...)
        self.perm = None         # slot offset of each leaf
        self.sid = None
        self.stats = dict(regrids=0, attempts=0, steps=0, retakes=0, cell_updates=0, patches=[], device_ms=0.0)
        self.tag_log = []        # the tags of every regrid, in leaf order
        # log(record): called once per regrid with a JSON-serialisable dict
        # (the time, dt, CFL and step counters at the regrid, the leaves per
        # level after it, whether the forest changed) -- the run's line of
        # ForestClaw's per-step production output (fc2d_geoclaw_run.c:522-527)
        self.log = log

    # -- layout -----------------------------------------------------------------------
    def _dx(self, L):
        """Cell sizes (dx, dy) of level L, or of every level of an integer
        array L (the same IEEE division per element)."""
        if np.ndim(L):
            L = np.asarray(L, dtype=np.int64)
            return (self.lx / ((np.left_shift(self.forest.nx0, L) * self.mx).astype(np.float64)),
                    self.ly / ((np.left_shift(self.forest.ny0, L) * self.mx).astype(np.float64)))
        return self.lx / ((self.forest.nx0 << L) * self.mx), self.ly / ((self.forest.ny0 << L) * self.mx)

    def cell_centres(self, L, i, j):
        """Cell centres (x, y) of leaves (L[k], i[k], j[k]), ghost cells
        included: two arrays of shape (n, mx+4, mx+4)."""
        m = self.mx
        c = np.arange(-MBC, m + MBC) + 0.5
        dx, dy = self._dx(np.asarray(L))
        dx, dy = dx[:, None], dy[:, None]
        xs = (np.asarray(i)[:, None] * m + c[None, :]) * dx
        ys = (np.asarray(j)[:, None] * m + c[None, :]) * dy
        return np.broadcast_to(xs[:, None, :], (len(L), m + 2 * MBC, m + 2 * MBC)), \
            np.broadcast_to(ys[:, :, None], (len(L), m + 2 * MBC, m + 2 * MBC))

    def _retune(self):
        """Re-select the step variant when the active population per attempt
        since the last check has moved to another bucket."""
        from .autotune import TuneCache
        st = self.pool.stats()
        prev, self._tune_stats = self._tune_stats, st
        if prev is None or st["attempts"] <= prev["attempts"]:
            return
        active = (st["active_patches"] - prev["active_patches"]) / (st["attempts"] - prev["attempts"])
        b = TuneCache.bucket(active)
        if b != self._tune_bucket:
            self.tune.choose(self.pool, active, levels=self.forest.counts())
            self._tune_bucket = b
            self._tune_stats = self.pool.stats()  # the candidates' launches are not attempts of the run

    def _describe(self, forest: Forest, base: int, perm: np.ndarray):
        """Describe the slots [base, base + len(forest)) of a forest whose
        leaf k sits in slot base + perm[k]; store its levels."""
        n = len(forest)
        if n > self.capacity:
            raise RuntimeError(f"forest of {n} leaves exceeds the capacity {self.capacity}")
        nbr, _ = forest.tables(self.mthbc)
        inv = np.empty(n, dtype=np.int64)  # leaf of each slot offset
        nb = np.empty((n, 8), dtype=np.int32)
        if _forest_lib().gff_slot_table(n, nbr, np.ascontiguousarray(perm, dtype=np.int64), base, inv, nb) < 0:
            raise ValueError("perm is not a permutation of the leaves")
        L, i, j = forest.L[inv], forest.i[inv], forest.j[inv]
        dx, dy = self._dx(L)
        xlo, ylo = i * self.mx * dx, j * self.mx * dy
        from . import check, dptr, iptr
        d = [np.ascontiguousarray(a, dtype=np.float64) for a in (dx, dy, xlo, ylo)]
        check(self.pool.lib.gf_pool_describe(self.pool.h, base, n, *[dptr(a) for a in d], iptr(np.ascontiguousarray(nb))),
              "gf_pool_describe")
        self.pool.n = max(self.pool.n, base + n)
        self.pool.set_levels(base, L.astype(np.int32))
        return inv

    def _links(self, forest: Forest, base: int, perm: np.ndarray):
        _, (c, f, s, g) = forest.tables(self.mthbc)
        self.pool.set_cf_links(base + perm[c], base + perm[f], s, g)

    def leaves_state(self, buf=None) -> tuple[np.ndarray, np.ndarray]:
        """(q, aux) of every leaf in leaf order (ghost cells included)."""
        info = self.pool.scenario(self.sid)
        base = self.half * self.capacity
        q = self.pool.download_aos(base, len(self.forest), buf=info["cur"] if buf is None else buf)
        a = self.pool.download_aux(base, len(self.forest))
        return q[self.perm], a[self.perm]

    # -- set-up -----------------------------------------------------------------------
    def start(self, t0=0.0, dt0=1.0, desired_cfl=0.9, max_cfl=1.0, state_fn=None, init_levels=True):
        """Register the scenario on the minlevel forest, then refine it by the
        tagging criterion on the initial state until it no longer changes
        (ForestClaw's initialisation: new leaves take the initial state, not
        an interpolation), alternating the pool halves as a regrid does.
        state_fn(L, i, j, aux) -> q (n, mx+4, mx+4, 3) gives the initial
        state (default: dry)."""
        f = self.forest
        self._upload_forest(f, 0, state_fn)
        self.sid = self.pool.lib.gf_pool_add_scenario(self.pool.h, 0, len(f), t0, dt0, desired_cfl, max_cfl, 0, 0.0)
        if self.sid < 0:
            from . import check
            check(self.sid, "gf_pool_add_scenario")
        if self.manning_n is not None:
            self.pool.set_friction(self.sid, self.manning_n)
        if self.dt_max > 0:
            self.pool.set_dt_max(self.sid, self.dt_max)
        if self.hydro is not None:
            h = self.hydro
            self.pool.set_inflow(self.sid, h["time"], h["hu"], h["b"], h["x0"], h["y0"])
        while init_levels:
            tags = self._tags(f, 0, t0)
            g = f.adapt(tags.astype(bool), np.zeros(len(f), dtype=bool))
            if g.same(f):
                break
            other = 1 - self.half
            self._upload_forest(g, other * self.capacity, state_fn)
            self.pool.set_scenario_slots(self.sid, other * self.capacity, len(g))
            self.half = other
            f = g
        self.stats["patches"].append(len(f))

    def _upload_forest(self, f: Forest, base: int, state_fn):
        self.perm = np.arange(len(f))
        self._describe(f, base, self.perm)
        aux = self.dem.aux(f.L, f.i, f.j)
        q = np.zeros(aux.shape[:3] + (3,)) if state_fn is None else state_fn(f.L, f.i, f.j, aux)
        self.pool.upload_aos(base, q, aux, buf=0)
        self._links(f, base, self.perm)
        self.forest = f

    def _tags(self, f: Forest, buf: int, t: float) -> np.ndarray:
        base = self.half * self.capacity
        self.pool.ghost_update(buf, t)
        return self.pool.tag_eta_gradient(base + self.perm, self.eta_tol, buf=buf)

    # -- stepping -----------------------------------------------------------------------
    def regrid(self, info=None):
        """Tag the current leaves, adapt the forest and lay it out in the
        other half of the pool (state in the scenario's current buffer).
        info: the scenario's current record, when the caller has just read it."""
        if info is None:
            info = self.pool.scenario(self.sid)
        cur, t = info["cur"], info["t"]
        old = self.forest
        tags = self._tags(old, cur, t).astype(bool)
        self.tag_log.append(tags)
        refine, coarsen = (tags, ~tags) if self.flags_fn is None else self.flags_fn(old, tags, self.stats["regrids"])
        new = old.adapt(refine, coarsen)
        self.stats["regrids"] += 1
        if self.log is not None:
            self.log(dict(regrid=self.stats["regrids"], t=float(t), dt=float(info["dt"]),
                          last_dt=float(info["last_dt"]), maxcfl=float(info["maxcfl"]), nsteps=int(info["nsteps"]),
                          retakes=int(info["retakes"]), attempts=int(info["attempts"]), tagged=int(tags.sum()),
                          leaves=len(new), levels={str(k): v for k, v in new.counts().items()},
                          changed=not new.same(old)))
        if new.same(old):
            return False
        ob, nb_ = self.half * self.capacity, (1 - self.half) * self.capacity
        (keep_new, keep_old), (rpar, rkids), (cpar, ckids) = regrid_arrays(old, new)
        # new slot order: unchanged leaves first, then the created ones
        fresh = np.ones(len(new), dtype=bool)
        fresh[keep_new] = False
        created = np.flatnonzero(fresh)
        perm = np.empty(len(new), dtype=np.int64)
        perm[keep_new] = np.arange(len(keep_new))
        perm[created] = len(keep_new) + np.arange(len(created))
        self._describe(new, nb_, perm)
        if len(created):
            aux = self.dem.aux(new.L[created], new.i[created], new.j[created])
            self.pool.upload_aos(nb_ + len(keep_new), np.zeros(aux.shape[:3] + (3,)), aux, buf=cur)
        if len(keep_new):
            self.pool.move(ob + self.perm[keep_old], nb_ + perm[keep_new], buf=cur)
        if len(rpar):
            self.pool.interpolate2fine(ob + self.perm[rpar], (nb_ + perm[rkids]).reshape(-1), buf=cur)
        if len(cpar):
            self.pool.average2coarse(nb_ + perm[cpar], (ob + self.perm[ckids]).reshape(-1), buf=cur)
        self.pool.set_scenario_slots(self.sid, nb_, len(new))
        self._links(new, nb_, perm)
        self.forest, self.perm, self.half = new, perm, 1 - self.half
        return True

    def run(self, steps: int, graph: bool = True) -> dict:
        """Advance by ``steps`` accepted steps, regridding every
        regrid_interval accepted steps."""
        info = self.pool.scenario(self.sid)
        target = info["nsteps"] + steps
        while info["nsteps"] < target:
            before = info["nsteps"]
            nxt = min(target, (before // self.regrid_interval + 1) * self.regrid_interval)
            if self._retune_due:
                # the autotuner times the patches of active scenarios only: set
                # the next stop first (the scenario stopped at the regrid step)
                self.pool.set_stop(self.sid, nxt)
                self._retune()
                self._retune_due = False
            # the scenario stops on the device at the next regrid step: replay
            # that many attempts, then single ones for any retakes
            if graph:
                info = self.pool.advance(self.sid, nxt)
                self.stats["device_ms"] += last_timing()[0]
            else:
                self.pool.set_stop(self.sid, nxt)
                k = nxt - before
                while True:
                    self.pool.run_plain(k)
                    self.stats["device_ms"] += last_timing()[0]
                    info = self.pool.scenario(self.sid)
                    if info["status"] != 0 or info["nsteps"] >= nxt:
                        break
                    k = 1
            if info["status"] != 0:
                raise RuntimeError(f"scenario stopped with status {info['status']} at t = {info['t']}")
            self.stats["attempts"] = info["attempts"]
            cells = len(self.forest) * self.mx * self.mx
            self.stats["cell_updates"] += (info["nsteps"] - before) * cells
            if info["nsteps"] % self.regrid_interval == 0:
                self.regrid(info)
                self.stats["patches"].append(len(self.forest))
                if self.tune is not None and self.stats["regrids"] % 4 == 1:
                    self._retune_due = True
        self.stats["steps"] = info["nsteps"]
        self.stats["retakes"] = info["retakes"]
        return info

    def close(self):
        self.pool.close()
