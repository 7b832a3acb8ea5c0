"""Resident-AMR runs on the oracle (test infrastructure).

The same cycle as geoflood_amd.amr.AmrRun on host patches: the same Forest
(leaves, neighbour table, coarse/fine links, regrid plan) and the same level
bathymetry, with the oracle's routines in place of the device kernels:

  ghost update (ForestClaw's fclaw2d_ghost_update with time_interp off, as
  restated for this package): every ghost cell set to NaN, same-level face
  and corner copies, fine -> coarse averaging over all links
  (gfo_average_ghost), then for each level from coarse to fine its bc2 sides
  (gfo_bc2_ex: walls, extrapolation, hydrograph at time t) and the
  interpolation to the next finer level (gfo_interpolate_ghost); a ghost cell
  left NaN fails the run;
  step: gfo_patch_update of every leaf with one dt, the outstyle_3 rule
  (fc2d_geoclaw_run.c:417-590) with retakes and the dt bound;
  regrid every regrid_interval accepted steps: ghost update, the
  surface-gradient tags (eta_gradient_tags), Forest.adapt, then unchanged
  leaves copied, refined ones from gfo_interpolate2fine of their parent,
  coarsened families from gfo_average2coarse.

Only tests and the benchmark's CPU leg use this module; it is the checker.
"""
from __future__ import annotations

import numpy as np

from geoflood_amd import GF_NB_EXTRAP, GF_NB_INFLOW, GF_NB_WALL
from geoflood_amd.amr import DIRS, Forest, LevelDem, regrid_plan
from oracle.oracle import Hydrograph

MBC = 2


def eta_gradient_tags(q, aux, mx, dx, dy, drytol, eta_tol) -> bool:
    """The surface-gradient criterion on one patch with filled ghost cells:
    any interior cell wet (h > drytol) or with
    ((eta(i+1,j) - eta(i-1,j)) / (2 dx))^2 + ((eta(i,j+1) - eta(i,j-1)) / (2 dy))^2 > eta_tol^2,
    eta = h + b (q, aux arrays (my+4, mx+4, .), row j, column i)."""
    h = q[..., 0]
    eta = h + aux[..., 0]
    s = slice(MBC, MBC + mx)
    wet = h[s, s] > drytol
    gx = (eta[s, MBC + 1:MBC + 1 + mx] - eta[s, MBC - 1:MBC - 1 + mx]) / (2.0 * dx)
    gy = (eta[MBC + 1:MBC + 1 + mx, s] - eta[MBC - 1:MBC - 1 + mx, s]) / (2.0 * dy)
    return bool(wet.any() or (gx * gx + gy * gy > eta_tol * eta_tol).any())


class OracleAmr:
    """The oracle twin of geoflood_amd.amr.AmrRun (same arguments, plus the
    oracle library ``oc`` with its parameters set)."""

    def __init__(self, oc, mx, nx0, ny0, lx, ly, minlevel, maxlevel, dem_fn, mthbc=(3, 3, 3, 3), hydro=None,
                 eta_tol=0.01, regrid_interval=4, dt_max=0.0, drytol=1e-3, flags_fn=None):
        self.oc, self.mx, self.lx, self.ly = oc, mx, lx, ly
        self.flags_fn, self.nregrids = flags_fn, 0
        self.mthbc, self.eta_tol, self.regrid_interval = tuple(mthbc), eta_tol, regrid_interval
        self.dt_max, self.drytol = dt_max, drytol
        self.hyd = None
        if hydro is not None:
            sides = tuple(1 if m == 0 else 0 for m in self.mthbc)
            self.hyd = Hydrograph(hydro["time"], hydro["hu"], hydro["b"], hydro["x0"], hydro["y0"], sides)
        self.forest = Forest(nx0, ny0, minlevel, maxlevel)
        self.dem = LevelDem(self.forest, mx, dem_fn, mthbc)
        self.q, self.aux = [], []
        self.t = self.dt = 0.0
        self.desired_cfl, self.max_cfl = 0.9, 1.0
        self.nsteps = self.retakes = self.attempts = 0
        self.dts, self.cfls, self.patches, self.tag_log = [], [], [], []

    def _dx(self, L):
        return self.lx / ((self.forest.nx0 << L) * self.mx), self.ly / ((self.forest.ny0 << L) * self.mx)

    def _fresh(self, f: Forest, state_fn):
        aux = self.dem.aux(f.L, f.i, f.j)
        q = np.zeros(aux.shape[:3] + (3,)) if state_fn is None else state_fn(f.L, f.i, f.j, aux)
        return [np.ascontiguousarray(x) for x in q], [np.ascontiguousarray(a) for a in aux]

    # -- ghost update -------------------------------------------------------------------
    def ghost_update(self, t):
        f, mx = self.forest, self.mx
        nbr, (lc, lf, ls, lg) = f.tables(self.mthbc)
        inner = (slice(MBC, MBC + mx), slice(MBC, MBC + mx))
        for q in self.q:
            g = np.ones(q.shape[:2], bool)
            g[inner] = False
            q[g] = np.nan
        r = range(1, MBC + 1)
        for s in range(len(f)):
            for d, (dx, dy) in enumerate(DIRS):
                tg = nbr[s, d]
                if tg < 0:
                    continue
                di = list(range(MBC, MBC + mx)) if dx == 0 else ([MBC - k for k in r] if dx < 0 else [MBC + mx - 1 + k for k in r])
                dj = list(range(MBC, MBC + mx)) if dy == 0 else ([MBC - k for k in r] if dy < 0 else [MBC + mx - 1 + k for k in r])
                self.q[s][np.ix_(dj, di)] = self.q[tg][np.ix_([y - dy * mx for y in dj], [x - dx * mx for x in di])]
        for c, fi, sd, ig in zip(lc, lf, ls, lg):
            self.oc.average_ghost(self.q[c], self.q[fi], self.aux[c], self.aux[fi], int(sd), int(ig))
        code = {GF_NB_WALL: 3, GF_NB_EXTRAP: 1, GF_NB_INFLOW: 0}
        for lev in range(f.maxlevel + 1):
            for s in np.flatnonzero(f.L == lev):
                mth = [code.get(int(nbr[s, k]), -1) for k in range(4)]
                if max(mth) >= 0:
                    dx, dy = self._dx(lev)
                    rc = self.oc.bc2_ex(mx, mx, MBC, self.q[s], self.aux[s], mth, f.i[s] * mx * dx, f.j[s] * mx * dy,
                                        dx, dy, t, self.hyd)
                    assert rc == 0, "bc2 failed"
            for c, fi, sd, ig in zip(lc, lf, ls, lg):
                if f.L[c] == lev:
                    self.oc.interpolate_ghost(self.q[c], self.q[fi], self.aux[c], self.aux[fi], int(sd), int(ig))
        for s, q in enumerate(self.q):
            assert np.isfinite(q).all(), f"a ghost region of leaf {s} (level {f.L[s]}) is not filled"

    def tags(self, t) -> np.ndarray:
        self.ghost_update(t)
        f = self.forest
        out = np.zeros(len(f), dtype=bool)
        for s in range(len(f)):
            dx, dy = self._dx(f.L[s])
            out[s] = eta_gradient_tags(self.q[s], self.aux[s], self.mx, dx, dy, self.drytol, self.eta_tol)
        return out

    # -- set-up and regrid --------------------------------------------------------------
    def start(self, t0=0.0, dt0=1.0, desired_cfl=0.9, max_cfl=1.0, state_fn=None, init_levels=True):
        self.t, self.dt, self.desired_cfl, self.max_cfl = t0, dt0, desired_cfl, max_cfl
        f = self.forest
        self.q, self.aux = self._fresh(f, state_fn)
        while init_levels:
            tags = self.tags(t0)
            g = f.adapt(tags, np.zeros(len(f), dtype=bool))
            if g.same(f):
                break
            self.forest = f = g
            self.q, self.aux = self._fresh(f, state_fn)
        self.patches.append(len(f))

    def regrid(self):
        old = self.forest
        tags = self.tags(self.t)
        refine, coarsen = (tags, ~tags) if self.flags_fn is None else self.flags_fn(old, tags, self.nregrids)
        new = old.adapt(refine, coarsen)
        self.nregrids += 1
        self.tag_log.append(tags)
        if new.same(old):
            return False
        (keep_new, keep_old), refined, coarsened = regrid_plan(old, new)
        aux = self.dem.aux(new.L, new.i, new.j)
        q = [np.zeros(a.shape[:2] + (3,)) for a in aux]
        au = [np.ascontiguousarray(a) for a in aux]
        for kn, ko in zip(keep_new, keep_old):
            q[kn] = self.q[ko].copy()
            au[kn] = self.aux[ko].copy()
        for p, kids in refined:
            for g, k in enumerate(kids):
                self.oc.interpolate2fine(self.q[p], q[k], self.aux[p], au[k], g)
        for p, kids in coarsened:
            for g, k in enumerate(kids):
                self.oc.average2coarse(q[p], self.q[k], au[p], self.aux[k], g)
        self.forest, self.q, self.aux = new, q, au
        return True

    # -- stepping -------------------------------------------------------------------------
    def attempt(self):
        self.ghost_update(self.t)
        f = self.forest
        saved = [q.copy() for q in self.q]
        cfl = 0.0
        for s in range(len(f)):
            dx, dy = self._dx(f.L[s])
            cfl = max(cfl, self.oc.patch_update(self.mx, self.mx, MBC, self.q[s], self.aux[s], dx, dy, self.t, self.dt))
        self.attempts += 1
        if cfl > self.max_cfl:
            self.q = saved
            self.retakes += 1
            self.dt = self.dt * self.desired_cfl / cfl
            return False
        self.t += self.dt
        self.dts.append(self.dt)
        self.cfls.append(cfl)
        self.nsteps += 1
        if self.dt_max > 0:
            d = self.dt * self.desired_cfl / cfl if cfl > 0 else self.dt
            self.dt = self.dt_max if (cfl == 0.0 or d > self.dt_max) else d
        else:
            self.dt = self.dt * self.desired_cfl / cfl
        return True

    def run(self, steps):
        target = self.nsteps + steps
        while self.nsteps < target:
            if self.attempt() and self.nsteps % self.regrid_interval == 0:
                self.regrid()
                self.patches.append(len(self.forest))

    def state(self):
        """Interior (h, hu, hv) of every leaf, leaf order: (n, 3, mx, mx)."""
        s = slice(MBC, MBC + self.mx)
        return np.stack([np.moveaxis(q[s, s], -1, 0) for q in self.q])
