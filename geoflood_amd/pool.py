"""Device-resident patch pools and the drop-in batch step (Python side of
the C ABI in include/geoflood_amd.h).

A *brick* is a uniform single-level block of mi x mj patches of mx x my cells
(the forest of the uniform configurations and of each ensemble scenario).
Global interior fields are arrays of shape (mj*my, mi*mx); patch (pi, pj) has
pool slot first + pj*mi + pi.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from . import (GF_NB_EXTRAP, GF_NB_INFLOW, GF_NB_WALL, GfRefine, GfScenarioInfo, check, dptr, iptr,
               load)

MBC = 2
_KIND = {3: GF_NB_WALL, 1: GF_NB_EXTRAP, 0: GF_NB_INFLOW}


@dataclass
class BrickLayout:
    mi: int
    mj: int
    mx: int
    my: int
    dx: float
    dy: float
    mthbc: tuple = (3, 3, 3, 3)  # left, right, bottom, top (bc2: 3 wall, 1 extrapolation, 0 inflow)
    first: int = 0
    xlo: float = 0.0  # lower-left corner of the brick
    ylo: float = 0.0

    @property
    def npatch(self) -> int:
        return self.mi * self.mj

    def origins(self):
        """Per-patch lower-left corners (xlower, ylower)."""
        pi, pj = np.meshgrid(np.arange(self.mi), np.arange(self.mj))
        return (np.ascontiguousarray(self.xlo + pi.ravel() * self.mx * self.dx, dtype=np.float64),
                np.ascontiguousarray(self.ylo + pj.ravel() * self.my * self.dy, dtype=np.float64))

    def neighbours(self) -> np.ndarray:
        """(npatch, 8) neighbour table: left, right, bottom, top, left-bottom,
        right-bottom, left-top, right-top; pool indices or GF_NB_* codes."""
        mi, mj, f = self.mi, self.mj, self.first
        pi, pj = np.meshgrid(np.arange(mi), np.arange(mj))
        pi, pj = pi.ravel(), pj.ravel()
        nb = np.empty((mi * mj, 8), dtype=np.int32)
        kind = [_KIND[self.mthbc[s]] for s in range(4)]

        def idx(a, b):
            return f + b * mi + a

        nb[:, 0] = np.where(pi > 0, idx(pi - 1, pj), kind[0])
        nb[:, 1] = np.where(pi < mi - 1, idx(pi + 1, pj), kind[1])
        nb[:, 2] = np.where(pj > 0, idx(pi, pj - 1), kind[2])
        nb[:, 3] = np.where(pj < mj - 1, idx(pi, pj + 1), kind[3])
        for c, (dxs, dys) in enumerate([(-1, -1), (1, -1), (-1, 1), (1, 1)]):
            ai, bj = pi + dxs, pj + dys
            ok = (ai >= 0) & (ai < mi) & (bj >= 0) & (bj < mj)
            nb[:, 4 + c] = np.where(ok, idx(ai, bj), GF_NB_WALL)
        return nb


def pad_bathymetry(b: np.ndarray, mthbc=(3, 3, 3, 3), mbc: int = MBC) -> np.ndarray:
    """Globally padded bathymetry with the ghost values bc2 leaves at the
    domain boundary (walls and hydrograph sides mirror, extrapolation copies
    the edge cell; sides applied left, right, bottom, top)."""
    ny, nx = b.shape
    out = np.zeros((ny + 2 * mbc, nx + 2 * mbc))
    out[mbc:mbc + ny, mbc:mbc + nx] = b
    for side in range(4):
        wall = mthbc[side] in (0, 3)
        for k in range(1, mbc + 1):
            if side == 0:
                out[:, mbc - k] = out[:, mbc + k - 1] if wall else out[:, mbc]
            elif side == 1:
                out[:, mbc + nx - 1 + k] = out[:, mbc + nx - k] if wall else out[:, mbc + nx - 1]
            elif side == 2:
                out[mbc - k, :] = out[mbc + k - 1, :] if wall else out[mbc, :]
            else:
                out[mbc + ny - 1 + k, :] = out[mbc + ny - k, :] if wall else out[mbc + ny - 1, :]
    return out


def to_patches(padded: np.ndarray, mi: int, mj: int, mx: int, my: int) -> np.ndarray:
    """Split a globally padded field (ny+4, nx+4) into overlapping patch
    tiles (mj*mi, my+4, mx+4)."""
    w = np.lib.stride_tricks.sliding_window_view(padded, (my + 2 * MBC, mx + 2 * MBC))
    return np.ascontiguousarray(w[::my, ::mx][:mj, :mi]).reshape(mi * mj, my + 2 * MBC, mx + 2 * MBC)


def global_to_aos(h, hu, hv, b, layout: BrickLayout):
    """Host AoS q (npatch, my+4, mx+4, 3) and aux (npatch, my+4, mx+4, 1) in
    the reference's layout.  q ghost cells are left zero (the ghost fill sets
    them); aux ghosts hold the padded bathymetry."""
    mi, mj, mx, my = layout.mi, layout.mj, layout.mx, layout.my
    assert h.shape == (mj * my, mi * mx)
    q = np.zeros((mi * mj, my + 2 * MBC, mx + 2 * MBC, 3))
    for m, f in enumerate((h, hu, hv)):
        q[:, MBC:MBC + my, MBC:MBC + mx, m] = f.reshape(mj, my, mi, mx).transpose(0, 2, 1, 3).reshape(
            mi * mj, my, mx)
    aux = to_patches(pad_bathymetry(b, layout.mthbc), mi, mj, mx, my)[..., None]
    return q, np.ascontiguousarray(aux)


def aos_to_global(q: np.ndarray, layout: BrickLayout):
    mi, mj, mx, my = layout.mi, layout.mj, layout.mx, layout.my
    inner = q[:, MBC:MBC + my, MBC:MBC + mx, :]
    out = inner.reshape(mj, mi, my, mx, 3).transpose(4, 0, 2, 1, 3).reshape(3, mj * my, mi * mx)
    return out[0].copy(), out[1].copy(), out[2].copy()


class Pool:
    """A device-resident patch pool (gf_pool_* of the C ABI)."""

    def __init__(self, mx: int, my: int, capacity: int, mbc: int = MBC):
        self.lib = load()
        self.mx, self.my, self.mbc, self.capacity = mx, my, mbc, capacity
        h = self.lib.gf_pool_create(mx, my, mbc, capacity)
        if not h:
            raise MemoryError(self.lib.gf_last_error().decode())
        self.h = ct.c_void_p(h)
        self.layouts: list[BrickLayout] = []
        self.n = 0

    def close(self):
        if self.h:
            self.lib.gf_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def bytes_for(mx, my, capacity, mbc=MBC) -> int:
        return int(load().gf_pool_bytes(mx, my, mbc, capacity))

    # -- layout ---------------------------------------------------------------
    def set_layout(self, layouts: list[BrickLayout]):
        """Place bricks one after another in the pool and describe them."""
        first = 0
        dx, dy, xl, yl, nb = [], [], [], [], []
        for L in layouts:
            assert (L.mx, L.my) == (self.mx, self.my)
            L.first = first
            dx.append(np.full(L.npatch, L.dx))
            dy.append(np.full(L.npatch, L.dy))
            ox, oy = L.origins()
            xl.append(ox)
            yl.append(oy)
            nb.append(L.neighbours())
            first += L.npatch
        if first > self.capacity:
            raise ValueError(f"{first} patches exceed the pool capacity {self.capacity}")
        self.layouts = list(layouts)
        self.n = first
        dxa = np.ascontiguousarray(np.concatenate(dx)) if dx else np.zeros(0)
        dya = np.ascontiguousarray(np.concatenate(dy)) if dy else np.zeros(0)
        xla = np.ascontiguousarray(np.concatenate(xl)) if xl else np.zeros(0)
        yla = np.ascontiguousarray(np.concatenate(yl)) if yl else np.zeros(0)
        nba = np.ascontiguousarray(np.concatenate(nb)) if nb else np.zeros((0, 8), np.int32)
        sc = np.full(first, -1, dtype=np.int32)
        check(self.lib.gf_pool_set_patches(self.h, first, dptr(dxa), dptr(dya), dptr(xla), dptr(yla),
                                           iptr(sc), iptr(nba)), "gf_pool_set_patches")

    def set_patches(self, dx, dy, xlo, ylo, nbr=None):
        """Describe slots [0, n) one by one (gf_pool_set_patches); nbr (n, 8)
        neighbour codes, None: every ghost cell is supplied by the host."""
        d = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1)) for a in (dx, dy, xlo, ylo)]
        n = len(d[0])
        nb = None if nbr is None else np.ascontiguousarray(np.asarray(nbr, dtype=np.int32).reshape(n, 8))
        check(self.lib.gf_pool_set_patches(self.h, n, *[dptr(a) for a in d], None,
                                           None if nb is None else iptr(nb)), "gf_pool_set_patches")
        self.layouts = []
        self.n = n

    def describe(self, layout: BrickLayout, first: int):
        """Describe one brick at pool slots [first, first+npatch) without
        touching the other slots (gf_pool_describe)."""
        assert (layout.mx, layout.my) == (self.mx, self.my)
        layout.first = first
        n = layout.npatch
        ox, oy = layout.origins()
        nb = np.ascontiguousarray(layout.neighbours())
        check(self.lib.gf_pool_describe(self.h, first, n, dptr(np.full(n, float(layout.dx))),
                                        dptr(np.full(n, float(layout.dy))), dptr(ox), dptr(oy),
                                        iptr(nb)), "gf_pool_describe")
        self.n = max(self.n, first + n)

    # -- data movement --------------------------------------------------------
    def upload_aos(self, first: int, q: np.ndarray, aux: np.ndarray | None, buf: int = 0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        n = q.shape[0]
        a = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
        # aux fields per cell: the last axis of a (n, ny, nx, maux) array (b is field 0)
        maux = 1 if a is None or a.ndim < 4 else int(a.shape[-1])
        check(self.lib.gf_pool_upload(self.h, first, n, buf, dptr(q), None if a is None else dptr(a), maux),
              "gf_pool_upload")

    def download_aos(self, first: int, n: int, buf: int = 0) -> np.ndarray:
        q = np.zeros((n, self.my + 2 * MBC, self.mx + 2 * MBC, 3))
        check(self.lib.gf_pool_download(self.h, first, n, buf, dptr(q)), "gf_pool_download")
        return q

    def download_aux(self, first: int, n: int) -> np.ndarray:
        a = np.zeros((n, self.my + 2 * MBC, self.mx + 2 * MBC, 1))
        check(self.lib.gf_pool_download_aux(self.h, first, n, dptr(a)), "gf_pool_download_aux")
        return a

    def upload_global(self, layout: BrickLayout, h, hu, hv, b, buf: int = 0):
        q, aux = global_to_aos(h, hu, hv, b, layout)
        self.upload_aos(layout.first, q, aux, buf)

    def download_global(self, layout: BrickLayout, buf: int = 0):
        return aos_to_global(self.download_aos(layout.first, layout.npatch, buf), layout)

    # -- stepping ---------------------------------------------------------------
    def ghost_fill(self, first: int = 0, n: int | None = None, buf: int = 0, t: float = 0.0):
        n = self.n - first if n is None else n
        check(self.lib.gf_pool_ghost_fill(self.h, first, n, buf, t), "gf_pool_ghost_fill")

    def step(self, dt: float, first: int = 0, n: int | None = None, src: int = 0, dst: int = 1,
             plain: bool = False):
        n = self.n - first if n is None else n
        cfl = np.zeros(n)
        m = ct.c_double(0)
        fn = self.lib.gf_pool_step_plain if plain else self.lib.gf_pool_step
        check(fn(self.h, first, n, src, dst, dt, dptr(cfl), ct.byref(m)), "gf_pool_step")
        return m.value, cfl

    def add_scenario(self, layout: BrickLayout, dt0: float, desired_cfl: float = 0.9,
                     max_cfl: float = 1.0, target_steps: int = 0, t_end: float = 0.0,
                     t0: float = 0.0) -> int:
        sid = self.lib.gf_pool_add_scenario(self.h, layout.first, layout.npatch, t0, dt0,
                                            desired_cfl, max_cfl, target_steps, t_end)
        if sid < 0:
            check(sid, "gf_pool_add_scenario")
        return sid

    def remove_scenario(self, sid: int):
        check(self.lib.gf_pool_remove_scenario(self.h, sid), "gf_pool_remove_scenario")

    def set_friction(self, sid: int, manning_n: float):
        check(self.lib.gf_pool_set_scenario_friction(self.h, sid, manning_n),
              "gf_pool_set_scenario_friction")

    def set_stop(self, sid: int, target_steps: int = 0, t_end: float = 0.0):
        """New stop condition; the scenario runs again when it has not reached
        it (gf_pool_set_scenario_stop)."""
        check(self.lib.gf_pool_set_scenario_stop(self.h, sid, target_steps, t_end), "gf_pool_set_scenario_stop")

    def set_dt_max(self, sid: int, dt_max: float):
        """Upper bound of the scenario's dt (0: none); with a zero CFL number
        the next dt is dt_max (gf_pool_set_scenario_dt_max)."""
        check(self.lib.gf_pool_set_scenario_dt_max(self.h, sid, dt_max), "gf_pool_set_scenario_dt_max")

    def set_inflow(self, sid: int, time, hu, b: float, x0: float, y0: float):
        """Hydrograph of the scenario's inflow sides (gf_pool_set_scenario_inflow)."""
        t = np.ascontiguousarray(time, dtype=np.float64)
        q = np.ascontiguousarray(hu, dtype=np.float64)
        assert t.shape == q.shape and t.ndim == 1
        check(self.lib.gf_pool_set_scenario_inflow(self.h, sid, len(t), dptr(t), dptr(q), b, x0, y0),
              "gf_pool_set_scenario_inflow")

    def clear_scenarios(self):
        check(self.lib.gf_pool_clear_scenarios(self.h), "gf_pool_clear_scenarios")

    def set_active(self, sid: int, active: bool):
        check(self.lib.gf_pool_set_scenario_active(self.h, sid, int(active)), "gf_pool_set_scenario_active")

    def scenario(self, sid: int) -> dict:
        info = GfScenarioInfo()
        check(self.lib.gf_pool_get_scenario(self.h, sid, ct.byref(info)), "gf_pool_get_scenario")
        return {k: getattr(info, k) for k, _ in GfScenarioInfo._fields_}

    def scenarios(self, first: int, n: int) -> np.ndarray:
        """Control records of scenario ids [first, first+n) as a structured
        array with the fields of gf_scenario_info (gf_pool_get_scenarios)."""
        arr = (GfScenarioInfo * max(n, 1))()
        check(self.lib.gf_pool_get_scenarios(self.h, first, n, arr), "gf_pool_get_scenarios")
        dt = np.dtype([(k, np.float64 if t is ct.c_double else np.int32) for k, t in GfScenarioInfo._fields_],
                      align=True)
        return np.frombuffer(bytes(arr), dtype=dt, count=n).copy()

    def advance(self, sid: int, target_steps: int, graph: bool = True, max_attempts: int = 0) -> dict:
        """Run until scenario sid has target_steps accepted steps (or stops);
        its control record afterwards (gf_pool_advance)."""
        info = GfScenarioInfo()
        check(self.lib.gf_pool_advance(self.h, sid, target_steps, int(graph), max_attempts, ct.byref(info)),
              "gf_pool_advance")
        return {k: getattr(info, k) for k, _ in GfScenarioInfo._fields_}

    def run(self, substeps: int, graph: bool = True) -> bool:
        act = ct.c_int(0)
        check(self.lib.gf_pool_run(self.h, substeps, int(graph), ct.byref(act)), "gf_pool_run")
        return bool(act.value)

    # -- refinement tagging ------------------------------------------------------
    @staticmethod
    def refine_criteria(wave_tolerance=0.1, deep_depth=100.0, max_level_deep=3,
                        max_velocity_depth_product=1e10, speed_tolerance=(1e10,), flowgrades=(),
                        sea_level=0.0, regions=()):
        """gf_refine from Python values; flowgrades rows (value, variable,
        minlevel); regions rows (xlow, ylow, xhi, yhi, tlow, thi, minlevel,
        maxlevel).  Returns (struct, arrays kept alive)."""
        sp = np.ascontiguousarray(speed_tolerance, dtype=np.float64)
        fg = np.asarray(flowgrades, dtype=np.float64).reshape(-1, 3)
        fv = np.ascontiguousarray(fg[:, 0]) if len(fg) else np.zeros(1)
        fvar = np.ascontiguousarray(fg[:, 1], dtype=np.int32) if len(fg) else np.zeros(1, np.int32)
        fmin = np.ascontiguousarray(fg[:, 2], dtype=np.int32) if len(fg) else np.zeros(1, np.int32)
        rg = np.ascontiguousarray(np.asarray(regions, dtype=np.float64).reshape(-1, 8)) if len(regions) \
            else np.zeros((1, 8))
        c = GfRefine()
        c.wave_tolerance, c.deep_depth, c.max_level_deep = wave_tolerance, deep_depth, max_level_deep
        c.max_velocity_depth_product = max_velocity_depth_product
        c.num_speeds, c.speed_tolerance = len(sp), dptr(sp)
        c.num_flowgrades, c.flowgradevalue = len(fg), dptr(fv)
        c.iflowgradevariable, c.iflowgrademinlevel = iptr(fvar), iptr(fmin)
        c.sea_level, c.mregions, c.region = sea_level, len(regions), dptr(rg)
        return c, (sp, fv, fvar, fmin, rg)

    def tag_refine(self, patches, t, level, maxlevel, init_flag=0, buf=0, **criteria):
        pt = np.ascontiguousarray(patches, dtype=np.int32)
        out = np.zeros(len(pt), dtype=np.int32)
        c, keep = self.refine_criteria(**criteria)
        check(self.lib.gf_pool_tag_refine(self.h, buf, t, ct.byref(c), level, maxlevel, init_flag, iptr(pt),
                                          len(pt), iptr(out)), "gf_pool_tag_refine")
        return out

    def tag_coarsen(self, families, t, level, maxlevel, init_flag=0, buf=0, **criteria):
        fm = np.ascontiguousarray(np.asarray(families, dtype=np.int32).reshape(-1, 4))
        out = np.zeros(len(fm), dtype=np.int32)
        c, keep = self.refine_criteria(**criteria)
        check(self.lib.gf_pool_tag_coarsen(self.h, buf, t, ct.byref(c), level, maxlevel, init_flag, iptr(fm),
                                           len(fm), iptr(out)), "gf_pool_tag_coarsen")
        return out

    # -- output: ASCII frames and gauges -----------------------------------------
    def write_frame(self, directory: str, frame: int, time: float, first: int = 0, n: int | None = None,
                    buf: int = 0, ngrids: int | None = None, header: bool = True, patch_num=None, level=None,
                    blockno=None, mpirank: int = 0):
        """Append patches [first, first+n) of buffer buf to directory/fort.qNNNN
        in the reference's format (and, with header, write fort.tNNNN and empty
        fort.qNNNN first); gf_pool_write_frame."""
        n = self.n - first if n is None else n

        def ints(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(n))
        pn, lv, bn = ints(patch_num), ints(level), ints(blockno)
        check(self.lib.gf_pool_write_frame(self.h, buf, first, n, directory.encode(), frame, time,
                                           n if ngrids is None else ngrids, int(header),
                                           iptr(pn) if pn is not None else None,
                                           iptr(lv) if lv is not None else None,
                                           iptr(bn) if bn is not None else None, mpirank),
              "gf_pool_write_frame")

    def sample_gauges(self, slots, xc, yc, buf: int = 0) -> np.ndarray:
        """(n, 4) array of h, hu, hv, b at the gauges (gf_pool_sample_gauges)."""
        sl = np.ascontiguousarray(np.asarray(slots, dtype=np.int32).reshape(-1))
        x = np.ascontiguousarray(np.asarray(xc, dtype=np.float64).reshape(-1))
        y = np.ascontiguousarray(np.asarray(yc, dtype=np.float64).reshape(-1))
        out = np.zeros((len(sl), 4))
        check(self.lib.gf_pool_sample_gauges(self.h, buf, len(sl), iptr(sl), dptr(x), dptr(y), dptr(out)),
              "gf_pool_sample_gauges")
        return out

    # -- regrid and time interpolation between levels ---------------------------
    def average2coarse(self, coarse, fine, buf: int = 0):
        """Coarse slots from their four children (gf_pool_average2coarse)."""
        c = np.ascontiguousarray(np.asarray(coarse, dtype=np.int32).reshape(-1))
        f = np.ascontiguousarray(np.asarray(fine, dtype=np.int32).reshape(-1))
        assert len(f) == 4 * len(c)
        check(self.lib.gf_pool_average2coarse(self.h, buf, len(c), iptr(c), iptr(f)), "gf_pool_average2coarse")

    def interpolate2fine(self, coarse, fine, buf: int = 0):
        """Children from their coarse patch (gf_pool_interpolate2fine)."""
        c = np.ascontiguousarray(np.asarray(coarse, dtype=np.int32).reshape(-1))
        f = np.ascontiguousarray(np.asarray(fine, dtype=np.int32).reshape(-1))
        assert len(f) == 4 * len(c)
        check(self.lib.gf_pool_interpolate2fine(self.h, buf, len(c), iptr(c), iptr(f)), "gf_pool_interpolate2fine")

    def copy(self, slots, src_buf: int, dst_buf: int):
        """State of slots from one buffer to the other (gf_pool_copy)."""
        sl = np.ascontiguousarray(np.asarray(slots, dtype=np.int32).reshape(-1))
        check(self.lib.gf_pool_copy(self.h, len(sl), iptr(sl), src_buf, dst_buf), "gf_pool_copy")

    def _cf_links(self, coarse, fine, side, igrid):
        c, f, s = (np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1)) for a in (coarse, fine, side))
        g = np.zeros_like(s) if igrid is None else np.ascontiguousarray(np.asarray(igrid, dtype=np.int32).reshape(-1))
        assert len(c) == len(f) == len(s) == len(g)
        return c, f, s, g

    def average_ghost(self, coarse, fine, side, igrid=None, buf: int = 0):
        """Coarse ghost cells from half-size neighbours (gf_pool_average_ghost):
        link k = (coarse[k], fine[k], side[k] 0..3 face / 4..7 corner, igrid[k])."""
        c, f, s, g = self._cf_links(coarse, fine, side, igrid)
        check(self.lib.gf_pool_average_ghost(self.h, buf, len(c), iptr(c), iptr(f), iptr(s), iptr(g)),
              "gf_pool_average_ghost")

    def interpolate_ghost(self, coarse, fine, side, igrid=None, buf: int = 0):
        """Fine ghost cells from the coarse patches (gf_pool_interpolate_ghost)."""
        c, f, s, g = self._cf_links(coarse, fine, side, igrid)
        check(self.lib.gf_pool_interpolate_ghost(self.h, buf, len(c), iptr(c), iptr(f), iptr(s), iptr(g)),
              "gf_pool_interpolate_ghost")

    def timeinterp(self, src, buf_last, buf_curr, alpha, dst, dst_buf):
        s = np.ascontiguousarray(np.asarray(src, dtype=np.int32).reshape(-1))
        d = np.ascontiguousarray(np.asarray(dst, dtype=np.int32).reshape(-1))
        check(self.lib.gf_pool_timeinterp(self.h, len(s), iptr(s), buf_last, buf_curr, alpha, iptr(d), dst_buf),
              "gf_pool_timeinterp")

    # -- resident AMR ---------------------------------------------------------------
    def set_levels(self, first: int, levels):
        lv = np.ascontiguousarray(levels, dtype=np.int32)
        check(self.lib.gf_pool_set_levels(self.h, first, len(lv), iptr(lv)), "gf_pool_set_levels")

    def set_cf_links(self, coarse, fine, side, igrid):
        """Store the forest's coarse/fine links: every attempt of run() and
        ghost_update() transfers over them (gf_pool_set_cf_links)."""
        c, f, sd, ig = self._cf_links(coarse, fine, side, igrid)
        n = len(c)
        check(self.lib.gf_pool_set_cf_links(self.h, n, iptr(c), iptr(f), iptr(sd), iptr(ig)),
              "gf_pool_set_cf_links")

    def ghost_update(self, buf: int = -1, t: float = 0.0):
        check(self.lib.gf_pool_ghost_update(self.h, buf, t), "gf_pool_ghost_update")

    def move(self, src, dst, buf: int):
        s_ = np.ascontiguousarray(src, dtype=np.int32)
        d_ = np.ascontiguousarray(dst, dtype=np.int32)
        assert s_.shape == d_.shape
        check(self.lib.gf_pool_move(self.h, len(s_), iptr(s_), iptr(d_), buf), "gf_pool_move")

    def set_scenario_slots(self, sid: int, first: int, n: int):
        check(self.lib.gf_pool_set_scenario_slots(self.h, sid, first, n), "gf_pool_set_scenario_slots")
        self.n = max(self.n, first + n)

    def tag_eta_gradient(self, patches, eta_tol: float, buf: int = 0) -> np.ndarray:
        pt = np.ascontiguousarray(patches, dtype=np.int32)
        out = np.zeros(len(pt), dtype=np.int32)
        check(self.lib.gf_pool_tag_eta_gradient(self.h, buf, eta_tol, iptr(pt), len(pt), iptr(out)),
              "gf_pool_tag_eta_gradient")
        return out

    # -- step variants ------------------------------------------------------------
    def variants(self):
        """(current variant, [(tile_x, tile_y, threads, patches_per_workgroup,
        freg), ...]); freg 1: x-face increments in registers, 0: in LDS."""
        cur, n = ct.c_int(0), ct.c_int(0)
        check(self.lib.gf_pool_variants(self.h, ct.byref(cur), ct.byref(n), None), "gf_pool_variants")
        d = np.zeros(5 * n.value, dtype=np.int32)
        check(self.lib.gf_pool_variants(self.h, None, None, iptr(d)), "gf_pool_variants")
        return cur.value, [tuple(int(x) for x in d[5 * k:5 * k + 5]) for k in range(n.value)]

    def set_variant(self, v: int):
        check(self.lib.gf_pool_set_variant(self.h, v), "gf_pool_set_variant")

    def autotune(self, reps: int = 3):
        """Time each variant on the current active patches and select the
        fastest; returns (best, [ms per variant])."""
        _, vs = self.variants()
        ms = np.zeros(len(vs))
        best = ct.c_int(0)
        check(self.lib.gf_pool_autotune(self.h, reps, ct.byref(best), dptr(ms)), "gf_pool_autotune")
        return best.value, ms.tolist()

    def stats(self) -> dict:
        a, d, n, s = (ct.c_longlong(0) for _ in range(4))
        check(self.lib.gf_pool_stats(self.h, ct.byref(a), ct.byref(d), ct.byref(n), ct.byref(s)), "gf_pool_stats")
        return dict(active_patches=a.value, dry_patches=d.value, attempts=n.value, steady_patches=s.value)

    def run_plain(self, substeps: int) -> bool:
        act = ct.c_int(0)
        check(self.lib.gf_pool_run_plain(self.h, substeps, ct.byref(act)), "gf_pool_run_plain")
        return bool(act.value)


def gauge_file_create(directory: str, gid: int, xc: float, yc: float):
    """Write the header of directory/gaugeNNNNN.txt (gf_gauge_file_create)."""
    check(load().gf_gauge_file_create(directory.encode(), gid, xc, yc), "gf_gauge_file_create")


def gauge_file_append(directory: str, gid: int, level, t, rec):
    """Append records (level, t, h hu hv b rows of rec) to directory/gaugeNNNNN.txt."""
    lv = np.ascontiguousarray(np.asarray(level, dtype=np.int32).reshape(-1))
    tt = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1))
    r = np.ascontiguousarray(np.asarray(rec, dtype=np.float64).reshape(-1, 4))
    check(load().gf_gauge_file_append(directory.encode(), gid, len(lv), iptr(lv), dptr(tt), dptr(r)),
          "gf_gauge_file_append")


class HostBatch:
    """Host patches in the reference layout (q (my+4, mx+4, 3), aux
    (my+4, mx+4, maux) per patch) bound once for repeated gf_step2_batch
    calls: the pointer arrays the C ABI takes are built here, as ForestClaw's
    plugin holds its patch pointers, not on every step.  The arrays are kept
    referenced and are updated in place by step()."""

    def __init__(self, q_list, aux_list, dx, dy):
        n = len(q_list)
        if n == 0 or len(aux_list) != n:
            raise ValueError("need the same, nonzero number of q and aux patches")
        shp, ashp = q_list[0].shape, aux_list[0].shape
        for q, a in zip(q_list, aux_list):
            if q.shape != shp or a.shape != ashp:
                raise ValueError("all patches must have the same shape")
            if q.dtype != np.float64 or a.dtype != np.float64 \
                    or not q.flags.c_contiguous or not a.flags.c_contiguous:
                raise ValueError("expected C-contiguous float64 patches")
        self.q_list, self.aux_list = list(q_list), list(aux_list)
        self.n = n
        self.my, self.mx = shp[0] - 2 * MBC, shp[1] - 2 * MBC
        self.maux = ashp[-1]
        self._qp = np.fromiter((q.ctypes.data for q in self.q_list), dtype=np.uintp, count=n)
        self._ap = np.fromiter((a.ctypes.data for a in self.aux_list), dtype=np.uintp, count=n)
        self._dx = np.ascontiguousarray(np.broadcast_to(np.asarray(dx, dtype=np.float64), (n,)))
        self._dy = np.ascontiguousarray(np.broadcast_to(np.asarray(dy, dtype=np.float64), (n,)))

    def step(self, t, dt):
        """One gf_step2_batch call over every patch; returns maxcfl."""
        pp = ct.POINTER(ct.POINTER(ct.c_double))
        m = ct.c_double(0)
        check(load().gf_step2_batch(self.mx, self.my, MBC, self.maux, self.n,
                                    self._qp.ctypes.data_as(pp), self._ap.ctypes.data_as(pp),
                                    dptr(self._dx), dptr(self._dy), t, dt, ct.byref(m)),
              "gf_step2_batch")
        return m.value


def step2_batch(q_list, aux_list, dx, dy, t, dt):
    """Drop-in batch step over host patches in the reference layout
    (gf_step2_batch).  Arrays are updated in place; returns maxcfl."""
    return HostBatch(q_list, aux_list, dx, dy).step(t, dt)
