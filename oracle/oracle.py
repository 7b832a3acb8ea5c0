"""ctypes bindings for the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Two implementations with the same entry points:
  * ``C``   -- oracle/liboracle.so, the C restatement (gf_oracle.c);
  * ``REF`` -- oracle/_ref/libgfref.so, the reference GeoClaw Fortran step
    compiled from its own sources (present only where oracle/Makefile's
    ``ref`` target was built).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time the reference algorithm.  Arrays follow the reference's host
layout (Fortran order): q has shape (my+2mbc, mx+2mbc, 3) in C order, i.e.
q[j, i, m], with Fortran index (m+1, i-mbc+1, j-mbc+1).
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
C_LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libgfref.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
UPDATE_FN = ct.CFUNCTYPE(ct.c_double, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                         _dp, _dp, ct.c_double, ct.c_double, ct.c_double, ct.c_double)


class GfoParams(ct.Structure):
    _fields_ = [("grav", ct.c_double), ("dry_tolerance", ct.c_double),
                ("manning_n", ct.c_double), ("friction_depth", ct.c_double),
                ("friction_forcing", ct.c_int), ("src_term", ct.c_int),
                ("order", ct.c_int), ("transverse", ct.c_int),
                ("mthlim", ct.c_int * 3)]


@dataclass
class Params:
    """Physics/method parameters shared by the oracle and the product."""
    grav: float = 9.81
    dry_tolerance: float = 1e-3
    manning_n: float = 0.025
    friction_depth: float = 1.0e6
    friction_forcing: int = 1
    src_term: int = 1
    order: int = 2
    transverse: int = 2
    mthlim: tuple = field(default=(4, 4, 4))

    def to_c(self) -> GfoParams:
        p = GfoParams()
        p.grav, p.dry_tolerance = self.grav, self.dry_tolerance
        p.manning_n, p.friction_depth = self.manning_n, self.friction_depth
        p.friction_forcing, p.src_term = self.friction_forcing, self.src_term
        p.order, p.transverse = self.order, self.transverse
        for k in range(3):
            p.mthlim[k] = self.mthlim[k]
        return p


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


@dataclass
class Hydrograph:
    """Discharge hydrograph of the bc2 inflow (hydrograh_module.f90): hu[k]
    is the discharge per unit width at time[k]; the channel strip has width
    b about x0 (bottom/top sides) or y0 (left/right sides); sides are the
    bc2 sides (left, right, bottom, top) whose mthbc = 0 uses it."""
    time: np.ndarray
    hu: np.ndarray
    b: float
    x0: float
    y0: float
    sides: tuple = (1, 0, 0, 0)

    def __post_init__(self):
        self.time = np.ascontiguousarray(self.time, dtype=np.float64)
        self.hu = np.ascontiguousarray(self.hu, dtype=np.float64)

    def to_c(self):
        h = GfoHydrograph()
        h.nrows = len(self.time)
        h.time = self.time.ctypes.data_as(_dp)
        h.hu = self.hu.ctypes.data_as(_dp)
        h.b, h.x0, h.y0 = self.b, self.x0, self.y0
        for k in range(4):
            h.side[k] = int(self.sides[k])
        return h


class GfoHydrograph(ct.Structure):
    _fields_ = [("nrows", ct.c_int), ("time", _dp), ("hu", _dp), ("b", ct.c_double),
                ("x0", ct.c_double), ("y0", ct.c_double), ("side", ct.c_int * 4)]


@dataclass
class Refine:
    """Refinement criteria (refinement_data_module, geoclaw sea_level and the
    regions common block).  regions: rows of (xlow, ylow, xhi, yhi, tlow,
    thi, minlevel, maxlevel); flowgrades: rows of (value, variable 1-3,
    minlevel)."""
    wave_tolerance: float = 0.1
    deep_depth: float = 100.0
    max_level_deep: int = 3
    max_velocity_depth_product: float = 1.0e10
    speed_tolerance: tuple = (1.0e10,)
    flowgrades: tuple = ()
    sea_level: float = 0.0
    regions: tuple = ()

    def arrays(self):
        self._sp = np.ascontiguousarray(self.speed_tolerance, dtype=np.float64)
        fg = np.asarray(self.flowgrades, dtype=np.float64).reshape(-1, 3)
        self._fgv = np.ascontiguousarray(fg[:, 0]) if len(fg) else np.zeros(1)
        self._fgvar = np.ascontiguousarray(fg[:, 1], dtype=np.int32) if len(fg) else np.zeros(1, np.int32)
        self._fgmin = np.ascontiguousarray(fg[:, 2], dtype=np.int32) if len(fg) else np.zeros(1, np.int32)
        self._reg = np.ascontiguousarray(np.asarray(self.regions, dtype=np.float64).reshape(-1, 8)) \
            if len(self.regions) else np.zeros((1, 8))
        return len(fg)

    def to_c(self):
        nfg = self.arrays()
        r = GfoRefine()
        r.wave_tolerance, r.deep_depth = self.wave_tolerance, self.deep_depth
        r.max_level_deep, r.max_velocity_depth_product = self.max_level_deep, self.max_velocity_depth_product
        r.num_speeds, r.speed_tolerance = len(self._sp), self._sp.ctypes.data_as(_dp)
        r.num_flowgrades = nfg
        r.flowgradevalue = self._fgv.ctypes.data_as(_dp)
        r.iflowgradevariable = self._fgvar.ctypes.data_as(_ip)
        r.iflowgrademinlevel = self._fgmin.ctypes.data_as(_ip)
        r.sea_level = self.sea_level
        r.mregions = len(self.regions)
        r.region = self._reg.ctypes.data_as(_dp)
        return r


class GfoRefine(ct.Structure):
    _fields_ = [("wave_tolerance", ct.c_double), ("deep_depth", ct.c_double),
                ("max_level_deep", ct.c_int), ("max_velocity_depth_product", ct.c_double),
                ("num_speeds", ct.c_int), ("speed_tolerance", _dp), ("num_flowgrades", ct.c_int),
                ("flowgradevalue", _dp), ("iflowgradevariable", _ip), ("iflowgrademinlevel", _ip),
                ("sea_level", ct.c_double), ("mregions", ct.c_int), ("region", _dp)]


class _Lib:
    """Common wrappers over either implementation."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ct.CDLL(path)
        self.prefix = prefix
        self.params = Params()

    def f(self, name):
        return getattr(self.lib, self.prefix + name)

    # --- point functions -------------------------------------------------
    def riemanntype(self, hL, hR, uL, uR, drytol, g, maxiter=1):
        out = np.zeros(3)
        rare = np.zeros(2, dtype=np.int32)
        fn = self.f("riemanntype")
        fn.argtypes = [ct.c_double] * 4 + [ct.c_int, ct.c_double, ct.c_double, _dp, _ip]
        fn(hL, hR, uL, uR, maxiter, drytol, g, ptr(out), iptr(rare))
        return out, rare

    def riemann_aug_jcp(self, vin, drytol, g, maxiter=1):
        vin = np.ascontiguousarray(vin, dtype=np.float64)
        sw = np.zeros(3)
        fw = np.zeros(9)
        fn = self.f("riemann_aug_jcp")
        fn.argtypes = [ct.c_int, _dp, ct.c_double, ct.c_double, _dp, _dp]
        fn(maxiter, ptr(vin), drytol, g, ptr(sw), ptr(fw))
        return sw, fw.reshape(3, 3)  # fw[mw, m]

    def philim(self, a, b, meth):
        fn = self.f("philim")
        fn.argtypes = [ct.c_double, ct.c_double, ct.c_int]
        fn.restype = ct.c_double
        return fn(a, b, meth)


class CLib(_Lib):
    def __init__(self, path: str = C_LIB_PATH):
        super().__init__(path, "gfo_")
        self._cp = self.params.to_c()

    def set_params(self, params: Params):
        self.params = params
        self._cp = params.to_c()

    def set_pow_mode(self, correctly_rounded: bool):
        """x**(7/3) and x**(2/3) by the libm's pow (False, the reference
        build's arithmetic) or correctly rounded (True, the device's)."""
        self.lib.gfo_set_pow_mode(1 if correctly_rounded else 0)

    def pow(self, x, y):
        fn = self.lib.gfo_pow
        fn.argtypes = [ct.c_double, ct.c_double]
        fn.restype = ct.c_double
        return np.array([fn(float(v), y) for v in np.ravel(x)])

    def rpn2(self, ixy, mx, mbc, q1d, aux2, maux=1):
        n = mx + 2 * mbc
        fw = np.zeros((n, 3, 3)); s = np.zeros((n, 3))
        amdq = np.zeros((n, 3)); apdq = np.zeros((n, 3))
        fn = self.lib.gfo_rpn2
        fn.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        fn(ct.byref(self._cp), ixy, mx, mbc, maux, ptr(q1d), ptr(aux2), ptr(fw), ptr(s), ptr(amdq), ptr(apdq))
        return fw, s, amdq, apdq

    def rpt2(self, ixy, imp, mx, mbc, q1d, aux1, aux2, aux3, asdq, maux=1):
        n = mx + 2 * mbc
        bm = np.zeros((n, 3)); bp = np.zeros((n, 3))
        fn = self.lib.gfo_rpt2
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 5 + [_dp] * 7
        fn(ct.byref(self._cp), ixy, imp, mx, mbc, maux, ptr(q1d), ptr(aux1), ptr(aux2), ptr(aux3),
           ptr(asdq), ptr(bm), ptr(bp))
        return bm, bp

    def limiter(self, mx, mbc, wave, s, mthlim):
        ml = np.ascontiguousarray(mthlim, dtype=np.int32)
        fn = self.lib.gfo_limiter
        fn.argtypes = [ct.c_int, ct.c_int, _dp, _dp, _ip]
        fn(mx, mbc, ptr(wave), ptr(s), iptr(ml))

    def flux2(self, ixy, mx, mbc, q1d, dtdx1d, aux1, aux2, aux3, maux=1):
        n = mx + 2 * mbc
        fm = np.zeros((n, 3)); fp = np.zeros((n, 3))
        gm = np.zeros((2, n, 3)); gp = np.zeros((2, n, 3))
        cfl = ct.c_double(0)
        fn = self.lib.gfo_flux2
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp] * 9 + [_dp]
        fn(ct.byref(self._cp), ixy, mx, mbc, maux, ptr(q1d), ptr(dtdx1d), ptr(aux1), ptr(aux2),
           ptr(aux3), ptr(fm), ptr(fp), ptr(gm), ptr(gp), ct.byref(cfl))
        return fm, fp, gm, gp, cfl.value

    def step2_wrap(self, mx, my, mbc, q, aux, dx, dy, dt, maux=1):
        fn = self.lib.gfo_step2_wrap
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp, _dp] + [ct.c_double] * 3
        fn.restype = ct.c_double
        return fn(ct.byref(self._cp), mx, my, mbc, maux, ptr(q), ptr(aux), dx, dy, dt)

    def b4step2(self, mx, my, mbc, q, aux=None, dt=0.0):
        fn = self.lib.gfo_b4step2
        fn.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _dp]
        fn(ct.byref(self._cp), mx, my, mbc, ptr(q))

    def src2(self, mx, my, mbc, q, aux, dt, maux=1):
        fn = self.lib.gfo_src2
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp, _dp, ct.c_double]
        fn(ct.byref(self._cp), mx, my, mbc, maux, ptr(q), ptr(aux), dt)

    def bc2(self, mx, my, mbc, q, aux, mthbc, maux=1):
        mb = np.ascontiguousarray(mthbc, dtype=np.int32)
        fn = self.lib.gfo_bc2
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp, _ip]
        fn(mx, my, mbc, maux, ptr(q), ptr(aux), iptr(mb))

    def patch_update(self, mx, my, mbc, q, aux, dx, dy, t, dt, maux=1):
        fn = self.lib.gfo_patch_update
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp, _dp] + [ct.c_double] * 4
        fn.restype = ct.c_double
        return fn(ct.byref(self._cp), mx, my, mbc, maux, ptr(q), ptr(aux), dx, dy, t, dt)

    # --- hydrograph inflow -------------------------------------------------
    def hydro_interpolate(self, hyd: Hydrograph, t):
        fn = self.lib.gfo_hydro_interpolate
        fn.argtypes = [ct.POINTER(GfoHydrograph), ct.c_double]
        fn.restype = ct.c_double
        h = hyd.to_c()
        return fn(ct.byref(h), t)

    def riemann_invariants(self, q0, q1, grav=9.81):
        a = np.ascontiguousarray(q0, dtype=np.float64).copy()
        c = np.ascontiguousarray(q1, dtype=np.float64)
        fn = self.lib.gfo_riemann_invariants
        fn.argtypes = [_dp, _dp, ct.c_double]
        fn(ptr(a), ptr(c), grav)
        return a

    def two_shock(self, q0, q1, grav=9.81):
        a = np.ascontiguousarray(q0, dtype=np.float64).copy()
        c = np.ascontiguousarray(q1, dtype=np.float64)
        fn = self.lib.gfo_two_shock
        fn.argtypes = [_dp, _dp, ct.c_double]
        fn(ptr(a), ptr(c), grav)
        return a

    def bc2_ex(self, mx, my, mbc, q, aux, mthbc, xlower, ylower, dx, dy, t, hyd=None, maux=1):
        mb = np.ascontiguousarray(mthbc, dtype=np.int32)
        fn = self.lib.gfo_bc2_ex
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp, _dp, _ip] + [ct.c_double] * 5 + \
            [ct.c_void_p]
        fn.restype = ct.c_int
        h = None if hyd is None else hyd.to_c()
        return fn(ct.byref(self._cp), mx, my, mbc, maux, ptr(q), ptr(aux), iptr(mb), xlower, ylower,
                  dx, dy, t, None if h is None else ct.cast(ct.byref(h), ct.c_void_p))

    # --- refinement tagging -----------------------------------------------
    def tag4refinement(self, rf: Refine, mx, my, mbc, q, aux, xlower, ylower, dx, dy, t, level,
                       maxlevel, init_flag=0, maux=1):
        fn = self.lib.gfo_tag4refinement
        fn.argtypes = [ct.c_void_p, ct.POINTER(GfoRefine)] + [ct.c_int] * 4 + [ct.c_double] * 5 + \
            [_dp, _dp] + [ct.c_int] * 3
        fn.restype = ct.c_int
        r = rf.to_c()
        return fn(ct.byref(self._cp), ct.byref(r), mx, my, mbc, maux, xlower, ylower, dx, dy, t,
                  ptr(q), ptr(aux), level, maxlevel, init_flag)

    def tag4coarsening(self, rf: Refine, mx, my, mbc, qs, auxs, xlower, ylower, dx, dy, t, level,
                       maxlevel, init_flag=0, maux=1):
        fn = self.lib.gfo_tag4coarsening
        fn.argtypes = [ct.c_void_p, ct.POINTER(GfoRefine)] + [ct.c_int] * 4 + [_dp, _dp] + \
            [ct.c_double] * 3 + [ct.POINTER(_dp), ct.POINTER(_dp)] + [ct.c_int] * 3
        fn.restype = ct.c_int
        r = rf.to_c()
        xl = np.ascontiguousarray(xlower, dtype=np.float64)
        yl = np.ascontiguousarray(ylower, dtype=np.float64)
        qp = (_dp * 4)(*[ptr(a) for a in qs])
        ap = (_dp * 4)(*[ptr(a) for a in auxs])
        return fn(ct.byref(self._cp), ct.byref(r), mx, my, mbc, maux, ptr(xl), ptr(yl), dx, dy, t, qp, ap,
                  level, maxlevel, init_flag)

    # --- ASCII frame output and gauges ---------------------------------------
    def format_E(self, v, w, d) -> str:
        fn = self.lib.gfo_format_E
        fn.argtypes = [ct.c_char_p, ct.c_double, ct.c_int, ct.c_int]
        fn.restype = None
        buf = ct.create_string_buffer(w + 1)
        fn(buf, v, w, d)
        return buf.raw[:w].decode()

    def frame_header(self, time, meqn, maux, ngrids, mbc) -> bytes:
        fn = self.lib.gfo_frame_header
        fn.argtypes = [ct.c_char_p, ct.c_int, ct.c_double] + [ct.c_int] * 4
        fn.restype = ct.c_int
        buf = ct.create_string_buffer(1024)
        n = fn(buf, 1024, time, meqn, maux, ngrids, mbc)
        assert n >= 0
        return buf.raw[:n]

    def frame_patch(self, q, aux, xlower, ylower, dx, dy, patch_num, level, blockno, mpirank) -> bytes:
        """One patch's block of fort.qNNNN; q (my+2mbc, mx+2mbc, meqn) AoS is
        modified in place (|q| < 1e-99 -> 0) as the reference does."""
        ny, nx, meqn = q.shape
        maux = aux.shape[-1]
        mbc = 2
        mx, my = nx - 2 * mbc, ny - 2 * mbc
        fn = self.lib.gfo_frame_patch
        fn.argtypes = [ct.c_char_p] + [ct.c_int] * 6 + [ct.c_double] * 4 + [_dp, _dp] + [ct.c_int] * 4
        fn.restype = ct.c_int
        cap = 512 + mx * my * 110 + my * 4
        buf = ct.create_string_buffer(cap)
        n = fn(buf, cap, mx, my, meqn, maux, mbc, xlower, ylower, dx, dy, ptr(q), ptr(aux), patch_num, level,
               blockno, mpirank)
        assert n >= 0
        return buf.raw[:n]

    def update_gauge(self, q, aux, xlower, ylower, dx, dy, xc, yc):
        ny, nx, _ = q.shape
        fn = self.lib.gfo_update_gauge
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [ct.c_double] * 4 + [_dp, _dp, ct.c_double, ct.c_double,
                                                                          _dp, _dp]
        fn.restype = None
        qv = np.zeros(3)
        av = np.zeros(1)
        fn(ct.byref(self._cp), nx - 4, ny - 4, 2, aux.shape[-1], xlower, ylower, dx, dy, ptr(q), ptr(aux), xc, yc,
           ptr(qv), ptr(av))
        return qv, av

    def gauge_header(self, gid, xc, yc) -> bytes:
        fn = self.lib.gfo_gauge_header
        fn.argtypes = [ct.c_char_p, ct.c_int, ct.c_int, ct.c_double, ct.c_double]
        fn.restype = ct.c_int
        buf = ct.create_string_buffer(256)
        n = fn(buf, 256, gid, xc, yc)
        assert n >= 0
        return buf.raw[:n]

    def gauge_line(self, level, t, qvar, avar) -> bytes:
        fn = self.lib.gfo_gauge_line
        fn.argtypes = [ct.c_char_p, ct.c_int, ct.c_int, ct.c_double, _dp, _dp]
        fn.restype = ct.c_int
        buf = ct.create_string_buffer(256)
        qv = np.ascontiguousarray(qvar, dtype=np.float64)
        av = np.ascontiguousarray(avar, dtype=np.float64)
        n = fn(buf, 256, level, t, ptr(qv), ptr(av))
        assert n >= 0
        return buf.raw[:n]

    # --- regrid and time interpolation ---------------------------------------
    def average2coarse(self, qcoarse, qfine, aux_coarse, aux_fine, igrid):
        """Quadrant igrid of qcoarse (modified) from the fine patch qfine."""
        ny, nx, _ = qcoarse.shape
        fn = self.lib.gfo_average2coarse
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp] * 4 + [ct.c_int]
        fn.restype = None
        fn(ct.byref(self._cp), nx - 4, ny - 4, 2, aux_coarse.shape[-1], ptr(qcoarse), ptr(qfine), ptr(aux_coarse),
           ptr(aux_fine), igrid)

    def interpolate2fine(self, qcoarse, qfine, aux_coarse, aux_fine, igrid):
        """Fine patch igrid (qfine, modified) from the coarse patch."""
        ny, nx, _ = qcoarse.shape
        fn = self.lib.gfo_interpolate2fine
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp] * 4 + [ct.c_int]
        fn.restype = None
        fn(ct.byref(self._cp), nx - 4, ny - 4, 2, aux_coarse.shape[-1], ptr(qcoarse), ptr(qfine), ptr(aux_coarse),
           ptr(aux_fine), igrid)

    def average_ghost(self, qcoarse, qfine, aux_coarse, aux_fine, side, igrid=0):
        """Ghost cells of qcoarse (modified) from the half-size neighbour qfine
        across side 0..3 (face, half igrid) or 4..7 (corner)."""
        ny, nx, _ = qcoarse.shape
        fn = self.lib.gfo_average_ghost
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp] * 4 + [ct.c_int] * 2
        fn.restype = ct.c_int
        r = fn(ct.byref(self._cp), nx - 4, ny - 4, 2, aux_coarse.shape[-1], ptr(qcoarse), ptr(qfine),
               ptr(aux_coarse), ptr(aux_fine), side, igrid)
        if r != 0:
            raise ValueError("bad side / igrid")

    def interpolate_ghost(self, qcoarse, qfine, aux_coarse, aux_fine, side, igrid=0):
        """Ghost cells of the half-size neighbour qfine (modified) across side
        0..3 (face, half igrid) or 4..7 (corner) of the coarse patch."""
        ny, nx, _ = qcoarse.shape
        fn = self.lib.gfo_interpolate_ghost
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp] * 4 + [ct.c_int] * 2
        fn.restype = ct.c_int
        r = fn(ct.byref(self._cp), nx - 4, ny - 4, 2, aux_coarse.shape[-1], ptr(qcoarse), ptr(qfine),
               ptr(aux_coarse), ptr(aux_fine), side, igrid)
        if r != 0:
            raise ValueError("bad side / igrid, or mx != my")

    def compute_slopes(self, sl, sr, mth=5):
        fn = self.lib.gfo_compute_slopes
        fn.argtypes = [ct.c_double, ct.c_double, ct.c_int]
        fn.restype = ct.c_double
        return fn(sl, sr, mth)

    def timeinterp(self, qcurr, qlast, qinterp, alpha, psize):
        ny, nx, meqn = qcurr.shape
        fn = self.lib.gfo_timeinterp
        fn.argtypes = [ct.c_int] * 5 + [_dp] * 3 + [ct.c_double]
        fn.restype = ct.c_int
        return fn(nx - 4, ny - 4, 2, meqn, psize, ptr(qcurr), ptr(qlast), ptr(qinterp), alpha)

    # --- brick driver -------------------------------------------------------
    def brick_run(self, brick: "Brick", nsteps, initial_dt, desired_cfl=0.9, max_cfl=1.0,
                  t0=0.0, update=None, nthreads=1, dt_max=0.0):
        """Run the outstyle_3 loop on a Brick in place.  ``update`` is None
        (C restatement) or a RefLib (reference Fortran patch update).
        dt_max > 0 adds the library's dt bound (gfo_brick_run_dtmax)."""
        np_ = brick.mi * brick.mj
        qp = (_dp * np_)(*[ptr(a) for a in brick.q])
        ap = (_dp * np_)(*[ptr(a) for a in brick.aux])
        mb = np.ascontiguousarray(brick.mthbc, dtype=np.int32)
        dt_out = np.zeros(nsteps); cfl_out = np.zeros(nsteps); mass_out = np.zeros(nsteps)
        fnp = None
        if update is not None:
            fnp = ct.cast(update.lib.gfref_patch_update, ct.c_void_p)
        h = None if brick.hydro is None else brick.hydro.to_c()
        fn = self.lib.gfo_brick_run_dtmax
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 6 + [ct.POINTER(_dp), ct.POINTER(_dp), _ip] + \
            [ct.c_double] * 4 + [ct.c_void_p] + [ct.c_double] * 5 + \
            [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, _dp, _dp, _dp]
        fn.restype = ct.c_int
        retakes = fn(ct.byref(self._cp), brick.mi, brick.mj, brick.mx, brick.my, brick.mbc, 1,
                     qp, ap, iptr(mb), brick.xlo, brick.ylo, brick.dx, brick.dy,
                     None if h is None else ct.cast(ct.byref(h), ct.c_void_p),
                     t0, initial_dt, desired_cfl, max_cfl, dt_max,
                     nsteps, fnp, None, nthreads, ptr(dt_out), ptr(cfl_out), ptr(mass_out))
        return dict(dt=dt_out, cfl=cfl_out, mass=mass_out, retakes=retakes)

    def brick_advance(self, brick: "Brick", t, dt, update=None, nthreads=1):
        np_ = brick.mi * brick.mj
        qp = (_dp * np_)(*[ptr(a) for a in brick.q])
        ap = (_dp * np_)(*[ptr(a) for a in brick.aux])
        mb = np.ascontiguousarray(brick.mthbc, dtype=np.int32)
        fnp = None if update is None else ct.cast(update.lib.gfref_patch_update, ct.c_void_p)
        fn = self.lib.gfo_brick_advance
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 6 + [ct.POINTER(_dp), ct.POINTER(_dp), _ip] + \
            [ct.c_double] * 4 + [ct.c_void_p, ct.c_void_p, ct.c_int]
        fn.restype = ct.c_double
        return fn(ct.byref(self._cp), brick.mi, brick.mj, brick.mx, brick.my, brick.mbc, 1,
                  qp, ap, iptr(mb), brick.dx, brick.dy, t, dt, fnp, None, nthreads)

    def brick_ghost_fill(self, brick: "Brick", t=0.0):
        np_ = brick.mi * brick.mj
        qp = (_dp * np_)(*[ptr(a) for a in brick.q])
        ap = (_dp * np_)(*[ptr(a) for a in brick.aux])
        mb = np.ascontiguousarray(brick.mthbc, dtype=np.int32)
        h = None if brick.hydro is None else brick.hydro.to_c()
        fn = self.lib.gfo_brick_ghost_fill_ex
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 6 + [ct.POINTER(_dp), ct.POINTER(_dp), _ip] + \
            [ct.c_double] * 5 + [ct.c_void_p]
        fn.restype = ct.c_int
        return fn(ct.byref(self._cp), brick.mi, brick.mj, brick.mx, brick.my, brick.mbc, 1, qp, ap,
                  iptr(mb), brick.xlo, brick.ylo, brick.dx, brick.dy, t,
                  None if h is None else ct.cast(ct.byref(h), ct.c_void_p))


class RefLib(_Lib):
    """The reference Fortran step (oracle/_ref/libgfref.so)."""

    def __init__(self, path: str = REF_LIB_PATH):
        super().__init__(path, "gfref_")

    def set_params(self, params: Params):
        self.params = params
        ml = np.ascontiguousarray(params.mthlim, dtype=np.int32)
        fn = self.lib.gfref_set_params
        fn.argtypes = [ct.c_double] * 4 + [ct.c_int] * 4 + [_ip]
        fn(params.grav, params.dry_tolerance, params.manning_n, params.friction_depth,
           params.friction_forcing, params.src_term, params.order, params.transverse, iptr(ml))

    def rpn2(self, ixy, mx, mbc, q1d, aux2, maux=1):
        n = mx + 2 * mbc
        fw = np.zeros((n, 3, 3)); s = np.zeros((n, 3))
        amdq = np.zeros((n, 3)); apdq = np.zeros((n, 3))
        fn = self.lib.gfref_rpn2
        fn.argtypes = [ct.c_int] * 4 + [_dp] * 6
        fn(ixy, mx, mbc, maux, ptr(q1d), ptr(aux2), ptr(fw), ptr(s), ptr(amdq), ptr(apdq))
        return fw, s, amdq, apdq

    def rpt2(self, ixy, imp, mx, mbc, q1d, aux1, aux2, aux3, asdq, maux=1):
        n = mx + 2 * mbc
        bm = np.zeros((n, 3)); bp = np.zeros((n, 3))
        fn = self.lib.gfref_rpt2
        fn.argtypes = [ct.c_int] * 5 + [_dp] * 7
        fn(ixy, imp, mx, mbc, maux, ptr(q1d), ptr(aux1), ptr(aux2), ptr(aux3), ptr(asdq), ptr(bm), ptr(bp))
        return bm, bp

    def limiter(self, mx, mbc, wave, s, mthlim):
        ml = np.ascontiguousarray(mthlim, dtype=np.int32)
        fn = self.lib.gfref_limiter
        fn.argtypes = [ct.c_int, ct.c_int, _dp, _dp, _ip]
        fn(mx, mbc, ptr(wave), ptr(s), iptr(ml))

    def flux2(self, ixy, mx, mbc, q1d, dtdx1d, aux1, aux2, aux3, maux=1):
        n = mx + 2 * mbc
        fm = np.zeros((n, 3)); fp = np.zeros((n, 3))
        gm = np.zeros((2, n, 3)); gp = np.zeros((2, n, 3))
        cfl = ct.c_double(0)
        fn = self.lib.gfref_flux2
        fn.argtypes = [ct.c_int] * 4 + [_dp] * 9 + [_dp]
        fn(ixy, mx, mbc, maux, ptr(q1d), ptr(dtdx1d), ptr(aux1), ptr(aux2), ptr(aux3),
           ptr(fm), ptr(fp), ptr(gm), ptr(gp), ct.byref(cfl))
        return fm, fp, gm, gp, cfl.value

    def step2_wrap(self, mx, my, mbc, q, aux, dx, dy, dt, maux=1):
        fn = self.lib.gfref_step2_wrap
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp] + [ct.c_double] * 3
        fn.restype = ct.c_double
        return fn(mx, my, mbc, maux, ptr(q), ptr(aux), dx, dy, dt)

    def b4step2(self, mx, my, mbc, q, aux, dt=0.0, maux=1):
        fn = self.lib.gfref_b4step2
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp, ct.c_double]
        fn(mx, my, mbc, maux, ptr(q), ptr(aux), dt)

    def src2(self, mx, my, mbc, q, aux, dt, maux=1):
        fn = self.lib.gfref_src2
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp, ct.c_double]
        fn(mx, my, mbc, maux, ptr(q), ptr(aux), dt)

    def bc2(self, mx, my, mbc, q, aux, mthbc, maux=1):
        mb = np.ascontiguousarray(mthbc, dtype=np.int32)
        fn = self.lib.gfref_bc2
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp, _ip]
        fn(mx, my, mbc, maux, ptr(q), ptr(aux), iptr(mb))

    def set_hydrograph(self, hyd: Hydrograph):
        fn = self.lib.gfref_set_hydrograph
        fn.argtypes = [ct.c_int, _dp, _dp, ct.c_double, ct.c_double, ct.c_double, _ip]
        sides = np.ascontiguousarray(hyd.sides, dtype=np.int32)
        fn(len(hyd.time), ptr(hyd.time), ptr(hyd.hu), hyd.b, hyd.x0, hyd.y0, iptr(sides))

    def hydro_interpolate(self, t):
        fn = self.lib.gfref_hydro_interpolate
        fn.argtypes = [ct.c_double]
        fn.restype = ct.c_double
        return fn(t)

    def riemann_invariants(self, q0, q1):
        a = np.ascontiguousarray(q0, dtype=np.float64).copy()
        c = np.ascontiguousarray(q1, dtype=np.float64)
        fn = self.lib.gfref_riemann_invariants
        fn.argtypes = [_dp, _dp]
        fn(ptr(a), ptr(c))
        return a

    def two_shock(self, q0, q1):
        a = np.ascontiguousarray(q0, dtype=np.float64).copy()
        c = np.ascontiguousarray(q1, dtype=np.float64)
        fn = self.lib.gfref_two_shock
        fn.argtypes = [_dp, _dp]
        fn(ptr(a), ptr(c))
        return a

    def bc2_ex(self, mx, my, mbc, q, aux, mthbc, xlower, ylower, dx, dy, t, maux=1):
        mb = np.ascontiguousarray(mthbc, dtype=np.int32)
        fn = self.lib.gfref_bc2_ex
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp, _ip] + [ct.c_double] * 5
        fn(mx, my, mbc, maux, ptr(q), ptr(aux), iptr(mb), xlower, ylower, dx, dy, t)

    def set_refinement(self, rf: Refine):
        nfg = rf.arrays()
        fn = self.lib.gfref_set_refinement
        fn.argtypes = [ct.c_double, ct.c_double, ct.c_int, ct.c_double, ct.c_int, _dp, ct.c_int, _dp, _ip,
                       _ip, ct.c_double, ct.c_int, _dp]
        fn(rf.wave_tolerance, rf.deep_depth, rf.max_level_deep, rf.max_velocity_depth_product,
           len(rf._sp), ptr(rf._sp), nfg, ptr(rf._fgv), iptr(rf._fgvar), iptr(rf._fgmin), rf.sea_level,
           len(rf.regions), ptr(rf._reg))

    def tag4refinement(self, mx, my, mbc, q, aux, xlower, ylower, dx, dy, t, level, maxlevel,
                       init_flag=0, maux=1):
        fn = self.lib.gfref_tag4refinement
        fn.argtypes = [ct.c_int] * 4 + [ct.c_double] * 5 + [_dp, _dp] + [ct.c_int] * 3
        fn.restype = ct.c_int
        return fn(mx, my, mbc, maux, xlower, ylower, dx, dy, t, ptr(q), ptr(aux), level, maxlevel, init_flag)

    def tag4coarsening(self, mx, my, mbc, qs, auxs, xlower, ylower, dx, dy, t, level, maxlevel,
                       init_flag=0, maux=1):
        fn = self.lib.gfref_tag4coarsening
        fn.argtypes = [ct.c_int] * 4 + [_dp, _dp] + [ct.c_double] * 3 + [_dp] * 8 + [ct.c_int] * 3
        fn.restype = ct.c_int
        xl = np.ascontiguousarray(xlower, dtype=np.float64)
        yl = np.ascontiguousarray(ylower, dtype=np.float64)
        return fn(mx, my, mbc, maux, ptr(xl), ptr(yl), dx, dy, t, *[ptr(a) for a in qs],
                  *[ptr(a) for a in auxs], level, maxlevel, init_flag)

    def write_frame(self, workdir, iframe, time, patches, meqn=3, maux=1, mbc=2, ngrids=None, mpirank=0):
        """Run the reference writers in workdir: the header, then each patch
        (q, aux, xlower, ylower, dx, dy, patch_num, level, blockno) in order.
        Returns (fort.t bytes, fort.q bytes).  The reference modifies q."""
        hdr = self.lib.gfref_write_header
        hdr.argtypes = [ct.c_int, ct.c_double] + [ct.c_int] * 4
        hdr.restype = None
        wf = self.lib.gfref_write_file
        wf.argtypes = [ct.c_int] * 5 + [ct.c_double] * 4 + [_dp, _dp] + [ct.c_int] * 5
        wf.restype = None
        cwd = os.getcwd()
        os.chdir(workdir)
        try:
            hdr(iframe, time, meqn, maux, len(patches) if ngrids is None else ngrids, mbc)
            for (q, aux, xl, yl, dx, dy, num, level, blockno) in patches:
                ny, nx, _ = q.shape
                wf(nx - 2 * mbc, ny - 2 * mbc, meqn, maux, mbc, xl, yl, dx, dy, ptr(q), ptr(aux), iframe, num,
                   level, blockno, mpirank)
            tag = "%04d" % (iframe % 10000)
            with open("fort.t" + tag, "rb") as f:
                tb = f.read()
            with open("fort.q" + tag, "rb") as f:
                qb = f.read()
        finally:
            os.chdir(cwd)
        return tb, qb

    def update_gauge(self, q, aux, xlower, ylower, dx, dy, xc, yc):
        ny, nx, _ = q.shape
        fn = self.lib.gfref_update_gauge
        fn.argtypes = [ct.c_int] * 4 + [ct.c_double] * 4 + [_dp, _dp, ct.c_double, ct.c_double, _dp, _dp]
        fn.restype = None
        qv = np.zeros(3)
        av = np.zeros(1)
        fn(nx - 4, ny - 4, 2, aux.shape[-1], xlower, ylower, dx, dy, ptr(q), ptr(aux), xc, yc, ptr(qv), ptr(av))
        return qv, av

    def patch_update(self, mx, my, mbc, q, aux, dx, dy, t, dt, maux=1):
        fn = self.lib.gfref_patch_update
        fn.argtypes = [ct.c_void_p] + [ct.c_int] * 4 + [_dp, _dp] + [ct.c_double] * 4
        fn.restype = ct.c_double
        return fn(None, mx, my, mbc, maux, ptr(q), ptr(aux), dx, dy, t, dt)


@dataclass
class Brick:
    """A uniform single-level brick of mi x mj patches in the reference's host
    layout: q[k] has shape (my+2mbc, mx+2mbc, 3), aux[k] (my+2mbc, mx+2mbc, 1),
    for patch k = pj*mi + pi."""
    mi: int
    mj: int
    mx: int
    my: int
    mbc: int
    dx: float
    dy: float
    q: list
    aux: list
    mthbc: tuple = (3, 3, 3, 3)
    xlo: float = 0.0
    ylo: float = 0.0
    hydro: Hydrograph | None = None

    @classmethod
    def from_global(cls, h, hu, hv, b, mi, mj, mbc, dx, dy, mthbc=(3, 3, 3, 3), xlo=0.0, ylo=0.0,
                    hydro=None):
        """Split global interior fields (shape (ny, nx)) into patches.  Ghost
        values of aux are taken from a globally padded b with the wall/extrap
        rule applied (bc2 then re-applies it on the boundary every step)."""
        ny, nx = h.shape
        mx, my = nx // mi, ny // mj
        bpad = pad_aux(b, mbc, mthbc)
        q, aux = [], []
        for pj in range(mj):
            for pi in range(mi):
                qq = np.zeros((my + 2 * mbc, mx + 2 * mbc, 3))
                sl = (slice(pj * my, (pj + 1) * my), slice(pi * mx, (pi + 1) * mx))
                qq[mbc:mbc + my, mbc:mbc + mx, 0] = h[sl]
                qq[mbc:mbc + my, mbc:mbc + mx, 1] = hu[sl]
                qq[mbc:mbc + my, mbc:mbc + mx, 2] = hv[sl]
                aa = np.ascontiguousarray(
                    bpad[pj * my:pj * my + my + 2 * mbc, pi * mx:pi * mx + mx + 2 * mbc, None])
                q.append(qq)
                aux.append(aa)
        return cls(mi, mj, mx, my, mbc, dx, dy, q, aux, tuple(mthbc), xlo, ylo, hydro)

    def to_global(self):
        """Interior fields (h, hu, hv) as global arrays of shape (ny, nx)."""
        mx, my, mbc = self.mx, self.my, self.mbc
        out = np.zeros((3, self.mj * my, self.mi * mx))
        for pj in range(self.mj):
            for pi in range(self.mi):
                qq = self.q[pj * self.mi + pi]
                out[:, pj * my:(pj + 1) * my, pi * mx:(pi + 1) * mx] = \
                    np.moveaxis(qq[mbc:mbc + my, mbc:mbc + mx, :], 2, 0)
        return out

    def copy(self):
        return Brick(self.mi, self.mj, self.mx, self.my, self.mbc, self.dx, self.dy,
                     [a.copy() for a in self.q], [a.copy() for a in self.aux], self.mthbc,
                     self.xlo, self.ylo, self.hydro)


def pad_aux(b, mbc, mthbc=(3, 3, 3, 3)):
    """Globally padded bathymetry: walls (3) and hydrograph sides (0, a wall
    outside the channel strip) mirror, extrapolation (1) copies the edge
    cell, applied left, right, bottom, top as bc2 does."""
    ny, nx = b.shape
    out = np.zeros((ny + 2 * mbc, nx + 2 * mbc))
    out[mbc:mbc + ny, mbc:mbc + nx] = b
    for side in range(4):
        kind = mthbc[side]
        for k in range(1, mbc + 1):
            if side == 0:
                out[:, mbc - k] = out[:, mbc + k - 1] if kind in (0, 3) else out[:, mbc]
            elif side == 1:
                out[:, mbc + nx - 1 + k] = out[:, mbc + nx - k] if kind in (0, 3) else out[:, mbc + nx - 1]
            elif side == 2:
                out[mbc - k, :] = out[mbc + k - 1, :] if kind in (0, 3) else out[mbc, :]
            else:
                out[mbc + ny - 1 + k, :] = out[mbc + ny - k, :] if kind in (0, 3) else out[mbc + ny - 1, :]
    return out


_C = None
_REF = None


def c_lib() -> CLib:
    global _C
    if _C is None:
        _C = CLib()
    return _C


class RefAmrLib:
    """The reference's regrid averaging and time interpolation
    (oracle/_ref/libgfref_amr.so), loaded with lazy binding: the same source
    files' face/corner routines call ForestClaw, which is absent, and are
    never called."""

    def __init__(self, path: str = os.path.join(os.path.dirname(REF_LIB_PATH), "libgfref_amr.so")):
        self.lib = ct.CDLL(path, mode=os.RTLD_LAZY)

    def set_dry_tolerance(self, drytol):
        fn = self.lib.gfref_amr_set_drytol
        fn.argtypes = [ct.c_double]
        fn.restype = None
        fn(drytol)

    def average2coarse(self, qcoarse, qfine, aux_coarse, aux_fine, igrid):
        ny, nx, _ = qcoarse.shape
        fn = self.lib.gfref_average2coarse
        fn.argtypes = [ct.c_int] * 4 + [_dp] * 4 + [ct.c_int]
        fn.restype = None
        fn(nx - 4, ny - 4, 2, aux_coarse.shape[-1], ptr(qcoarse), ptr(qfine), ptr(aux_coarse), ptr(aux_fine), igrid)

    def timeinterp(self, qcurr, qlast, qinterp, alpha, psize):
        ny, nx, meqn = qcurr.shape
        fn = self.lib.gfref_timeinterp
        fn.argtypes = [ct.c_int] * 5 + [_dp] * 3 + [ct.c_double]
        fn.restype = ct.c_int
        return fn(nx - 4, ny - 4, 2, meqn, psize, ptr(qcurr), ptr(qlast), ptr(qinterp), alpha)


def ref_amr_lib():
    """The reference regrid build, or None where it was not built."""
    path = os.path.join(os.path.dirname(REF_LIB_PATH), "libgfref_amr.so")
    return RefAmrLib(path) if os.path.exists(path) else None


def ref_lib():
    """The reference Fortran build, or None where it was not built."""
    global _REF
    if _REF is None and os.path.exists(REF_LIB_PATH):
        _REF = RefLib()
    return _REF
