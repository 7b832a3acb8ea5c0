"""geoflood_amd -- MI355X-native GeoFlood patch step.

Python binding over the C ABI of ``libgeoflood_amd.so`` (include/geoflood_amd.h).
The library is the only compute path: if it is missing this package raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgeoflood_amd.so")

GF_OK = 0
GF_NB_WALL, GF_NB_EXTRAP, GF_NB_STORED, GF_NB_INFLOW = -1, -2, -3, -4

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)


class GfParams(ct.Structure):
    _fields_ = [("grav", ct.c_double), ("dry_tolerance", ct.c_double),
                ("manning_n", ct.c_double), ("friction_depth", ct.c_double),
                ("friction_forcing", ct.c_int), ("src_term", ct.c_int),
                ("order", ct.c_int), ("transverse", ct.c_int),
                ("mthlim", ct.c_int * 3), ("mcapa", ct.c_int)]


class GfScenarioInfo(ct.Structure):
    _fields_ = [("t", ct.c_double), ("dt", ct.c_double), ("maxcfl", ct.c_double),
                ("last_dt", ct.c_double), ("nsteps", ct.c_int), ("retakes", ct.c_int),
                ("attempts", ct.c_int), ("active", ct.c_int), ("cur", ct.c_int), ("status", ct.c_int)]


class GfRefine(ct.Structure):
    _fields_ = [("wave_tolerance", ct.c_double), ("deep_depth", ct.c_double),
                ("max_level_deep", ct.c_int), ("max_velocity_depth_product", ct.c_double),
                ("num_speeds", ct.c_int), ("speed_tolerance", ct.POINTER(ct.c_double)),
                ("num_flowgrades", ct.c_int), ("flowgradevalue", ct.POINTER(ct.c_double)),
                ("iflowgradevariable", ct.POINTER(ct.c_int)), ("iflowgrademinlevel", ct.POINTER(ct.c_int)),
                ("sea_level", ct.c_double), ("mregions", ct.c_int), ("region", ct.POINTER(ct.c_double))]


class GeoFloodError(RuntimeError):
    pass


# every function declared in include/geoflood_amd.h: name -> (restype, argtypes)
PROTOTYPES = {
    "gf_init": (ct.c_int, [ct.c_int]),
    "gf_finalize": (ct.c_int, []),
    "gf_check_parameters": (ct.c_int, [ct.c_int]),
    "gf_set_params": (ct.c_int, [ct.POINTER(GfParams)]),
    "gf_get_params": (ct.c_int, [ct.POINTER(GfParams)]),
    "gf_last_error": (ct.c_char_p, []),
    "gf_pool_bytes": (ct.c_longlong, [ct.c_int, ct.c_int, ct.c_int, ct.c_int]),
    "gf_step2_batch": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                  ct.POINTER(_dp), ct.POINTER(_dp), _dp, _dp,
                                  ct.c_double, ct.c_double, _dp]),
    "gf_step2_batch_noalloc": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                  ct.POINTER(_dp), ct.POINTER(_dp), _dp, _dp,
                                  ct.c_double, ct.c_double, _dp]),
    "gf_pool_create": (ct.c_void_p, [ct.c_int, ct.c_int, ct.c_int, ct.c_int]),
    "gf_pool_destroy": (ct.c_int, [ct.c_void_p]),
    "gf_pool_set_patches": (ct.c_int, [ct.c_void_p, ct.c_int, _dp, _dp, _dp, _dp, _ip, _ip]),
    "gf_pool_upload": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _dp, _dp, ct.c_int]),
    "gf_pool_download": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _dp]),
    "gf_pool_describe": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _ip]),
    "gf_pool_download_aux": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _dp]),
    "gf_pool_ghost_fill": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_double]),
    "gf_pool_step": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                ct.c_double, _dp, _dp]),
    "gf_pool_step_plain": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                      ct.c_double, _dp, _dp]),
    "gf_pool_add_scenario": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double,
                                        ct.c_double, ct.c_double, ct.c_double, ct.c_int,
                                        ct.c_double]),
    "gf_pool_clear_scenarios": (ct.c_int, [ct.c_void_p]),
    "gf_pool_remove_scenario": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "gf_pool_set_scenario_inflow": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _dp, _dp, ct.c_double,
                                               ct.c_double, ct.c_double]),
    "gf_pool_set_scenario_friction": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double]),
    "gf_pool_set_scenario_active": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int]),
    "gf_pool_get_scenario": (ct.c_int, [ct.c_void_p, ct.c_int, ct.POINTER(GfScenarioInfo)]),
    "gf_pool_set_scenario_dt_max": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double]),
    "gf_pool_set_scenario_stop": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double]),
    "gf_eval_powers": (ct.c_int, [ct.c_int, ct.c_int, _dp, _dp]),
    "gf_pool_get_scenarios": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.POINTER(GfScenarioInfo)]),
    "gf_pool_run": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip]),
    "gf_pool_run_plain": (ct.c_int, [ct.c_void_p, ct.c_int, _ip]),
    "gf_pool_advance": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(GfScenarioInfo)]),
    "gf_pool_tag_refine": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double, ct.POINTER(GfRefine), ct.c_int,
                                      ct.c_int, ct.c_int, _ip, ct.c_int, _ip]),
    "gf_pool_tag_coarsen": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double, ct.POINTER(GfRefine), ct.c_int,
                                       ct.c_int, ct.c_int, _ip, ct.c_int, _ip]),
    "gf_pool_variants": (ct.c_int, [ct.c_void_p, _ip, _ip, _ip]),
    "gf_pool_set_variant": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "gf_pool_autotune": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, _dp]),
    "gf_synchronize": (ct.c_int, []),
    "gf_last_timing": (ct.c_int, [_dp, _dp, _ip]),
    "gf_pool_stats": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong),
                                 ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong)]),
    "gf_pool_write_frame": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_char_p, ct.c_int,
                                       ct.c_double, ct.c_int, ct.c_int, _ip, _ip, _ip, ct.c_int]),
    "gf_pool_sample_gauges": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _dp, _dp, _dp]),
    "gf_pool_average2coarse": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _ip]),
    "gf_pool_interpolate2fine": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _ip]),
    "gf_pool_copy": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, ct.c_int, ct.c_int]),
    "gf_pool_average_ghost": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _ip, _ip, _ip]),
    "gf_pool_interpolate_ghost": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip, _ip, _ip, _ip]),
    "gf_pool_timeinterp": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, ct.c_int, ct.c_int, ct.c_double, _ip,
                                      ct.c_int]),
    "gf_pool_set_levels": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, _ip]),
    "gf_pool_set_cf_links": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, _ip, _ip, _ip]),
    "gf_pool_ghost_update": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double]),
    "gf_pool_move": (ct.c_int, [ct.c_void_p, ct.c_int, _ip, _ip, ct.c_int]),
    "gf_pool_set_scenario_slots": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int]),
    "gf_pool_tag_eta_gradient": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_double, _ip, ct.c_int, _ip]),
    "gf_gauge_file_create": (ct.c_int, [ct.c_char_p, ct.c_int, ct.c_double, ct.c_double]),
    "gf_gauge_file_append": (ct.c_int, [ct.c_char_p, ct.c_int, ct.c_int, _ip, _dp, _dp]),
    "gf_plugin_create": (ct.c_void_p, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int]),
    "gf_plugin_destroy": (ct.c_int, [ct.c_void_p]),
    "gf_plugin_single_step_update": (ct.c_int, [ct.c_void_p, _dp, _dp, ct.c_double, ct.c_double,
                                                ct.c_int, ct.c_int, ct.c_double, ct.c_double, _dp]),
}

_LIB = None


def load(path: str = LIB_PATH):
    """Load libgeoflood_amd.so and attach the C prototypes.  Raises when the
    library has not been built: there is no other implementation."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise GeoFloodError(
            f"{path} not found: build it with `python -m geoflood_amd.build` "
            "(this package has no non-HIP fallback)")
    lib = ct.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != GF_OK:
        msg = load().gf_last_error().decode()
        raise GeoFloodError(f"{what} failed ({rc}): {msg}")
    return rc


def dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(_ip)


def make_params(grav=9.81, dry_tolerance=1e-3, manning_n=0.025, friction_depth=1.0e6,
                friction_forcing=1, src_term=1, order=2, transverse=2, mthlim=(4, 4, 4)):
    p = GfParams()
    p.grav, p.dry_tolerance, p.manning_n, p.friction_depth = grav, dry_tolerance, manning_n, friction_depth
    p.friction_forcing, p.src_term, p.order, p.transverse = friction_forcing, src_term, order, transverse
    for k in range(3):
        p.mthlim[k] = mthlim[k]
    p.mcapa = 0
    return p


def init(device: int = 0):
    check(load().gf_init(device), "gf_init")


def set_params(**kw):
    p = make_params(**kw)
    check(load().gf_set_params(ct.byref(p)), "gf_set_params")
    return p


def last_timing():
    t = ct.c_double(0)
    s = ct.c_double(0)
    n = ct.c_int(0)
    load().gf_last_timing(ct.byref(t), ct.byref(s), ct.byref(n))
    return t.value, s.value, n.value


from .pool import HostBatch, Pool, step2_batch  # noqa: E402,F401
