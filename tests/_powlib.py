"""Host build of the step's correctly rounded powers (geoflood_amd/csrc/
gf_pow.h, the same source the device compiles) for the tests: eval(which, x)
with which 0 -> x^(7.0/3.0), 1 -> x^(2.0/3.0)."""
import ctypes as ct
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

_SRC = r'''
#include <cmath>
#define GF_HOST_DEVICE
#include "%s"
extern "C" void gf_host_powers(int which, int n, const double* x, double* out) {
    for (int k = 0; k < n; k++) out[k] = which == 0 ? gf::pow_7_3(x[k]) : gf::pow_2_3(x[k]);
}
'''


def _lib():
    global _LIB
    if _LIB is None:
        d = tempfile.mkdtemp(prefix="gf_pow_")
        src = os.path.join(d, "p.cpp")
        with open(src, "w") as f:
            f.write(_SRC % os.path.join(ROOT, "geoflood_amd", "csrc", "gf_pow.h"))
        so = os.path.join(d, "p.so")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-o", so, src], check=True)
        _LIB = ct.CDLL(so)
    return _LIB


def host_powers(which, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    dp = ct.POINTER(ct.c_double)
    _lib().gf_host_powers(which, len(x), x.ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


def libm_pow(x, y):
    """The host libm's pow, one call per value (numpy's array power may use
    its own vectorised routine instead)."""
    m = ct.CDLL("libm.so.6")
    m.pow.restype = ct.c_double
    m.pow.argtypes = [ct.c_double, ct.c_double]
    return np.array([m.pow(float(v), y) for v in x])
