"""The C-ABI library: it loads, exports every entry point declared in
include/*.h, the Python prototypes cover exactly those, and argument/state
errors come back as status codes with a message (no GPU needed)."""
import ctypes as ct
import glob
import os
import re

import numpy as np
import pytest

import geoflood_amd as gf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(gf_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = gf.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(gf.PROTOTYPES), names ^ set(gf.PROTOTYPES)


def test_missing_library_fails_loudly(tmp_path):
    saved = gf._LIB
    gf._LIB = None
    try:
        with pytest.raises(gf.GeoFloodError):
            gf.load(str(tmp_path / "libgeoflood_amd.so"))
    finally:
        gf._LIB = saved


def test_errors_are_status_codes():
    lib = gf.load()
    assert lib.gf_check_parameters(3) == 0
    assert lib.gf_check_parameters(4) == -4
    assert b"mwaves" in lib.gf_last_error()
    # no gf_init in this process: state errors, not crashes
    p = gf.make_params()
    assert lib.gf_set_params(ct.byref(p)) == -5
    assert lib.gf_pool_create(32, 32, 2, 4) is None
    assert b"gf_init" in lib.gf_last_error()
    assert lib.gf_pool_destroy(None) == 0
    assert lib.gf_pool_set_patches(None, 0, None, None, None, None, None, None) == -1


def test_pool_bytes_accounting():
    lib = gf.load()
    b1 = lib.gf_pool_bytes(32, 32, 2, 1000)
    b2 = lib.gf_pool_bytes(32, 32, 2, 2000)
    per = b2 - b1
    # two q buffers (3 fields) + aux, 36x36 FP64, plus descriptors and control
    assert 7 * 36 * 36 * 8 * 1000 <= per <= 7 * 36 * 36 * 8 * 1000 + 300 * 1000


def test_host_batch_binding_checks():
    """HostBatch binds patch pointers once; mismatched or non-contiguous
    patches are rejected before any library call."""
    q = [np.zeros((12, 12, 3)) for _ in range(3)]
    a = [np.zeros((12, 12, 1)) for _ in range(3)]
    hb = gf.HostBatch(q, a, 0.1, 0.2)
    assert (hb.n, hb.mx, hb.my, hb.maux) == (3, 8, 8, 1)
    assert [int(x) for x in hb._qp] == [x.ctypes.data for x in q]
    with pytest.raises(ValueError):
        gf.HostBatch(q, a[:2], 0.1, 0.1)
    with pytest.raises(ValueError):
        gf.HostBatch(q[:2] + [np.zeros((12, 10, 3))], a, 0.1, 0.1)
    with pytest.raises(ValueError):
        gf.HostBatch([x[:, ::-1] for x in q], a, 0.1, 0.1)
    with pytest.raises(ValueError):
        gf.HostBatch([], [], 0.1, 0.1)


def test_cf_ghost_entry_points_reject_without_a_pool():
    """The coarse/fine ghost transfers and the regrid calls return status
    codes for a NULL pool instead of touching the device."""
    lib = gf.load()
    z = (ct.c_int * 1)(0)
    assert lib.gf_pool_average_ghost(None, 0, 1, z, z, z, z) == -1  # GF_EINVAL
    assert lib.gf_pool_interpolate_ghost(None, 0, 1, z, z, z, z) == -1  # GF_EINVAL
    assert lib.gf_pool_average2coarse(None, 0, 1, z, z) == -1  # GF_EINVAL
    assert lib.gf_pool_interpolate2fine(None, 0, 1, z, z) == -1  # GF_EINVAL
    assert b"NULL" in lib.gf_last_error()


def test_header_compiles_as_c(tmp_path):
    """The reference plugin that includes include/geoflood_amd.h is C: the
    header must compile as C99, every declaration usable from C."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = tmp_path / "use.c"
    src.write_text('#include "geoflood_amd.h"\n'
                   "int main(void) {\n"
                   "    gf_pool *pool = gf_pool_create(32, 32, 2, 4);\n"
                   "    int c[1] = {0}, f[1] = {1}, s[1] = {1}, g[1] = {0};\n"
                   "    gf_pool_average_ghost(pool, 0, 1, c, f, s, g);\n"
                   "    gf_pool_copy(pool, 1, c, 1, 0);\n"
                   "    return gf_pool_destroy(pool);\n"
                   "}\n")
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I", inc, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
