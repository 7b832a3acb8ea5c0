import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgeoflood_amd.so)")


@pytest.fixture(autouse=True)
def _oracle_pow_mode(request):
    """Device tests compare with the oracle evaluating x**(7/3) and x**(2/3)
    correctly rounded, as the device does (oracle/gf_oracle.c, gfo_pow): the
    rest of the oracle is pinned bit for bit to the reference build on the
    CPU, where it keeps the libm's pow.  On dry-bed runs the libm's rare
    misroundings (about 0.08 % of calls) alone move fronts apart: two oracle
    runs differing only in them reach 1.0e-10 at 4096^2 after 100 steps."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    from oracle.oracle import CLib
    lib = CLib()
    lib.set_pow_mode(True)
    yield
    lib.set_pow_mode(False)


@pytest.fixture(scope="session")
def oracle_c():
    from oracle.oracle import CLib, Params
    lib = CLib()
    lib.set_params(Params())
    return lib


@pytest.fixture(scope="session")
def oracle_ref():
    """The reference Fortran build, when oracle/_ref was built."""
    from oracle.oracle import Params, ref_lib
    lib = ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref/libgfref.so not built (needs the reference sources)")
    lib.set_params(Params())
    return lib


@pytest.fixture(scope="session")
def gf():
    import geoflood_amd
    geoflood_amd.load()
    geoflood_amd.init(0)
    geoflood_amd.set_params()
    return geoflood_amd
