import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgeoflood_amd.so)")


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
