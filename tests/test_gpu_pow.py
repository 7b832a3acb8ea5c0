"""The step's powers on the device (gf_eval_powers: the code src2 and the
hydrograph inflow run) against the same source built for the host and
against the host libm's pow, over 2^20 arguments in [1e-3, 20] (the depths
and discharges of the workloads): bit-identical to the host build
(correctly rounded, whichever cbrt and log the library provides), and
different from glibc's pow only where glibc misrounds (about 0.08 %; the
device's own pow differs from glibc in about 24 % of these calls)."""
import ctypes as ct

import numpy as np
import pytest

from _powlib import host_powers, libm_pow

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,y", [(0, 7.0 / 3.0), (1, 2.0 / 3.0)])
def test_device_powers_correctly_rounded(gf, which, y):
    print(f"case device powers which={which}", flush=True)
    rng = np.random.default_rng(100 + which)
    x = np.exp(rng.uniform(np.log(1e-3), np.log(20.0), 1 << 20))
    dev = np.zeros_like(x)
    lib = gf.load()
    assert lib.gf_eval_powers(which, len(x), gf.dptr(x), gf.dptr(dev)) == 0
    assert np.array_equal(dev, host_powers(which, x))
    sample = x[:100000]
    lm = libm_pow(sample, y)
    rate = float((dev[:100000] != lm).mean())
    print(f"device vs glibc pow: {rate:.5%} of 100000", flush=True)
    assert rate < 0.002
