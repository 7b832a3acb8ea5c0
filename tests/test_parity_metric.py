"""The per-cell comparison used by the GPU parity tests flags what it must."""
import numpy as np
import pytest

from _parity import state_diff


def _state():
    rng = np.random.default_rng(0)
    h = np.where(rng.random((16, 16)) < 0.5, 0.0, 1 + rng.random((16, 16)))
    return np.stack([h, rng.normal(0, 1, h.shape) * h, rng.normal(0, 1, h.shape) * h])


def test_identical_states():
    s = _state()
    assert state_diff(s, s.copy()) == (0.0, 0.0, True)


def test_mask_flip_is_caught():
    s = _state()
    t = s.copy()
    j, i = np.argwhere(s[0] == 0.0)[0]
    t[0, j, i] = 2e-3
    assert state_diff(s, t)[2] is False


def test_wet_relative_and_momentum_scale():
    s = _state()
    t = s.copy()
    j, i = np.argwhere(s[0] > 0)[0]
    t[0, j, i] *= 1 + 1e-9
    assert 5e-10 < state_diff(s, t)[0] < 2e-9
    # a momentum at rounding level in still water is measured against h sqrt(g h)
    u = s.copy()
    u[1, j, i] = 1e-17
    v = u.copy()
    v[1, j, i] = -1e-17
    assert state_diff(u, v)[0] < 1e-16


def test_dry_cells_absolute():
    s = _state()
    t = s.copy()
    j, i = np.argwhere(s[0] == 0.0)[0]
    t[0, j, i] = 5e-18
    rel, dry, mask = state_diff(s, t)
    assert mask and rel == 0.0 and 0 < dry < 1e-10


def test_bench_state_parity_matches_the_test_metric():
    """bench.py's parity_vs_cpu uses the floored per-cell metric of the
    parity tests: the same numbers on the same states, a near-dry cell at
    rounding level does not count as a relative difference of 1."""
    import bench
    from _parity import state_diff
    rng = np.random.default_rng(11)
    ref = np.stack([np.abs(rng.normal(1, 0.5, (40, 40))), rng.normal(0, 1, (40, 40)), rng.normal(0, 1, (40, 40))])
    dry = rng.random((40, 40)) < 0.2
    ref[:, dry] = 0.0                      # a dry cell carries no momentum after b4step2
    got = ref * (1 + 1e-13 * rng.normal(size=ref.shape))
    got[0][ref[0] == 0.0] = 5e-18
    r = bench.state_parity(got, ref)
    rel, dry, mask = state_diff(got, ref)
    assert r["mask_equal"] == mask is True
    assert r["wet_max_rel"] == pytest.approx(rel, rel=1e-12)
    assert r["dry_max_abs_over_drytol"] == pytest.approx(dry, rel=1e-12)
    assert r["wet_max_rel"] < 1e-12 and r["dry_max_abs_over_drytol"] < 1e-12
