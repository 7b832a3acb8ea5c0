"""The per-cell comparison used by the GPU parity tests flags what it must."""
import numpy as np

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
