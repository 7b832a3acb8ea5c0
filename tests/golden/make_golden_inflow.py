"""Write tests/golden/inflow_vectors.npz from the reference hydrograph module
and bc2 (hydrograph/hydrograh_module.f90, fortran_source/
geoflood_bc2_fort.f90) compiled by oracle/Makefile (target ``ref``).

Cases: the discharge interpolation over a sine hydrograph table, the
Riemann-invariant and two-shock Newton solves on random boundary states, and
bc2 on random patches with one hydrograph side (mthbc = 0) combined with
walls, extrapolation and interior sides.

    make -C oracle ref && python tests/golden/make_golden_inflow.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Hydrograph, Params, RefLib  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inflow_vectors.npz")
MX, MY, MBC = 16, 12, 2


def table():
    t = np.linspace(0.0, 3600.0, 61)
    return t, 2.0 * np.sin(np.pi * t / 3600.0)


def random_patch(rng):
    shape = (MY + 2 * MBC, MX + 2 * MBC)
    q = np.zeros(shape + (3,))
    h = np.where(rng.random(shape) < 0.4, 0.0, rng.uniform(0, 3, shape))
    h[rng.random(shape) < 0.1] = 1e-3
    q[..., 0] = h
    q[..., 1] = rng.normal(0, 1, shape)
    q[..., 2] = rng.normal(0, 1, shape)
    aux = rng.normal(0, 1, shape + (1,))
    return q, aux


def main():
    ref = RefLib()
    ref.set_params(Params())
    rng = np.random.default_rng(20240917)
    out = {}
    t, hu = table()
    out["table_t"], out["table_hu"] = t, hu

    # interpolation (the channel geometry does not enter)
    ref.set_hydrograph(Hydrograph(t, hu, 100.0, 0.0, 1000.0, (1, 0, 0, 0)))
    ts = np.concatenate([rng.uniform(-100.0, 3700.0, 300), t, [-1.0, 3601.0]])
    out["interp_t"] = ts
    out["interp_hu"] = np.array([ref.hydro_interpolate(x) for x in ts])

    # Newton solves: q0 = (0, hu0), q1 = (h1, hu1)
    n = 400
    q0 = np.stack([np.zeros(n), rng.uniform(0.001, 30.0, n)], 1)
    q1 = np.stack([rng.uniform(0.001, 6.0, n), rng.uniform(-5.0, 15.0, n)], 1)
    out["newton_q0"], out["newton_q1"] = q0, q1
    out["ri_h0"] = np.array([ref.riemann_invariants(a, b)[0] for a, b in zip(q0, q1)])
    out["ts_h0"] = np.array([ref.two_shock(a, b)[0] for a, b in zip(q0, q1)])

    # bc2 with one hydrograph side
    ncase = 48
    keys = ["q_in", "aux_in", "q_out", "aux_out", "mthbc", "sides", "geom", "strip"]
    rec = {k: [] for k in keys}
    for k in range(ncase):
        side = k % 4
        sides = [0, 0, 0, 0]
        sides[side] = 1
        mthbc = [int(rng.choice([-1, 1, 3])) for _ in range(4)]
        mthbc[side] = 0
        strip = np.array([rng.uniform(2, 30), rng.uniform(-5, 20), rng.uniform(-5, 20)])
        ref.set_hydrograph(Hydrograph(t, hu, strip[0], strip[1], strip[2], tuple(sides)))
        q, aux = random_patch(rng)
        tt = float(rng.choice([0.0, 3600.0, rng.uniform(0, 3600), -5.0]))
        geom = np.array([rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(0.5, 2),
                         rng.uniform(0.5, 2), tt])
        qo, ao = q.copy(), aux.copy()
        ref.bc2_ex(MX, MY, MBC, qo, ao, mthbc, *geom)
        for name, v in zip(keys, [q, aux, qo, ao, mthbc, sides, geom, strip]):
            rec[name].append(np.asarray(v))
    for name in keys:
        out["bc2_" + name] = np.stack(rec[name])
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
