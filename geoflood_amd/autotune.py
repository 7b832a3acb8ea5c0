"""Step-variant choices per population, kept across runs.

gf_pool_autotune times every step variant of a pool's patch shape (tile
shape, threads per tile, patches per workgroup, LDS split) on the pool's
current active patches and keeps the fastest.  Which variant wins depends on
how many patches are active (a few hundred tiles leave most of the chip idle
whatever the shape; tens of thousands fill it), so a run whose population
changes -- a flood front spreading over refined levels -- re-selects as the
population moves, and this cache remembers each choice by (patch size,
method, population bucket) so that a bucket seen before, in this run or an
earlier one (the JSON file), costs no timing at all.  The reference has one
fixed shape (cudaclaw_flux2.cu:68-120: one 128-thread block per patch).
"""
from __future__ import annotations

import json
import math
import os


class TuneCache:
    def __init__(self, path: str | None = None):
        self.path = path
        self.table: dict[str, dict] = {}
        self.log: list[dict] = []
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    @staticmethod
    def bucket(active: float) -> int:
        """Population bucket: log2 of the active patches, rounded (each bucket
        spans a factor of about 1.4 either side)."""
        return int(round(math.log2(max(active, 1.0))))

    @staticmethod
    def key(mx: int, method: str, active: float) -> str:
        return f"{mx}:{method}:{TuneCache.bucket(active)}"

    def choose(self, pool, active: float, method: str = "default", reps: int = 2, levels: dict | None = None):
        """Select the variant for this population: the cached choice, or an
        autotune whose result is cached (and saved when a path is set).
        Returns the variant index."""
        k = self.key(pool.mx, method, active)
        hit = k in self.table
        if hit:
            v = int(self.table[k]["variant"])
            pool.set_variant(v)
        else:
            v, ms = pool.autotune(reps=reps)
            self.table[k] = dict(variant=int(v), ms=[round(m, 5) for m in ms], active=round(float(active), 1),
                                 levels=levels or {})
            self.save()
        self.log.append(dict(key=k, active=round(float(active), 1), variant=int(v), cached=hit))
        return v

    def save(self):
        if self.path:
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
