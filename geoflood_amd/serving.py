"""Continuous-batching engine for ensembles of flood scenarios.

Requests (flood scenarios of different domain sizes) arrive over time.  Each
admitted scenario owns a contiguous range of slots of one device-resident
patch pool and its own device time-step controller, Manning coefficient and
hydrograph; every attempt of the engine is one replay of the pool's hipGraph,
which advances every active scenario by one step of its own dt in the same
persistent launch.  A request is admitted as soon as its HBM footprint (its
patch slots) fits in the pool; scenarios that reach their end time stop on
the device, and the host harvests them, frees their slots and admits the
next requests between replays.

Each scenario is an outstyle_3 run (fc2d_geoclaw_run.c:417-590) of the
single-level forest of ``overland_flood``: Gaussian-sum DEM, hydrograph
inflow on the west edge (bc2 mthbc = 0, hydrograph/hydrograh_module.f90),
extrapolation on the east edge, walls north and south.

    python -m geoflood_amd.serving --rate 32 --requests 2000
"""
from __future__ import annotations

import argparse
import json
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor
from concurrent.futures import wait as futures_wait
from dataclasses import dataclass, field

import numpy as np

from . import GeoFloodError, init, set_params
from .pool import MBC, BrickLayout, Pool, global_to_aos
from .scenarios import gaussian_sum_dem, sine_hydrograph

MX = 32
LENGTH = 2000.0      # every scenario covers the same 2 km x 2 km domain
T_END = 1800.0       # simulated seconds per scenario
T_START = 0.0        # every scenario starts dry at t = 0, where the hydrograph's Q(0) = 0
DT0 = 1.0
DT_MAX = 10.0        # dt bound: a step with no wave anywhere (CFL 0) is followed by dt = DT_MAX
STRIP = 100.0        # inflow channel width (m), centred on the west edge
RETUNE_FACTOR = 2.0  # re-run the autotuner when the live patches change by this factor


@dataclass
class Request:
    rid: int
    n: int            # cells per side (a multiple of 32)
    seed: int         # DEM seed
    manning: float
    peak_q: float     # m^3/s
    arrival: float    # seconds after the start of the stream
    t_end: float = T_END

    @property
    def npatch(self) -> int:
        return (self.n // MX) ** 2


@dataclass
class Result:
    rid: int
    n: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    retakes: int
    t: float
    max_depth: float
    wet_cells: int
    volume: float
    status: int = 0   # gf_scenario_info status: 0, or GF_SCEN_NONFINITE when it stopped on a NaN state


    def record(self) -> dict:
        """One JSON-lines record of the request: arrival, admission and finish
        times (stream seconds), latency, steps, retakes, the final state's
        summary and status."""
        return dict(rid=self.rid, n=self.n, arrival=round(self.arrival, 6), admitted=round(self.admitted, 6),
                    finished=round(self.finished, 6), latency=round(self.latency, 6), steps=self.steps,
                    retakes=self.retakes, t=self.t, cell_updates=self.steps * self.n * self.n,
                    max_depth=self.max_depth, wet_cells=self.wet_cells, volume=self.volume, status=self.status)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def make_requests(count: int, rate: float, seed: int = 0, sizes=(128, 256, 512, 1024),
                  weights=(0.4, 0.3, 0.2, 0.1), t_end: float = T_END) -> list[Request]:
    """Poisson arrivals at ``rate`` per second (seed ``seed``); domain size
    drawn with ``weights``; Manning n ~ U[0.02, 0.06]; inflow peak
    log-uniform in [20, 2000] m^3/s; request k uses DEM seed k."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    sizes_k = rng.choice(np.asarray(sizes), p=np.asarray(weights) / np.sum(weights), size=count)
    manning = rng.uniform(0.02, 0.06, count)
    peak = np.exp(rng.uniform(np.log(20.0), np.log(2000.0), count))
    return [Request(k, int(sizes_k[k]), k, float(manning[k]), float(peak[k]), float(arrivals[k]), t_end)
            for k in range(count)]


class SlotAllocator:
    """First-fit allocator of contiguous pool slot ranges."""

    def __init__(self, capacity: int):
        self.free = [(0, capacity)]  # sorted (start, length)

    def alloc(self, n: int):
        for k, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, ln - n)
                return s
        return None

    def release(self, start: int, n: int):
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free = merged

    def largest(self) -> int:
        return max((ln for _, ln in self.free), default=0)


@dataclass
class _Live:
    req: Request
    sid: int
    first: int
    layout: BrickLayout
    admitted: float


@dataclass
class Engine:
    """Serving loop over one pool of ``capacity`` 32x32 patch slots (the HBM
    budget: ``Pool.bytes_for(32, 32, capacity)`` bytes)."""
    capacity: int
    substeps: int = 16          # graph replays between host polls (each poll leaves the GPU idle)
    keep_states: bool = False   # keep each result's final (h, hu, hv)
    pool: Pool = field(init=False)
    states: dict = field(default_factory=dict)

    def __post_init__(self):
        self.pool = Pool(MX, MX, self.capacity)
        self.alloc = SlotAllocator(self.capacity)
        self.live: dict[int, _Live] = {}
        self.attempts = 0
        self.cell_updates = 0
        # host-side DEM/layout of queued requests (numpy releases the GIL in
        # the array work, so the threads run in parallel with the stepping)
        self.prep = ThreadPoolExecutor(max_workers=6)
        self.ready: dict[int, object] = {}
        self.tuned_at = 0  # live patches when the step variant was last chosen

    def close(self):
        self.prep.shutdown(wait=True)
        self.pool.close()

    @staticmethod
    def prepare(req: Request):
        """Host arrays of a request's initial state (dry, Gaussian-sum DEM) in
        the pool's upload layout."""
        m = req.n // MX
        dx = LENGTH / req.n
        layout = BrickLayout(m, m, MX, MX, dx, dx, mthbc=(0, 1, 3, 3))
        b = gaussian_sum_dem(req.n, LENGTH, seed=req.seed)
        z = np.zeros((req.n, req.n))
        q, aux = global_to_aos(z, z, z, b, layout)
        return layout, q, aux

    def enqueue(self, req: Request):
        self.ready[req.rid] = self.prep.submit(self.prepare, req)

    @staticmethod
    def footprint_bytes(req: Request) -> int:
        return Pool.bytes_for(MX, MX, req.npatch)

    def admit(self, req: Request, now: float, wait: bool = True) -> bool:
        """Place the request in free slots if its footprint fits (and, with
        wait False, only once its host-side preparation has finished, so that
        the engine keeps stepping the live scenarios meanwhile)."""
        if req.npatch > self.alloc.largest():
            return False
        pending = self.ready.get(req.rid)
        if not wait and pending is not None and not pending.done():
            return False
        fut = self.ready.pop(req.rid, None)
        layout, q, aux = fut.result() if fut is not None else self.prepare(req)
        first = self.alloc.alloc(req.npatch)
        self.pool.describe(layout, first)
        self.pool.upload_aos(first, q, aux, buf=0)
        sid = self.pool.add_scenario(layout, dt0=DT0, desired_cfl=0.9, max_cfl=1.0, t_end=req.t_end,
                                     t0=T_START)
        self.pool.set_dt_max(sid, DT_MAX)
        self.pool.set_friction(sid, req.manning)
        t, hu = sine_hydrograph(req.peak_q, STRIP)
        self.pool.set_inflow(sid, t, hu, STRIP, 0.0, 0.5 * LENGTH)
        self.live[sid] = _Live(req, sid, first, layout, now)
        return True

    def _harvest(self, sid: int, info, now: float) -> Result:
        lv = self.live.pop(sid)
        h, hu, hv = self.pool.download_global(lv.layout, buf=int(info["cur"]))
        if self.keep_states:
            self.states[lv.req.rid] = np.stack([h, hu, hv])
        dx = lv.layout.dx
        self.pool.remove_scenario(sid)
        self.alloc.release(lv.first, lv.req.npatch)
        self.cell_updates += int(info["nsteps"]) * lv.req.n * lv.req.n
        return Result(lv.req.rid, lv.req.n, lv.req.arrival, lv.admitted, now, int(info["nsteps"]),
                      int(info["retakes"]), float(info["t"]), float(np.nanmax(h)), int((h > 1e-3).sum()),
                      float(np.nansum(h) * dx * dx), int(info["status"]))

    def poll(self, now: float) -> list[Result]:
        if not self.live:
            return []
        hi = max(self.live) + 1
        recs = self.pool.scenarios(0, hi)
        done = []
        for sid in list(self.live):
            if not recs[sid]["active"]:
                done.append(self._harvest(sid, recs[sid], now))
        return done

    def serve(self, requests: list[Request], clock=None, log=None) -> list[Result]:
        """Run the stream to completion.  ``clock`` maps the engine's
        progress to stream time: None uses the wall clock (arrivals are
        admitted once their arrival time has passed); a callable taking the
        number of attempts so far gives a deterministic schedule."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue: list[Request] = []
        results: list[Result] = []
        t_start = time.perf_counter()

        def now():
            return (time.perf_counter() - t_start) if clock is None else clock(self.attempts)

        while pending or queue or self.live:
            tnow = now()
            while pending and pending[0].arrival <= tnow:
                queue.append(pending.pop(0))
                self.enqueue(queue[-1])
            # admit in arrival order; a request that does not fit waits, later
            # smaller ones may pass it (backfill).  On the wall clock a request
            # whose DEM is still being prepared waits the same way; on a
            # deterministic clock admission waits for the preparation.
            still = []
            for r in queue:
                if r.npatch > self.capacity:
                    raise ValueError(f"request {r.rid} needs {r.npatch} slots > capacity {self.capacity}")
                if not self.admit(r, tnow, wait=clock is not None):
                    still.append(r)
            queue = still
            if not self.live:
                if clock is None:
                    # nothing to step: wait for a queued request's preparation
                    # or for the next arrival
                    waiting = [self.ready[r.rid] for r in queue if r.rid in self.ready]
                    if waiting:
                        futures_wait(waiting, timeout=0.01, return_when=FIRST_COMPLETED)
                    elif pending:
                        time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
                elif pending:
                    self.attempts += 1  # idle tick of the deterministic clock
                continue
            live = sum(lv.req.npatch for lv in self.live.values())
            if live > RETUNE_FACTOR * self.tuned_at or RETUNE_FACTOR * live < self.tuned_at:
                try:
                    self.pool.autotune(reps=1)  # choose the step variant for this population
                except GeoFloodError:
                    pass                        # nothing listed to time (all steady): keep the variant
                self.tuned_at = live
            self.pool.run(self.substeps, graph=True)
            self.attempts += self.substeps
            done = self.poll(now())
            results.extend(done)
            if log and done:
                log(f"t={now():8.3f}s done {len(results)} live {len(self.live)} queued {len(queue)}")
        self.wall = time.perf_counter() - t_start
        return results


def summarize(results: list[Result], wall: float, cell_updates: int) -> dict:
    lat = np.array([r.latency for r in results])
    out = dict(requests=len(results), failed=sum(1 for r in results if r.status != 0), wall_s=round(wall, 3),
               p50_latency_s=round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
               p99_latency_s=round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
               mean_latency_s=round(float(lat.mean()), 4) if len(lat) else None,
               cell_updates_per_s=cell_updates / wall if wall > 0 else None)
    by = {}
    for n in sorted({r.n for r in results}):
        ln = np.array([r.latency for r in results if r.n == n])
        by[str(n)] = dict(count=int(len(ln)), p50=round(float(np.percentile(ln, 50)), 4),
                          p99=round(float(np.percentile(ln, 99)), 4))
    out["by_size"] = by
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--rate", type=float, default=32.0)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--capacity", type=int, default=65536, help="pool slots (32x32 patches)")
    ap.add_argument("--substeps", type=int, default=16)
    ap.add_argument("--t-end", type=float, default=T_END)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--jsonl", default=None, help="write one JSON line per finished request to this file")
    args = ap.parse_args()
    init(0)
    set_params(manning_n=0.035)
    reqs = make_requests(args.requests, args.rate, seed=args.seed, t_end=args.t_end)
    eng = Engine(args.capacity, substeps=args.substeps)
    res = eng.serve(reqs, log=lambda m: print(m, flush=True))
    if args.jsonl:
        with open(args.jsonl, "w") as f:
            for r in sorted(res, key=lambda r: r.rid):
                f.write(json.dumps(r.record()) + "\n")
    s = summarize(res, eng.wall, eng.cell_updates)
    s.update(rate=args.rate, capacity_patches=args.capacity,
             hbm_budget_bytes=Pool.bytes_for(MX, MX, args.capacity), t_end=args.t_end)
    print(json.dumps(s))
    eng.close()


if __name__ == "__main__":
    main()
