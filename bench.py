"""Benchmark: FP64 cell-updates/s of the GeoFlood patch step on one MI355X.

Workload (uniform large batch): 4096 x 4096 cells as 16,384 patches of
32 x 32 (mbc = 2), single level, seeded fractal DEM (spectral slope -2,
RMS 3 m, seed 0) plus a 1-in-500 slope descending in +x, dx = 1 m; dam-break
reservoir h = 10 m over the upper-left eighth, dry elsewhere; Manning
n = 0.03; walls; initial dt 0.01 s, desired CFL 0.9, max CFL 1.0.

A "step" is one step attempt over all 16,384 patches: device work lists
(patches whose 2-cell halo holds water are active, the rest dry; dry patches
that stayed below dry_tolerance for two steps are steady and left as they
are), one persistent kernel that steps the active patches (halo pulled from
the neighbours) and then the dry ones, and the device dt controller.  One
cell-update is one interior cell advanced by one accepted step (retaken
attempts do not count).

Runs, all from the same initial state, inputs resident in HBM:
  batched  -- W warm-up attempts (graph capture, autotuning of the step
              variant on the population they reach), then K timed attempts
              that continue the same trajectory, each one hipGraph launch
              (the reported value);
  events   -- the same W + K attempts, the K launched without a graph, with
              HIP events around every fused-step launch (per-kernel time ->
              roofline);
  plain    -- the first accepted steps with one kernel launch per patch from
              a host loop and the dt rule on the host;
  cpu      -- the reference GeoClaw Fortran step (oracle/_ref) on the host
              cores over the same first steps, when it is present (the
              reported cpu_baseline), then the C restatement on the same
              sample (cpu_port); the plain run's state is compared with the
              CPU state cell by cell (parity_vs_cpu);
  and, for context: the ensemble-serving stream (2,000 Poisson requests at
  32/s by default), the host-resident drop-in
  batch (PCIe included), ASCII frame and gauge output, and the regrid and
  coarse/fine ghost transfers.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s at 4096² (batched vs plain) + ensemble p50/p99 latency"
N = 4096
MX = 32
BYTES_PER_CELL_UPDATE = 56  # read h, hu, hv, b; write h, hu, hv (FP64)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (FMA = 2 FLOPs)
VALU_LANE_PEAK_T = 39.3          # 256 CUs x 64 lanes x 2.4 GHz, lane-instructions per second (T)
SIMDS = 1024                     # 256 CUs x 4 SIMDs
PARAMS = dict(grav=9.81, dry_tolerance=1e-3, manning_n=0.03, friction_depth=1.0e6,
              friction_forcing=1, src_term=1, order=2, transverse=2, mthlim=(4, 4, 4))
DT0, DESIRED_CFL, MAX_CFL = 0.01, 0.9, 1.0
RETUNE_FACTOR = 1.5


def profile_values(name):
    """key=value lines of a summary under profiles/ (empty when absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    return dict(line.split("=", 1) for line in open(path).read().split("\n")
                if "=" in line and " " not in line.split("=", 1)[0])


def state_parity(got, ref, drytol=1e-3, grav=9.81):
    """Per-cell comparison of two (3, ny, nx) states: the largest relative
    difference over wet cells (h against max |h|; hu, hv against the larger
    of max |hu| and h sqrt(g h)), the largest dry-cell difference in units of
    drytol, and whether the wet/dry masks agree."""
    wg, wr = got[0] > drytol, ref[0] > drytol
    wet, dry = wg & wr, ~wg & ~wr
    h = np.maximum(np.abs(got[0][wet]), np.abs(ref[0][wet]))
    rel = float(np.max(np.abs(got[0][wet] - ref[0][wet]) / h)) if wet.any() else 0.0
    for m in (1, 2):
        den = np.maximum(np.maximum(np.abs(got[m][wet]), np.abs(ref[m][wet])), h * np.sqrt(grav * h))
        if wet.any():
            rel = max(rel, float(np.max(np.abs(got[m][wet] - ref[m][wet]) / den)))
    dry_abs = max(float(np.max(np.abs(got[m][dry] - ref[m][dry]))) for m in range(3)) / drytol if dry.any() else 0.0
    return dict(wet_max_rel=rel, dry_max_abs_over_drytol=dry_abs, mask_equal=bool(np.array_equal(wg, wr)))


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_state():
    from geoflood_amd.scenarios import uniform_dam_break
    return uniform_dam_break(N, dx=1.0, h0=10.0, rms=3.0, regional_slope=1.0 / 500.0, seed=0)


def reset(pool, layout, state, target_steps):
    h, hu, hv, b, dx = state
    pool.clear_scenarios()
    pool.upload_global(layout, h, hu, hv, b, buf=0)
    sid = pool.add_scenario(layout, dt0=DT0, desired_cfl=DESIRED_CFL, max_cfl=MAX_CFL,
                            target_steps=target_steps)
    return sid


def cpu_baseline(state, nsteps, nthreads, port_only=False, cr_pow=False):
    """The reference Fortran step (or, where it was not built or port_only,
    the C restatement) over the first nsteps accepted steps of the same
    workload.  cr_pow: the C restatement evaluates its two powers correctly
    rounded, as the device does."""
    from oracle.oracle import Brick, CLib, Params, ref_lib
    h, hu, hv, b, dx = state
    p = Params(manning_n=PARAMS["manning_n"])
    c = CLib()
    c.set_params(p)
    c.set_pow_mode(cr_pow)
    ref = None if port_only else ref_lib()
    kind = "port"
    if ref is not None:
        ref.set_params(p)
        kind = "reference"
    br = Brick.from_global(h, hu, hv, b, N // MX, N // MX, 2, dx, dx)
    t0 = time.perf_counter()
    rec = c.brick_run(br, nsteps, DT0, DESIRED_CFL, MAX_CFL, update=ref, nthreads=nthreads)
    el = time.perf_counter() - t0
    c.set_pow_mode(False)
    return dict(value=nsteps * N * N / el, seconds=el, kind=kind, rec=rec, state=br.to_global())


def amr_overland(steps, maxlevel=4, tune=True):
    """The AMR overland flood: 2 km square, 16 x 16 patches at levels 1-4
    (level L: 4 * 2^L patches a side, 1.95 m cells at level 4), Gaussian-sum
    DEM (seed 0), west hydrograph Q = 200 sin(pi t / 3600) m^3/s through a
    100 m strip, east extrapolation, north/south walls, Manning n 0.035,
    tagged on h > dry_tolerance or |grad eta| > 0.01, regrid every 4 steps,
    dry at t = 0 (dt bound 10 s).  Every attempt is one graph replay (ghost
    update across levels, one persistent step of all levels, dt control);
    the whole loop, regrids included, is timed.  With tune, the step variant
    is selected by timing every variant on the population the warm-up
    reaches (and again if the population moves to another bucket)."""
    import geoflood_amd
    from geoflood_amd.amr import AmrRun
    from geoflood_amd.autotune import TuneCache
    from geoflood_amd.scenarios import gaussian_sum_dem, sine_hydrograph
    geoflood_amd.set_params(**dict(PARAMS, manning_n=0.035))
    t, hu = sine_hydrograph(200.0, 100.0)
    hydro = dict(time=t, hu=hu, b=100.0, x0=0.0, y0=1000.0)
    cache = TuneCache() if tune else None
    run = AmrRun(16, 4, 4, 2000.0, 2000.0, 1, maxlevel, lambda nx, ny, L: gaussian_sum_dem(nx, 2000.0, seed=0),
                 4 ** maxlevel * 16, mthbc=(0, 1, 3, 3), hydro=hydro, eta_tol=0.01, regrid_interval=4, dt_max=10.0,
                 tune_cache=cache)
    try:
        t0 = time.perf_counter()
        run.start(0.0, 1.0)
        t_init = time.perf_counter() - t0
        counts0 = run.forest.counts()
        log(f"amr: initial forest {len(run.forest)} leaves {counts0} in {t_init:.2f} s; {steps} steps ...")
        run.run(8)  # warm-up: graph capture, first regrids
        c0, s0, d0 = run.stats["cell_updates"], run.stats["steps"], run.stats["device_ms"]
        p0, r0 = len(run.stats["patches"]), run.stats["regrids"]
        st0 = run.pool.stats()
        t0 = time.perf_counter()
        info = run.run(steps)
        wall = time.perf_counter() - t0
        st1 = run.pool.stats()
        cu = run.stats["cell_updates"] - c0
        pts = run.stats["patches"][p0:]
        q, _ = run.leaves_state()
        wet = int((q[:, 2:-2, 2:-2, 0] > 1e-3).sum())
        dev_s = (run.stats["device_ms"] - d0) * 1e-3
        out = dict(value=cu / wall, unit="cell-updates/s", steps=run.stats["steps"] - s0,
                   device_s=round(dev_s, 4), device_cell_updates_per_s=cu / dev_s if dev_s > 0 else None,
                   attempts=info["attempts"], retakes=info["retakes"], wall_s=round(wall, 3),
                   simulated_time_s=round(info["t"], 3),
                   regrids=run.stats["regrids"] - r0, ms_per_step=round(1e3 * wall / max(run.stats["steps"] - s0, 1), 4),
                   patches=dict(min=int(min(pts)), mean=round(float(np.mean(pts)), 1), max=int(max(pts)),
                                final_by_level=run.forest.counts(), initial_by_level=counts0),
                   active_patches_per_attempt=round((st1["active_patches"] - st0["active_patches"]) /
                                                    max(st1["attempts"] - st0["attempts"], 1), 1),
                   wet_cells_final=wet,
                   variant=list(run.pool.variants()[1][run.pool.variants()[0]]),
                   tuning=[dict(e) for e in cache.log] if cache else None,
                   note="whole loop timed: graph replays (multi-level ghost update, one persistent step of "
                        "every level, device dt control) and every regrid (device tags, forest adapt on the host, "
                        "slot moves, interpolate2fine, average2coarse, scenario re-targeting)")
        log(f"amr: {out['steps']} steps in {wall:.2f} s ({out['ms_per_step']} ms/step), "
            f"{cu / wall / 1e9:.3f} G cell-updates/s, patches {out['patches']['min']}-{out['patches']['max']}, "
            f"{out['regrids']} regrids, t = {info['t']:.1f} s")
        return out
    finally:
        run.close()
        geoflood_amd.set_params(**PARAMS)


_DEM_CACHE = {}


def _dem2km(nx, ny, L):
    """The overland-flood DEM on an nx x nx lattice over 2 km (cached: every
    patch size of the sweep uses the same level lattices)."""
    from geoflood_amd.scenarios import gaussian_sum_dem
    if nx not in _DEM_CACHE:
        _DEM_CACHE[nx] = gaussian_sum_dem(nx, 2000.0, seed=0)
    return _DEM_CACHE[nx]


def autotune_sweep(steps, sizes=(8, 16, 32, 64), maxlevel=5):
    """Autotuner sweep under dynamic AMR: for each patch size mx, a 2 km
    square of level-0 bricks of 64/mx patches (levels 1-5, finest cells
    0.98 m at every mx), the overland-flood DEM, walls, a 30 m deep
    reservoir over x < 250 m, y >= 1750 m released at t = 0 (Manning n
    0.035); leaves tagged where wet, regrid every 2 steps; the step variant
    (tile shape, threads per tile, patches per workgroup, LDS split) is
    re-selected whenever the active population per attempt moves to another
    power-of-two bucket, from a cache of earlier choices (persisted to a JSON
    file for the run) or by timing every variant."""
    import tempfile
    import geoflood_amd
    from geoflood_amd.amr import AmrRun
    from geoflood_amd.autotune import TuneCache
    from geoflood_amd.scenarios import reservoir_state
    geoflood_amd.set_params(**dict(PARAMS, manning_n=0.035))
    out = {}
    cache_path = os.path.join(tempfile.mkdtemp(prefix="gf_tune_"), "variants.json")
    cache = TuneCache(cache_path)
    try:
        for mx in sizes:
            nx0 = 64 // mx
            cap = (nx0 << maxlevel) ** 2
            run = AmrRun(mx, nx0, nx0, 2000.0, 2000.0, 1, maxlevel, _dem2km, cap, eta_tol=1e30,
                         regrid_interval=2, tune_cache=cache)
            try:
                run.start(0.0, 0.01, state_fn=reservoir_state(run))
                n0 = len(run.forest)
                _, variants = run.pool.variants()
                t0 = time.perf_counter()
                info = run.run(steps)
                wall = time.perf_counter() - t0
                pts = run.stats["patches"]
                tuned = [dict(e, variant_desc=list(variants[e["variant"]])) for e in cache.log if e["key"].startswith(f"{mx}:")]
                timings = {k: v for k, v in cache.table.items() if k.startswith(f"{mx}:")}
                gain = []
                for k, v in sorted(timings.items(), key=lambda kv: int(kv[0].split(":")[-1])):
                    ms = v["ms"]
                    gain.append(dict(bucket=int(k.split(":")[-1]), active=v["active"], chosen=v["variant"],
                                     chosen_ms=ms[v["variant"]], variant0_ms=ms[0],
                                     speedup_vs_variant0=round(ms[0] / ms[v["variant"]], 3), ms_by_variant=ms))
                dev_s = run.stats["device_ms"] * 1e-3
                out[str(mx)] = dict(steps=info["nsteps"], attempts=info["attempts"], retakes=info["retakes"],
                                    wall_s=round(wall, 3), device_s=round(dev_s, 4),
                                    simulated_time_s=round(info["t"], 3),
                                    cell_updates_per_s=run.stats["cell_updates"] / wall,
                                    device_cell_updates_per_s=run.stats["cell_updates"] / dev_s if dev_s > 0 else None,
                                    patches=dict(initial=n0, final=len(run.forest), max=int(max(pts)),
                                                 final_by_level=run.forest.counts()),
                                    selections=tuned, tuning_by_bucket=gain,
                                    variants=[list(v) for v in variants])
                log(f"sweep mx={mx}: {info['nsteps']} steps in {wall:.1f} s, patches {n0} -> {len(run.forest)} "
                    f"(max {max(pts)}), {run.stats['cell_updates'] / wall / 1e9:.3f} G cell-updates/s, "
                    f"selections {[(e['key'].split(':')[-1], e['variant'], e['cached']) for e in tuned]}")
            finally:
                run.close()
    finally:
        geoflood_amd.set_params(**PARAMS)
    out["cache_entries"] = len(cache.table)
    out["note"] = ("variant = (tile x, tile y, threads per tile, patches per workgroup, x-face increments in "
                   "registers 1 / LDS 0); speedup_vs_variant0: the persistent step kernel's time with the chosen "
                   "variant against the fixed default one-workgroup-per-patch shape, timed by the autotuner on the "
                   "same population")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=24,
                    help="accepted steps of the per-patch launch loop (compared with the CPU leg when equal)")
    ap.add_argument("--output-patches", type=int, default=1024,
                    help="patches of the final state written as one ASCII frame (0: skip)")
    ap.add_argument("--patch-size", type=int, default=32, choices=(8, 16, 32, 64),
                    help="patch size of the workload (the same 4096^2 cells)")
    ap.add_argument("--regrid-families", type=int, default=1024,
                    help="coarse patches refined and averaged back in the regrid measurement (0: skip)")
    ap.add_argument("--host-steps", type=int, default=2,
                    help="steps of the host-resident drop-in batch (gf_step2_batch over all patches)")
    ap.add_argument("--retune-every", type=int, default=100,
                    help="attempts between checks of the active population (autotune on a change)")
    ap.add_argument("--cpu-steps", type=int, default=24,
                    help="accepted steps of the CPU baseline sample (about 10 s on 16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-requests", type=int, default=2000)
    ap.add_argument("--ensemble-rate", type=float, default=32.0)
    ap.add_argument("--ensemble-capacity", type=int, default=65536)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--sweep-steps", type=int, default=3000,
                    help="accepted steps per patch size of the autotuner sweep under dynamic AMR (0: skip; "
                         "the workload's full length is 3000)")
    ap.add_argument("--amr-steps", type=int, default=2000,
                    help="accepted steps of the AMR overland flood (levels 1-4, regrid every 4; 0: skip)")
    ap.add_argument("--amr-fixed-variant", action="store_true",
                    help="AMR overland flood with the default step variant (no autotune)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="use this step variant instead of the autotuner's choice (counter runs, where "
                         "the profiler's serialisation makes the autotuner's timings unreliable)")
    ap.add_argument("--no-events", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy kernel (torch) three times; "
                         "used to scale the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    global MX
    MX = args.patch_size

    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of this path is a design, not a measured run)",
              file=sys.stderr)
        sys.exit(2)

    import geoflood_amd
    from geoflood_amd.pool import BrickLayout, Pool

    geoflood_amd.init(0)
    geoflood_amd.set_params(**PARAMS)
    log("building workload ...")
    state = build_state()
    mi = N // MX
    layout = BrickLayout(mi, mi, MX, MX, state[4], state[4], (3, 3, 3, 3))
    pool = Pool(MX, MX, mi * mi)
    pool.set_layout([layout])
    cells = N * N

    # ---- warm-up (includes graph capture), then the timed batched run ----
    # The warm-up attempts are the first W attempts of the workload (the
    # first two also step every dry patch, before those become steady); the
    # step variant (tile shape, threads per workgroup) is then chosen by the
    # autotuner on the population they reached (it does not advance the
    # state), and chosen again inside the timed run whenever the active
    # population has changed by more than RETUNE_FACTOR.  The K timed
    # attempts continue the same trajectory.
    sid = reset(pool, layout, state, 0)
    pool.run(args.warmup, graph=True)
    best, tune_ms = pool.autotune(reps=3)
    _, variants = pool.variants()
    if args.variant >= 0:
        best = args.variant
        pool.set_variant(best)
    log(f"autotune: variants {variants}, k_all ms {['%.4f' % m for m in tune_ms]}, using {best}")
    pool.run(0, graph=True)  # capture outside the timed region
    geoflood_amd.load().gf_synchronize()
    info_w = pool.scenario(sid)
    log(f"timed batched run: {args.steps} attempts after {args.warmup} warm-up attempts "
        f"(t = {info_w['t']:.4f} s) ...")
    retunes = []
    t0 = time.perf_counter()
    done, last_act, st_prev = 0, None, pool.stats()
    st_start = st_prev
    while done < args.steps:
        k = min(args.retune_every, args.steps - done)
        pool.run(k, graph=True)
        done += k
        st = pool.stats()
        act = (st["active_patches"] - st_prev["active_patches"]) / max(st["attempts"] - st_prev["attempts"], 1)
        st_prev = st
        if last_act is None:
            last_act = act
        elif done < args.steps and args.variant < 0 and (act > RETUNE_FACTOR * last_act
                                                         or act * RETUNE_FACTOR < last_act):
            v, _ = pool.autotune(reps=2)
            retunes.append(dict(after_attempts=done, active_patches=round(act), variant=v))
            last_act = act
    wall = time.perf_counter() - t0
    total_ms, _, _ = geoflood_amd.last_timing()
    info_b = pool.scenario(sid)
    st_end = pool.stats()
    active_per_attempt = (st_end["active_patches"] - st_start["active_patches"]) / max(
        st_end["attempts"] - st_start["attempts"], 1)
    dry_per_attempt = (st_end["dry_patches"] - st_start["dry_patches"]) / max(
        st_end["attempts"] - st_start["attempts"], 1)
    accepted = info_b["nsteps"] - info_w["nsteps"]
    value = accepted * cells / wall
    ms_per_step = 1e3 * wall / args.steps
    log(f"batched: {active_per_attempt:.1f} active + {dry_per_attempt:.1f} dry patches per attempt")
    log(f"batched: {accepted} accepted / {args.steps} attempts, {wall:.3f} s wall, "
        f"{total_ms:.1f} ms device, t = {info_b['t']:.4f} s, {value / 1e9:.2f} G cell-updates/s")

    # ---- same attempts without a graph, HIP events around each launch ----
    # The dominant kernel is k_all, the persistent step over the attempt's
    # active tiles and then its non-steady dry patches.  Its algorithmic bytes
    # per launch are 56 B x MX*MX cells x the patches it reads and writes
    # (active + dry - steady): steady dry patches, whose step leaves both
    # buffers as they are, count as advanced in the cell-update rate but are
    # neither read nor written, so they are left out of the bandwidth figure
    # (the all-cells figure is reported separately as "effective").
    roofline = None
    composite = None
    if not args.no_events:
        sid = reset(pool, layout, state, 0)
        pool.run(args.warmup, graph=True)
        s0 = pool.stats()
        t0e = time.perf_counter()
        pool.run(args.steps, graph=False)
        ev_wall = time.perf_counter() - t0e
        ev_total, step_ms, launches = geoflood_amd.last_timing()
        s1 = pool.stats()
        info_e = pool.scenario(sid)
        act = s1["active_patches"] - s0["active_patches"]
        dry = s1["dry_patches"] - s0["dry_patches"]
        steady = s1["steady_patches"] - s0["steady_patches"]
        launches = max(launches, 1)
        k_avg_ms = step_ms / launches
        touched = (act + dry - steady) / launches
        k_bytes = BYTES_PER_CELL_UPDATE * MX * MX * touched
        achieved = k_bytes / (k_avg_ms * 1e-3) / 1e9
        eff_bytes = BYTES_PER_CELL_UPDATE * MX * MX * (act + dry) / launches
        effective = eff_bytes / (k_avg_ms * 1e-3) / 1e9
        # measured HBM bytes and VALU work per launch (rocprofv3 --pmc over
        # the 20 event-timed launches of the --steps 20 --warmup 5 trajectory,
        # profiles/); another trajectory scales the bytes by its touched
        # patches per launch
        traffic = valu = None
        kv = profile_values("k_all_32x32_uniform4096_traffic.txt") if MX == 32 else {}
        if "hbm_bytes_per_launch" in kv:
            traffic = float(kv["hbm_bytes_per_launch"])
            if "touched_patches_per_launch" in kv and touched > 0:
                traffic *= touched / float(kv["touched_patches_per_launch"])
            traffic = round(traffic)
        kv = profile_values("k_all_32x32_uniform4096_valu.txt") if MX == 32 else {}
        valu_issue_frac = fp64_frac = None
        if "fp64_flops_per_launch" in kv:
            fl = float(kv["fp64_flops_per_launch"]) / (k_avg_ms * 1e-3) / 1e12
            li = float(kv["valu_lane_instr_per_launch"]) / (k_avg_ms * 1e-3) / 1e12
            valu_issue_frac, fp64_frac = round(li / VALU_LANE_PEAK_T, 4), round(fl / FP64_PEAK_TFLOPS, 4)
            valu = dict(fp64_tflops=round(fl, 3), fp64_peak_tflops=FP64_PEAK_TFLOPS,
                        fp64_frac=fp64_frac, valu_lane_tinstr_per_s=round(li, 3),
                        valu_lane_peak_tinstr_per_s=VALU_LANE_PEAK_T, valu_issue_frac=valu_issue_frac,
                        note="FP64 FLOPs (FMA = 2) and VALU lane-instructions per launch from the "
                             "profiles/k_all_32x32_uniform4096_valu.txt counters, over this run's k_all time")
            if "valu_active_quad_cycles_per_launch" in kv and "clock_ghz" in kv:
                # SQ_ACTIVE_INST_VALU over the SIMD quad-cycles of this run's
                # k_all time at the clock measured in the counter runs
                simd_quads = SIMDS * k_avg_ms * 1e-3 * float(kv["clock_ghz"]) * 1e9 / 4.0
                valu["valu_busy_frac"] = round(float(kv["valu_active_quad_cycles_per_launch"]) / simd_quads, 4)
                valu["clock_ghz"] = float(kv["clock_ghz"])
        # The active tiles are bound by the FP64 VALU (two augmented Riemann
        # solves with ~17 correctly rounded divisions and ~8 square roots per
        # face): "bound" names that; achieved / frac are the HBM figures of the
        # bytes the launch really reads and writes.
        roofline = dict(bound="valu", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                        valu_issue_frac=valu_issue_frac, fp64_frac=fp64_frac,
                        valu_busy_frac=(valu or {}).get("valu_busy_frac"),
                        kernel="k_all<%d,%d,%d,%d,%d,%d,%d>" % ((MX, MX) + tuple(variants[pool.variants()[0]])),
                        kernel_avg_ms=round(k_avg_ms, 5), bytes_per_launch=round(k_bytes),
                        touched_patches_per_launch=round(touched, 1),
                        active_patches_per_launch=round(act / launches, 1),
                        dry_patches_per_launch=round(dry / launches, 1),
                        steady_patches_per_launch=round(steady / launches, 1), launches=launches,
                        traffic_over_algorithmic=round(traffic / k_bytes, 3) if traffic else None,
                        effective=dict(achieved=round(effective, 2), frac=round(effective / HBM_PEAK_GBS, 4),
                                       bytes_per_launch=round(eff_bytes),
                                       note="56 B for every cell the launch advances, steady dry patches "
                                            "(not read or written) included"),
                        valu=valu)
        composite = dict(attempt_avg_ms_without_graph=round(1e3 * ev_wall / launches, 5),
                         note="inflow (when present) -> classify -> k_all -> control per attempt; "
                              "the timed value uses one hipGraph replay per attempt")
        same = (info_e["nsteps"] == info_b["nsteps"] and info_e["t"] == info_b["t"])
        log(f"events run: k_all avg {k_avg_ms:.4f} ms ({act / launches:.0f} active + {dry / launches:.0f} dry "
            f"patches, {steady / launches:.0f} of them steady), {achieved:.1f} GB/s of the touched patches' "
            f"algorithmic bytes ({effective:.1f} GB/s all cells); same trajectory as graph run: {same}")

    # ---- output: one ASCII frame of 1,024 patches and 4,096 gauge samples ----
    # of the final state (fort.qNNNN bytes in the reference's format, written
    # to a scratch directory that is removed afterwards)
    output = None
    if args.output_patches > 0:
        import shutil
        import tempfile
        info_o = pool.scenario(sid)
        m = min(args.output_patches, layout.npatch)
        d = tempfile.mkdtemp(prefix="gf_frame_")
        try:
            t0o = time.perf_counter()
            pool.write_frame(d, 1, info_o["t"], first=0, n=m, buf=info_o["cur"], ngrids=layout.npatch)
            wo = time.perf_counter() - t0o
            nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        finally:
            shutil.rmtree(d, ignore_errors=True)
        rng = np.random.default_rng(0)
        ng = 4096
        gs = rng.integers(0, layout.npatch, ng)
        ox, oy = layout.origins()
        gx = ox[gs] + rng.uniform(0, MX, ng) * layout.dx
        gy = oy[gs] + rng.uniform(0, MX, ng) * layout.dy
        pool.sample_gauges(gs, gx, gy, buf=info_o["cur"])
        t0g = time.perf_counter()
        for _ in range(10):
            pool.sample_gauges(gs, gx, gy, buf=info_o["cur"])
        wg = (time.perf_counter() - t0g) / 10
        output = dict(frame_patches=m, frame_bytes=nbytes, frame_seconds=round(wo, 4),
                      frame_MB_per_s=round(nbytes / wo / 1e6, 1), gauges=ng, gauge_sample_ms=round(1e3 * wg, 4),
                      note="gf_pool_write_frame (device pack and flush, D2H, formatting on host threads) and "
                           "gf_pool_sample_gauges (one launch, host round trip included)")
        log(f"output: {m} patches -> {nbytes / 1e6:.1f} MB in {wo:.3f} s ({nbytes / wo / 1e6:.0f} MB/s); "
            f"{ng} gauges in {1e3 * wg:.3f} ms")

    # ---- ensemble serving sample (continuous batching of flood scenarios) ----
    ensemble = None
    if args.ensemble_requests > 0:
        from geoflood_amd.pool import Pool as _P
        from geoflood_amd.serving import Engine, make_requests, summarize
        geoflood_amd.set_params(**dict(PARAMS, manning_n=0.035))
        reqs = make_requests(args.ensemble_requests, args.ensemble_rate, seed=0)
        eng = Engine(args.ensemble_capacity, substeps=16)
        log(f"ensemble: {len(reqs)} requests at {args.ensemble_rate}/s ...")
        res = eng.serve(reqs)
        ensemble = summarize(res, eng.wall, eng.cell_updates)
        ensemble.update(rate_per_s=args.ensemble_rate, t_end_s=1800.0, capacity_patches=args.ensemble_capacity,
                        hbm_budget_bytes=_P.bytes_for(32, 32, args.ensemble_capacity),
                        sample=f"{len(reqs)} requests of the Poisson stream (seed 0); domains 128^2-1024^2 "
                               "of 32x32 patches over 2 km, Gaussian-sum DEM (seed = request id), Manning n "
                               "U[0.02,0.06], inflow peak log-U[20,2000] m^3/s, dry at t = 0 -> 1800 s "
                               "(dt bound 10 s)")
        eng.close()
        geoflood_amd.set_params(**PARAMS)
        log(f"ensemble: p50 {ensemble['p50_latency_s']} s, p99 {ensemble['p99_latency_s']} s, "
            f"{ensemble['cell_updates_per_s'] / 1e9:.2f} G cell-updates/s")

    # ---- plain per-patch loop ----
    plain = None
    gpu_first = None
    if not args.no_plain:
        sid = reset(pool, layout, state, args.plain_steps)
        t0 = time.perf_counter()
        while pool.run_plain(args.plain_steps):
            pass
        pw = time.perf_counter() - t0
        info_p = pool.scenario(sid)
        pv = info_p["nsteps"] * cells / pw
        plain = dict(value=pv, unit="cell-updates/s", steps=info_p["nsteps"],
                     attempts=info_p["attempts"], ms_per_step=1e3 * pw / max(info_p["attempts"], 1))
        gpu_first = np.stack(pool.download_global(layout, buf=info_p["cur"]))
        log(f"plain: {info_p['nsteps']} steps in {pw:.3f} s, {pv / 1e9:.4f} G cell-updates/s")

    # ---- host-resident drop-in batch (the reference's architecture) ----
    # gf_step2_batch over all 16,384 host patches in the reference's AoS
    # layout: pack, H2D, step, D2H, unpack per call.  The host owns the ghost
    # cells in this mode (ForestClaw fills them between steps); here they stay
    # those of the initial state, which changes the values but not the work.
    hostbatch = None
    if args.host_steps > 0:
        from geoflood_amd.pool import HostBatch, global_to_aos
        h, hu, hv, b, dx = state
        q_aos, aux_aos = global_to_aos(h, hu, hv, b, layout)
        q_list = list(q_aos)
        aux_list = list(aux_aos)
        hb = HostBatch(q_list, aux_list, dx, dx)  # patch pointers, bound once
        hb.step(0.0, DT0)  # warm-up: staging and kernels
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            hb.step(0.0, DT0)
        hw = time.perf_counter() - t0
        pcie = (q_aos.nbytes + aux_aos.nbytes) + q_aos.nbytes
        hostbatch = dict(value=args.host_steps * cells / hw, unit="cell-updates/s", steps=args.host_steps,
                         ms_per_step=1e3 * hw / args.host_steps, pcie_bytes_per_step=pcie,
                         note="gf_step2_batch (host AoS patches: pack, H2D, step, D2H, unpack per call); "
                              "PCIe-inclusive, not comparable to the device-resident value")
        log(f"host batch: {args.host_steps} steps in {hw:.3f} s, {hostbatch['value'] / 1e9:.3f} G cell-updates/s "
            f"({pcie / 1e6:.0f} MB over PCIe per step)")
        del q_aos, aux_aos, q_list, aux_list, hb

    # ---- regrid between levels on resident patches ----
    # F coarse patches of the workload (ghost cells as stored) refined into
    # 4F children and averaged back, in one call each
    regrid = None
    if args.regrid_families > 0:
        from geoflood_amd.pool import Pool as _RP, global_to_aos
        F = min(args.regrid_families, layout.npatch)
        h, hu, hv, b, dx = state
        q_aos, aux_aos = global_to_aos(h, hu, hv, b, layout)
        ox, oy = layout.origins()
        rp = _RP(MX, MX, 5 * F)
        try:
            cx = np.concatenate([ox[:F], np.repeat(ox[:F], 4) + np.tile([0, 0.5 * MX, 0, 0.5 * MX], F) * dx])
            cy = np.concatenate([oy[:F], np.repeat(oy[:F], 4) + np.tile([0, 0, 0.5 * MX, 0.5 * MX], F) * dx])
            cd = np.concatenate([np.full(F, dx), np.full(4 * F, 0.5 * dx)])
            rp.set_patches(cd, cd, cx, cy)
            rp.upload_aos(0, q_aos[:F], aux_aos[:F])
            rp.upload_aos(F, np.zeros((4 * F,) + q_aos.shape[1:]), np.repeat(aux_aos[:F], 4, axis=0))
            cs, fs = np.arange(F), F + np.arange(4 * F)
            rp.interpolate2fine(cs, fs)
            rp.average2coarse(cs, fs)
            reps = 5
            t0r = time.perf_counter()
            for _ in range(reps):
                rp.interpolate2fine(cs, fs)
            ti = (time.perf_counter() - t0r) / reps
            t0r = time.perf_counter()
            for _ in range(reps):
                rp.average2coarse(cs, fs)
            ta = (time.perf_counter() - t0r) / reps
            # coarse/fine ghost transfers: every coarse patch linked across its
            # 8 face halves and 4 corners to the children slots (12 links each)
            j = np.tile(np.arange(12), F)
            lc = np.repeat(cs, 12)
            lf = F + 4 * lc + j % 4
            ls = np.where(j < 8, j // 2, 4 + (j - 8))
            lg = np.where(j < 8, j % 2, 0)
            rp.average_ghost(lc, lf, ls, lg)
            rp.interpolate_ghost(lc, lf, ls, lg)
            t0r = time.perf_counter()
            for _ in range(reps):
                rp.average_ghost(lc, lf, ls, lg)
            tag = (time.perf_counter() - t0r) / reps
            t0r = time.perf_counter()
            for _ in range(reps):
                rp.interpolate_ghost(lc, lf, ls, lg)
            tig = (time.perf_counter() - t0r) / reps
        finally:
            rp.close()
        del q_aos, aux_aos
        regrid = dict(families=F, interpolate2fine_ms=round(1e3 * ti, 4), average2coarse_ms=round(1e3 * ta, 4),
                      cf_ghost_links=int(len(lc)), average_ghost_ms=round(1e3 * tag, 4),
                      interpolate_ghost_ms=round(1e3 * tig, 4),
                      note="one call each over all families / links (32x32 patches), host round trip included")
        log(f"regrid: {F} families, interpolate2fine {1e3 * ti:.3f} ms, average2coarse {1e3 * ta:.3f} ms; "
            f"{len(lc)} coarse/fine ghost links, average {1e3 * tag:.3f} ms, interpolate {1e3 * tig:.3f} ms")

    # ---- AMR overland flood on resident patches ----
    amr = None
    if args.amr_steps > 0:
        amr = amr_overland(args.amr_steps, tune=not args.amr_fixed_variant)

    # ---- autotuner sweep under dynamic AMR ----
    sweep = None
    if args.sweep_steps > 0:
        sweep = autotune_sweep(args.sweep_steps)

    # ---- CPU baseline (reference Fortran step on the host cores) ----
    cpu = None
    cpu_port = None
    cpu_host = None
    parity = None
    if not args.no_cpu:
        nthreads = min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline: {args.cpu_steps} steps on {nthreads} threads ...")
        cb = cpu_baseline(state, args.cpu_steps, nthreads)
        cpu = dict(value=cb["value"], unit="cell-updates/s", cores=nthreads, kind=cb["kind"],
                   sample=f"first {args.cpu_steps} accepted steps of the full 4096^2 workload "
                          f"({args.cpu_steps * cells} cell-updates, {cb['seconds']:.1f} s)")
        if gpu_first is not None and args.cpu_steps == args.plain_steps:
            parity = state_parity(gpu_first, cb["state"])
            parity.update(steps=args.cpu_steps, bar=1e-10,
                          metric="wet cells (h > drytol in both): |dh| / max|h|, and |d(hu)|, |d(hv)| relative to "
                                 "max(|hu|, h sqrt(g h)) of the cell; dry cells: max |d(h, hu, hv)| / drytol; "
                                 "plus identical wet/dry masks")
        log(f"cpu: {cb['value'] / 1e6:.2f} M cell-updates/s ({cb['kind']}); parity {parity}")
        if cb["kind"] == "reference":
            # the C restatement on the same sample and threads, for comparison
            # (with the device's correctly rounded powers: its state must equal
            # the device's bit for bit)
            cp = cpu_baseline(state, args.cpu_steps, nthreads, port_only=True, cr_pow=True)
            cpu_port = dict(value=cp["value"], unit="cell-updates/s", cores=nthreads, kind="port",
                            sample=cpu["sample"].split(" (")[0] + f" ({cp['seconds']:.1f} s)")
            if gpu_first is not None and args.cpu_steps == args.plain_steps:
                cpu_port["device_state_identical"] = bool(np.array_equal(gpu_first, cp["state"]))
            log(f"cpu: {cp['value'] / 1e6:.2f} M cell-updates/s (port); device state identical to it: "
                f"{cpu_port.get('device_state_identical')}")
        cpu_host = dict(model=cpu_model(), nproc=os.cpu_count(), threads_used=nthreads)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "uniform 4096x4096 dam break over a seeded fractal DEM "
                        f"({(N // MX) ** 2} patches of {MX}x{MX}, mbc=2, single level)",
            "dem": "spectral slope -2, RMS 3 m, seed 0, plus 1/500 slope down +x; dx = 1 m",
            "initial_state": "h = 10 m for x < L/2, y >= 3L/4; dry elsewhere; u = v = 0",
            "numerics": "g 9.81, drytol 1e-3, Manning n 0.03, MC limiter, transverse 2, walls",
            "dt_control": f"initial dt {DT0} s, desired CFL {DESIRED_CFL}, max CFL {MAX_CFL}, on device",
            "mode": "batched: one hipGraph replay per attempt (classify, fused step with halo pull, dt control)",
            "accepted_steps": accepted,
            "warmup_accepted_steps": info_w["nsteps"],
            "simulated_time_s": [info_w["t"], info_b["t"]],
            "step_variants": [dict(tile=[v[0], v[1]], threads=v[2], patches_per_workgroup=v[3],
                                   xface_increments="registers" if v[4] else "lds") for v in variants],
            "autotune_initial": dict(variant=best, k_all_ms=[round(m, 5) for m in tune_ms]),
            "autotune_retunes": retunes,
            "active_patches_per_attempt": round(active_per_attempt, 1),
            "dry_patches_per_attempt": round(dry_per_attempt, 1),
        },
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "host_batch": hostbatch,
        "output": output,
        "regrid": regrid,
        "plain": plain,
        "roofline": roofline,
        "step_composite": composite,
        "cpu_baseline": cpu,
        "cpu_port": cpu_port,
        "cpu_host": cpu_host,
        "parity_vs_cpu": parity,
        "ensemble": ensemble,
        "amr_overland": amr,
        "autotune_sweep": sweep,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
