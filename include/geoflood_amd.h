/*
 * geoflood_amd.h -- C ABI of the MI355X (gfx950) GeoFlood patch-step library
 * (libgeoflood_amd.so).
 *
 * The library replaces the GPU route of GeoFlood's fc2d_cpu_cuda solver
 * plugin: the per-patch finite-volume step (b4step2, rpn2 + rpt2, MC-limited
 * flux2 with the conservative update, src2 Manning friction) and the CFL
 * reduction, in FP64, behind plain C entry points.  Each declaration names
 * the reference interface it replaces (paths relative to the GeoFlood tree).
 *
 * Conventions
 *   - Every entry point returns an int status (GF_OK = 0, negative on error)
 *     instead of calling exit() as the reference's CHECK macro does
 *     (src/fc2d_cpu_cuda/fc2d_cudaclaw_check.h:3-11); gf_last_error() gives
 *     the message of the last failure on the calling thread.  No C++
 *     exception leaves the library.
 *   - Host patch arrays use the reference's layout: q(meqn=3, 1-mbc:mx+mbc,
 *     1-mbc:my+mbc) and aux(maux, ...) in Fortran order, ghost cells included.
 *   - Only mbc = 2, meqn = 3, mwaves = 3, mcapa = 0 and Cartesian coordinates
 *     are supported (the configurations of the GeoFlood GPU route), with
 *     (mx, my) in {(8,8), (16,16), (32,32), (64,64)}.
 *   - One device, one stream and one set of library state per process: call
 *     the library from one host thread at a time, as the reference plugin
 *     does (one host thread per MPI rank, fc2d_cpucuda.cpp:971-976).  The
 *     library runs host threads of its own for staging copies and text
 *     formatting; they are joined by gf_finalize.
 */
#ifndef GEOFLOOD_AMD_H
#define GEOFLOOD_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GF_OK = 0,
    GF_EINVAL = -1,       /* invalid argument */
    GF_EHIP = -2,         /* HIP runtime error */
    GF_ENOMEM = -3,       /* device, pinned or host memory allocation failed */
    GF_EUNSUPPORTED = -4, /* configuration outside the supported set */
    GF_ESTATE = -5,       /* call out of order (e.g. no gf_init), or another host failure */
    GF_EIO = -6           /* an output file could not be written */
};

/* Neighbour codes for gf_pool_set_patches (values >= 0 are patch indices):
 * the bc2 side kinds mthbc = 3 (wall), 1 (extrapolation) and 0 (hydrograph
 * inflow, geoflood_bc2_fort.f90:52-102; face entries only), or ghosts that
 * the host stores. */
enum { GF_NB_WALL = -1, GF_NB_EXTRAP = -2, GF_NB_STORED = -3, GF_NB_INFLOW = -4 };

/* Physics and method parameters.  Replaces the __constant__ GeofloodVars of
 * src/fc2d_cpu_cuda/riemann_source/variables.h:20-37 and the method symbols
 * order/mthlim/use_fwaves of cuda_source/cudaclaw_flux2.cu:40-59. */
typedef struct gf_params {
    double grav;           /* geoclaw_module::grav */
    double dry_tolerance;  /* geoclaw_module::dry_tolerance */
    double manning_n;      /* manning_coefficient(1) */
    double friction_depth; /* geoclaw_module::friction_depth */
    int friction_forcing;  /* apply Manning friction in src2 */
    int src_term;          /* call src2 (fc2d_cpucuda_options::src_term) */
    int order;             /* method(2): 1 or 2 */
    int transverse;        /* method(3): 0, 1 or 2 */
    int mthlim[3];         /* limiter per wave family: 1 minmod .. 4 MC, 5 BW */
    int mcapa;             /* capacity index; must be 0 */
} gf_params;

/* Snapshot of a scenario's device-side time-step control record. */
typedef struct gf_scenario_info {
    double t;        /* simulated time of the accepted state */
    double dt;       /* dt of the next attempt */
    double maxcfl;   /* max CFL number of the last attempt */
    double last_dt;  /* dt of the last accepted step */
    int nsteps;      /* accepted steps */
    int retakes;     /* rejected attempts (maxcfl > max_cfl) */
    int attempts;    /* all attempts */
    int active;      /* 1 while stepping */
    int cur;         /* q buffer holding the accepted state */
    int status;      /* GF_SCEN_OK, or why the scenario stopped */
} gf_scenario_info;

/* gf_scenario_info::status.  GF_SCEN_NONFINITE: an attempt produced a NaN or
 * infinite state or CFL number; the scenario stopped without accepting it
 * (t, dt and the accepted buffer are those before the attempt).  The
 * reference has no such check (a NaN CFL number fails "maxcfl > max_cfl" in
 * fc2d_geoclaw_run.c:529 and the step is accepted). */
enum { GF_SCEN_OK = 0, GF_SCEN_NONFINITE = 1 };

/* Refinement criteria of the tagging entry points (refinement_data_module,
 * geoclaw_module::sea_level and the regions common block of the reference:
 * geolib_source/refinement_data_module.f90, fclaw2d_source/regions.i).
 * region[8*m .. 8*m+7] = xlow, ylow, xhi, yhi, tlow, thi, minlevel,
 * maxlevel of region m. */
typedef struct gf_refine {
    double wave_tolerance;
    double deep_depth;
    int max_level_deep;
    double max_velocity_depth_product;
    int num_speeds;                 /* >= 1 */
    const double *speed_tolerance;
    int num_flowgrades;
    const double *flowgradevalue;
    const int *iflowgradevariable;  /* 1 depth, 2 momentum, 3 surface elevation */
    const int *iflowgrademinlevel;
    double sea_level;
    int mregions;
    const double *region;
} gf_refine;

typedef struct gf_pool gf_pool;

/* ---- device and parameters ---------------------------------------------- */

/* Select the device and create the library's stream.  Replaces
 * fc2d_cudaclaw_initialize_GPUs (cuda_source/cudaclaw_initialize.cu:37-74). */
int gf_init(int device);

/* Release everything gf_init created. */
int gf_finalize(void);

/* Validate mwaves against the supported solver.  Replaces
 * cudaclaw_check_parameters (fc2d_cpucuda_options.h:93,
 * cuda_source/cudaclaw_flux2.cu:48-51). */
int gf_check_parameters(int mwaves);

/* Set physics/method parameters for subsequent steps.  Replaces setprob_cuda
 * (riemann_source/variables.cu:72-113) and cudaclaw_set_method_parameters
 * (fc2d_cpucuda_options.h:94, cuda_source/cudaclaw_flux2.cu:53-59). */
int gf_set_params(const gf_params *params);

/* Read back the parameters.  Replaces cudaclaw_get_method_parameters
 * (fc2d_cudaclaw_cuda.h:130). */
int gf_get_params(gf_params *params);

/* Message of the last error on this thread ("" if none). */
const char *gf_last_error(void);

/* Number of bytes of HBM a pool of `capacity` patches of mx x my occupies
 * (both q buffers, aux, descriptors, per-patch scratch). */
long long gf_pool_bytes(int mx, int my, int mbc, int capacity);

/* ---- drop-in batch step (reference architecture) ------------------------ */

/* Advance n host-resident patches by one step: pack AoS -> SoA, copy to the
 * device, run the fused step, reduce the CFL number, copy back and unpack
 * (interior cells are updated; ghost cells are left as given).  q[k], aux[k]
 * point to patch k's arrays in the reference layout with ghost cells filled
 * by the caller; dx[k], dy[k] are its cell sizes.  *maxcfl receives the
 * maximum over the patches of the per-patch CFL numbers, each capped to 0
 * above 1000 as geoclaw_update does (fc2d_cpucuda.cpp:392).  Returns
 * GF_EINVAL for n <= 0, NULL arrays, maux < 1, dt <= 0 or a cell size <= 0,
 * and GF_EUNSUPPORTED for mbc != 2 or a patch shape other than 8x8, 16x16,
 * 32x32 or 64x64, before any device work.  Replaces cudaclaw_step2_batch
 * (cuda_source/cudaclaw_step2.cu:57-324). */
int gf_step2_batch(int mx, int my, int mbc, int maux, int n,
                   double *const *q, double *const *aux,
                   const double *dx, const double *dy,
                   double t, double dt, double *maxcfl);

/* gf_step2_batch for callers that cannot allocate memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * has no counterpart in the reference.  Same arguments and results as
 * gf_step2_batch, but it calls no hipMalloc, hipFree or other allocating
 * routine, on the device or the host: its device memory is a fixed 32 MiB
 * array per GPU (loaded with the library and held until the process ends),
 * and the host patches are copied straight between the caller's arrays and
 * that array.  A batch needs about n * (10 + maux) * (mx+4) * (my+4) * 8
 * bytes of it (up to about 290 patches of 32x32 with maux = 1); a larger
 * batch returns GF_EINVAL.  Every call on one GPU uses the same array and leaves its last
 * contents there: calls that overlap on one GPU (from several host threads)
 * race. */
int gf_step2_batch_noalloc(int mx, int my, int mbc, int maux, int n,
                           double *const *q, double *const *aux,
                           const double *dx, const double *dy,
                           double t, double dt, double *maxcfl);

/* ---- plugin slot: per-patch callback with batching ---------------------- */

typedef struct gf_plugin gf_plugin;

/* Batching state of the single_step_update slot for patches of mx x my with
 * maux aux fields, flushing every buffer_len patches.  Replaces the
 * flux_array allocated by cudaclaw_update at iter == 0 (fc2d_cpucuda.cpp:
 * 448-458) and the buffer-len option of the plugin (fc2d_cpucuda_options.c).
 * Returns NULL on invalid arguments. */
gf_plugin *gf_plugin_create(int mx, int my, int mbc, int maux, int buffer_len);

/* Free the batching state. */
int gf_plugin_destroy(gf_plugin *plugin);

/* The single_step_update callback for patch number iter of total_count in
 * the current sweep (ForestClaw's buffer_data): queues the host patch (q,
 * aux with ghost cells filled, cell sizes dx, dy) and, when buffer_len
 * patches are queued or iter is the last, advances the queued patches with
 * gf_step2_batch.  *maxcfl is the batch's maximum per-patch CFL number
 * (capped per patch) on flushing calls and 0 otherwise, so the caller's
 * running maximum over all calls is the level's CFL number.  Replaces
 * cudaclaw_update (fc2d_cpucuda.cpp:409-495) with cudaclaw_store_buffer
 * (cuda_source/cudaclaw_store_patches.cu:13-42). */
int gf_plugin_single_step_update(gf_plugin *plugin, double *q, double *aux, double dx,
                                 double dy, int iter, int total_count, double t,
                                 double dt, double *maxcfl);

/* ---- device-resident patch pool ----------------------------------------- */

/* Allocate a pool of `capacity` patches: two SoA q buffers with ghost cells,
 * aux, descriptors, per-patch CFL/depth and scenario control records.
 * Replaces fc2d_cudaclaw_allocate_buffers (cuda_source/cudaclaw_allocate.cu:
 * 118-138) and the per-patch cudaclaw_allocate_fluxes (:21-85), which
 * allocated 12 scratch arrays per patch; the fused kernel needs none. */
gf_pool *gf_pool_create(int mx, int my, int mbc, int capacity);

/* Free a pool.  Replaces fc2d_cudaclaw_deallocate_buffers and
 * cudaclaw_deallocate_fluxes (cuda_source/cudaclaw_allocate.cu:88-145). */
int gf_pool_destroy(gf_pool *pool);

/* Patch shapes: mx = my in {8, 16, 32, 64}, mbc = 2. */

/* Describe patches [0, n) and forget every scenario: cell sizes, scenario
 * entries (NULL or all -1: membership is assigned by gf_pool_add_scenario;
 * patches of no active scenario are skipped by gf_pool_run), and 8 neighbour
 * entries per patch (left, right, bottom, top, left-bottom, right-bottom,
 * left-top, right-top), each a patch index or GF_NB_*.  This is the
 * information ForestClaw's ghost exchange uses (fc2d_geoclaw_copy_fort.f:
 * 11-107) together with the per-side mthbc of geoclaw_bc2
 * (fc2d_cpucuda.cpp:131-186). */
int gf_pool_set_patches(gf_pool *pool, int n, const double *dx, const double *dy,
                        const double *xlower, const double *ylower, const int *scenario,
                        const int *nbr);

/* Describe pool slots [first, first+n) without touching the others: the
 * admission path of a serving engine, which places a new scenario's forest in
 * free slots while other scenarios keep stepping.  nbr holds 8 entries per
 * patch as in gf_pool_set_patches (indices are pool slots); xlower/ylower
 * (nullable) are the patches' lower-left corners, used by hydrograph sides
 * (cudaclaw_patch_data_t's xlower/ylower, cudaclaw_allocate.h:18-64).  The
 * slots must belong to no scenario and stay so until gf_pool_add_scenario.
 * The slots belong
 * to no scenario until gf_pool_add_scenario.  Same reference interface as
 * gf_pool_set_patches, per forest instead of per pool. */
int gf_pool_describe(gf_pool *pool, int first, int n, const double *dx, const double *dy,
                     const double *xlower, const double *ylower, const int *nbr);

/* Copy n patches from host AoS arrays (contiguous, one patch after another,
 * reference layout with ghosts) into pool slots [first, first+n), q into
 * buffer `buf` (0 or 1).  aux_aos may be NULL to keep the pool's aux.
 * Replaces CUDACLAW_SWAP_DATA with geoclaw2cudaclaw = 1 (cuda_source/
 * data_swap.f90:22-40) plus the H2D copy (cudaclaw_step2.cu:158). */
int gf_pool_upload(gf_pool *pool, int first, int n, int buf, const double *q_aos,
                   const double *aux_aos, int maux);

/* Copy pool slots [first, first+n) of buffer `buf` back to host AoS arrays.
 * Replaces the D2H copy and CUDACLAW_SWAP_DATA with geoclaw2cudaclaw = 0
 * (cudaclaw_step2.cu:298-315; data_swap.f90:41-53). */
int gf_pool_download(gf_pool *pool, int first, int n, int buf, double *q_aos);

/* Copy the aux (bathymetry, maux = 1, ghosts included) of pool slots
 * [first, first+n) back to host AoS arrays: bc2's wall branches of a
 * hydrograph side rewrite ghost aux on the device (geoflood_bc2_fort.f90:
 * 82-90), so the device copy is the current one.  The reference keeps aux on
 * the host (CUDACLAW_SWAP_DATA copies it forward only, data_swap.f90:22-40). */
int gf_pool_download_aux(gf_pool *pool, int first, int n, double *aux_aos);

/* Fill the ghost cells of buffer `buf` on the device for patches [first,
 * first+n): same-level face and corner copies, then solid-wall /
 * extrapolation / hydrograph sides in the order left, right, bottom, top
 * (geoflood_bc2_fort.f90:7-499), the hydrograph at time t.  GF_NB_STORED
 * sides keep the values already stored (e.g. coarse/fine ghosts filled by
 * the host).  Hydrograph sides are filled for patches of a scenario only. */
int gf_pool_ghost_fill(gf_pool *pool, int first, int n, int buf, double t);

/* One step attempt of patches [first, first+n) from buffer `src` to buffer
 * `dst` with time step dt, all patches in one launch.  dst == src steps in
 * place, which only a step variant with one tile per patch can do (its one
 * workgroup holds the whole patch before writing); with a tiled variant
 * (gf_pool_variants) it returns GF_EUNSUPPORTED.  cfl (optional, host,
 * length n) receives the per-patch capped CFL numbers (NaN for a patch whose
 * new state is not finite); *maxcfl (optional) their maximum, NaN if any is.
 * Ghost cells must already be filled. */
int gf_pool_step(gf_pool *pool, int first, int n, int src, int dst, double dt,
                 double *cfl, double *maxcfl);

/* The same step issued as one kernel launch per patch from a host loop (the
 * "plain" per-patch step). */
int gf_pool_step_plain(gf_pool *pool, int first, int n, int src, int dst, double dt,
                       double *cfl, double *maxcfl);

/* Register a scenario over pool slots [first_patch, first_patch+npatch) with
 * its own device-side time-step control (fc2d_geoclaw_run.c:417-590): start
 * time, initial dt, desired/max CFL, and the stop condition (target_steps
 * accepted steps and/or simulated time t_end; <= 0 disables each).  The
 * accepted state starts in q buffer 0.  Returns the scenario id (>= 0). */
int gf_pool_add_scenario(gf_pool *pool, int first_patch, int npatch, double t0,
                         double dt0, double desired_cfl, double max_cfl,
                         int target_steps, double t_end);

/* Deactivate every scenario and forget them (slots keep their data). */
int gf_pool_clear_scenarios(gf_pool *pool);

/* Forget one scenario: its slots return to no scenario and its id may be
 * reused by the next gf_pool_add_scenario.  Other scenarios are unaffected.
 * Ends one outstyle_3 run (fc2d_geoclaw_run.c:417-590) of an ensemble. */
int gf_pool_remove_scenario(gf_pool *pool, int scenario);

/* Hydrograph inflow of a scenario's GF_NB_INFLOW sides: discharge per unit
 * width hu[k] at time[k] (nrows rows, linear in between, clamped outside),
 * channel strip of width b centred at y0 (left/right sides) or x0
 * (bottom/top).  Inside the strip the ghost state is the reference's
 * inflow state (critical depth for a dry boundary cell, Riemann-invariant /
 * two-shock depth for a wet one, wall for zero discharge); outside it is a
 * wall.  A scenario may use one inflow side.  Replaces read_hydrograph_data
 * and the mthbc = 0 branch of fc2d_geoclaw_bc2 (hydrograph/
 * hydrograh_module.f90:35-331, fortran_source/geoflood_bc2_fort.f90:52-102
 * and the right/bottom/top twins) with hydrograph_type 'discharge'. */
int gf_pool_set_scenario_inflow(gf_pool *pool, int scenario, int nrows, const double *time,
                                const double *hu, double b, double x0, double y0);

/* Per-scenario Manning coefficient (the scenario starts with gf_params'
 * manning_n).  Replaces manning_coefficient(1) of friction_module
 * (fortran_source/geoclaw_src2_fort.f90:53) for one member of an ensemble. */
int gf_pool_set_scenario_friction(gf_pool *pool, int scenario, double manning_n);

/* New stop condition of a scenario: target_steps accepted steps and/or
 * simulated time t_end (<= 0 disables each), as in gf_pool_add_scenario; the
 * scenario is active again when it has not reached them (and has not
 * stopped on a non-finite state).  A driver that regrids every few steps
 * (fc2d_geoclaw_run.c:569-575) moves the target to the next regrid and
 * replays attempts until the scenario stops there. */
int gf_pool_set_scenario_stop(gf_pool *pool, int scenario, int target_steps, double t_end);

/* Upper bound of a scenario's dt (0, the default: none).  After an accepted
 * attempt with CFL number m the next dt is dt * desired_cfl / m
 * (fc2d_geoclaw_run.c:577-580), capped at dt_max; with m = 0 (no wave: every
 * cell dry, e.g. a dry start with a hydrograph at zero discharge) it is
 * dt_max, or the same dt when no bound is set, where the reference's rule
 * gives an infinite dt.  The reference's outstyle_3 has no bound. */
int gf_pool_set_scenario_dt_max(gf_pool *pool, int scenario, double dt_max);

/* Stop or restart a scenario without removing it. */
int gf_pool_set_scenario_active(gf_pool *pool, int scenario, int active);

/* ---- refinement tagging on resident patches ------------------------------ */

/* tags[k] = 1 when any interior cell of pool slot patches[k] (buffer buf,
 * time t, refinement level `level`) flags for refinement.  Replaces
 * fc2d_geoclaw_fort_tag4refinement + fc2d_geoclaw_flag2refine
 * (fclaw2d_source/fc2d_geoclaw_tag4refinement_fort.f90:1-59,
 * fc2d_geoflood_flag2refine.f90:7-130) behind the plugin's tag4refinement
 * hook (fc2d_cpucuda.cpp:512-560), without moving the patch to the host. */
int gf_pool_tag_refine(gf_pool *pool, int buf, double t, const gf_refine *crit, int level,
                       int maxlevel, int init_flag, const int *patches, int n, int *tags);

/* tags[k] = 1 when the four siblings families[4k .. 4k+3] may coarsen (no
 * cell of any of them flags under the coarsening thresholds).  Replaces
 * fc2d_geoclaw_fort_tag4coarsening (fc2d_geoclaw_tag4coarsening_fort.f90:
 * 1-110) behind the tag4coarsening hook (fc2d_cpucuda.cpp:562-614). */
int gf_pool_tag_coarsen(gf_pool *pool, int buf, double t, const gf_refine *crit, int level,
                        int maxlevel, int init_flag, const int *families, int n, int *tags);

/* ---- step variants and the autotuner ------------------------------------ */

/* Step variants of the pool's patch shape: *count variants, desc[5k..5k+4]
 * = tile width, tile height, threads per tile, patches per workgroup of the
 * batched step kernel, and where a face's flux increments wait between the
 * x and y sweeps (1 the owning thread's registers, 0 LDS: the LDS split) of
 * variant k.  A tile smaller than the patch splits each patch into tiles with
 * their own 2-cell halo; several patches per workgroup step side by side,
 * one per group of threads.  The results do not depend on the variant.
 * *current is the variant in use.  Any pointer may be NULL.  The reference
 * fixes one shape (cudaclaw_flux2.cu:68-120: one block of 128 threads per
 * patch). */
int gf_pool_variants(gf_pool *pool, int *current, int *count, int *desc);

/* Select a step variant (the next gf_pool_run recaptures its graph). */
int gf_pool_set_variant(gf_pool *pool, int variant);

/* Time every variant's persistent step kernel `reps` times on the active
 * patches of the pool's current states (no scenario advances) and select
 * the fastest; *best (nullable) receives it, ms[k] (nullable, one entry per
 * variant) the average time of variant k in milliseconds.  Re-run when the
 * active population changes (after regrid, or as a flood front grows).
 * Only active scenarios' patches are stepped: with none listed (every
 * scenario stopped, or all its patches steady and dry) it returns GF_ESTATE
 * and keeps the variant. */
int gf_pool_autotune(gf_pool *pool, int reps, int *best, double *ms);

/* Copy a scenario's control record to the host. */
int gf_pool_get_scenario(gf_pool *pool, int scenario, gf_scenario_info *info);

/* Control records of scenario ids [first, first+n) in one device-to-host
 * copy (ids with no scenario report active = 0): the poll of a serving loop
 * that runs many scenarios (the per-step host read of maxcfl in
 * fc2d_geoclaw_run.c:504-515, for an ensemble). */
int gf_pool_get_scenarios(gf_pool *pool, int first, int n, gf_scenario_info *info);

/* Advance all active scenarios by up to `substeps` attempts: each attempt is
 * ghost fill -> fused step of every patch -> per-scenario CFL reduction and
 * accept/retake decision on the device.  With use_graph != 0 the attempt is
 * captured once in a hipGraph and replayed.  Returns after the device has
 * finished; *any_active (optional) tells whether a scenario is still active. */
int gf_pool_run(gf_pool *pool, int substeps, int use_graph, int *any_active);

/* Advance scenario s to target_steps accepted steps: set its stop there (as
 * gf_pool_set_scenario_stop with t_end off), run one attempt per missing
 * step, then single attempts until it stops (each retake costs one), and
 * copy its control record to *info (optional).  Every other active scenario
 * of the pool advances with it.  max_attempts > 0 bounds the attempts
 * (GF_ESTATE beyond it); the scenario stops early on a non-finite state
 * (info->status).  One call per regrid interval of the outstyle_3 loop
 * (fc2d_geoclaw_run.c:505-575: steps and retakes up to the next regrid),
 * in place of gf_pool_set_scenario_stop, gf_pool_run and
 * gf_pool_get_scenario. */
int gf_pool_advance(gf_pool *pool, int scenario, int target_steps, int use_graph, int max_attempts,
                    gf_scenario_info *info);

/* The plain loop for the same scenarios: per attempt, a host-issued launch
 * per patch, a host read of the CFL numbers and the dt rule on the host.
 * Semantics equal gf_pool_run. */
int gf_pool_run_plain(gf_pool *pool, int substeps, int *any_active);

/* The step's two powers evaluated on the device, for n host values x:
 * which = 0 gives x^y with y = 7.0/3.0 (src2's h**(7.d0/3.d0),
 * geoclaw_src2_fort.f90:62), which = 1 with y = 2.0/3.0 (the hydrograph's
 * (q0(2)/sqrt(grav))**(2.0d0/3.0d0), geoflood_bc2_fort.f90:63).  Both are
 * correctly rounded on the device; this entry point has no counterpart in
 * the reference and exists to check the device arithmetic against the host
 * libm's pow. */
int gf_eval_powers(int which, int n, const double *x, double *out);

/* Wait for all work queued on the library stream. */
int gf_synchronize(void);

/* Device time (ms) of the last gf_pool_run / gf_pool_run_plain /
 * gf_pool_step* call and of its step launches alone (for gf_pool_run: the
 * persistent step kernel of each attempt, summed), from HIP events on the
 * library stream; step_launches counts the timed step launches.  Graph runs
 * report the total only. */
int gf_last_timing(double *total_ms, double *step_ms, int *step_launches);

/* Patches stepped by the full kernel, patches handled as dry, attempts, and
 * of the dry patches those left untouched as steady (every interior depth
 * below dry_tolerance for two steps in a row, so both state buffers already
 * hold the step's result), accumulated over all gf_pool_run calls on this
 * pool. */
int gf_pool_stats(gf_pool *pool, long long *active_patches, long long *dry_patches,
                  long long *attempts, long long *steady_patches);

/* ---- Output: ASCII frames and gauges --------------------------------------
 * Replaces fc2d_geoclaw_output_ascii (fc2d_geoclaw_output_ascii.c:39-114)
 * with its Fortran writers fc2d_geoclaw_fort_write_header / _write_file
 * (fc2d_geoclaw_output_ascii_fort.f:1-117), for device-resident patches.
 *
 * Appends patches [first, first+n) of buffer buf to dir/fort.qNNNN (NNNN =
 * frame mod 10000) in the reference's byte format; with write_header, first
 * writes dir/fort.tNNNN (time, meqn+1 = 4, ngrids, maux = 1, 2, mbc) and
 * empties fort.qNNNN, as the reference's header routine does.  Per patch:
 * patch_num[k] (NULL: first + k), level[k] and blockno[k] (NULL: 0), mpirank,
 * mx, my, xlower, ylower, dx, dy (from gf_pool_set_patches /
 * gf_pool_describe), then h, hu, hv, eta per interior cell.  As the reference
 * does to the patch it writes, values with |q| < 1e-99 are set to 0 in the
 * pool's state.  GF_EIO when a file cannot be written. */
int gf_pool_write_frame(gf_pool *pool, int buf, int first, int n, const char *dir, int frame,
                        double time, int ngrids, int write_header, const int *patch_num,
                        const int *level, const int *blockno, int mpirank);

/* Gauge samples of buffer buf: gauge k at (xc[k], yc[k]) inside patch slot[k]
 * (xlower <= xc <= xlower + mx*dx, likewise y, as the reference asserts,
 * fc2d_geoclaw_gauges_default.c:250-251); out[4k..4k+3] = h, hu, hv, b by
 * fc2d_geoclaw_update_gauge (geoclaw_gauges_fort.f90:1-77): bilinear in the
 * four cells around the gauge, or the containing cell when one of their
 * depths is below dry_tolerance/10.  A cell beyond the last interior row or
 * column takes the value a ghost fill would give it (the neighbour's
 * interior, wall mirror or extrapolation; the stored ghost for host-filled
 * sides), where the reference reads the ghost cells its step left. */
int gf_pool_sample_gauges(gf_pool *pool, int buf, int n, const int *slot, const double *xc,
                          const double *yc, double *out);

/* ---- Regrid and time interpolation between levels --------------------------
 * Refinement ratio 2 between slots of one pool (same patch size): family f
 * is the coarse patch in slot coarse[f] and its four children in slots
 * fine[4f + g], g = igrid = 0..3 (x fastest), all in buffer buf.  Slots must
 * be described (gf_pool_set_patches / gf_pool_describe, the children with
 * their own dx, dy, corners and bathymetry uploaded) and distinct.  Both
 * recompute the written patches' depth maxima for the batched path.
 *
 * gf_pool_average2coarse replaces fc2d_geoclaw_fort_average2coarse
 * (fc2d_geoclaw_average_fort.f:482-603, mcapa = 0): each child averaged into
 * its quadrant of the coarse interior (wet-cell mean of eta, mean depth
 * capped by it, momentum rescaled), as ForestClaw's coarsening calls it.
 *
 * gf_pool_interpolate2fine replaces fc2d_geoclaw_fort_interpolate2fine
 * (fc2d_geoclaw_interpolate_fort.f:405-579): each child's interior (and its
 * ghost layers on the family's inner sides) from limited linear
 * reconstructions of eta (dry cells at sea level 0) and of hu, hv, with
 * ForestClaw's fclaw2d_clawpatch_compute_slopes at mth = 5.  Coarse cells
 * beyond the coarse interior are read as the halo pull reads them; a child
 * may therefore not be a same-level neighbour of a coarse patch of the call. */
int gf_pool_average2coarse(gf_pool *pool, int buf, int n, const int *coarse, const int *fine);
int gf_pool_interpolate2fine(gf_pool *pool, int buf, int n, const int *coarse, const int *fine);

/* ---- Coarse/fine ghost transfers --------------------------------------------
 * Ghost cells across a level change, between slots of one pool (buffer buf,
 * square patches, refinement ratio 2).  Link l joins the coarse patch in slot
 * coarse[l] and a half-size neighbour of the same orientation in slot
 * fine[l] across side[l] of the coarse patch: 0..3 = face 0..3 (left, right,
 * bottom, top) with the neighbour in half igrid[l] along it (0: the lower
 * coordinate); 4..7 = corner 0..3 (lower-left, lower-right, upper-left,
 * upper-right), igrid[l] not read.  The index map is that of neighbours
 * within one block or across brick blocks without rotation, in place of
 * ForestClaw's transform_face_half / transform_corner_half.  Each written
 * region (a coarse patch's face half or corner for averaging, a fine
 * patch's for interpolation) may appear once per call.  Both reset the
 * written patches' steady-dry state.  The sides that these fill are
 * GF_NB_STORED in the patch descriptors.  A run calls gf_pool_average_ghost,
 * then gf_pool_interpolate_ghost (level by level from coarse to fine), then
 * steps: same-level ghosts and physical sides are read from the neighbours
 * when a step loads its tiles (or written by gf_pool_ghost_fill, called after
 * the transfers), and a ghost corner between a physical side and a stored
 * side takes bc2's mirror of the stored strip, as in ForestClaw's ghost
 * update, which applies the physical boundaries after the exchanges.
 *
 * gf_pool_average_ghost replaces fc2d_geoclaw_fort_average_face and
 * fc2d_geoclaw_fort_average_corner (fc2d_geoclaw_average_fort.f:33-480,
 * mcapa = 0), which the plugin's average_face / average_corner hooks call
 * per (coarse, fine) pair (fc2d_cpucuda.cpp:698-745, :779-823): the coarse ghost cells
 * that the fine neighbour covers, from its interior.
 *
 * gf_pool_interpolate_ghost replaces fc2d_geoclaw_fort_interpolate_face and
 * fc2d_geoclaw_fort_interpolate_corner (fc2d_geoclaw_interpolate_fort.f:
 * 24-403; hooks fc2d_cpucuda.cpp:746-778, :824-857): the fine neighbour's ghost cells
 * that the coarse boundary cells cover (a face also fills the 2x2 ghost
 * corner away from the neighbour's sibling), from limited reconstructions of
 * eta and hu, hv as in gf_pool_interpolate2fine, the depth clipped at 0.
 * Coarse cells beyond the coarse interior are read as the halo pull reads
 * them (the stored ghosts on GF_NB_STORED sides); no slot may be both a
 * coarse and a fine patch of one call.  At a fine substep between coarse
 * steps the hooks read the coarse patch's time-interpolated copy
 * (fclaw2d_clawpatch_timesync_data): pass the gf_pool_timeinterp
 * destination slot, described like the coarse patch, as the coarse slot. */
int gf_pool_average_ghost(gf_pool *pool, int buf, int n, const int *coarse, const int *fine, const int *side,
                          const int *igrid);
int gf_pool_interpolate_ghost(gf_pool *pool, int buf, int n, const int *coarse, const int *fine,
                              const int *side, const int *igrid);

/* Copy the state of slots[0..n) (q, ghost cells included) from buffer src_buf
 * to buffer dst_buf.  Replaces fclaw2d_clawpatch_save_current_step, ForestClaw's
 * copy of a patch's state into its saved copy that the plugin's update makes
 * before each step (fc2d_cpucuda.cpp:306, :443), for resident patches: a
 * subcycled run steps the coarse and the fine level into different buffers
 * and brings them into one buffer at a synchronisation time with this copy,
 * and the saved copy is what gf_pool_timeinterp reads as q_last. */
int gf_pool_copy(gf_pool *pool, int n, const int *slots, int src_buf, int dst_buf);

/* Time interpolation for coarse ghost sources, replaces
 * fc2d_geoclaw_fort_timeinterp (fc2d_geoclaw_timeinterp_fort.f:1-69): for
 * k < n, the two interior layers next to each face of slot dst[k] of buffer
 * dst_buf become q_last + alpha (q_curr - q_last) of slot src[k] (buffers
 * buf_last, buf_curr). */
int gf_pool_timeinterp(gf_pool *pool, int n, const int *src, int buf_last, int buf_curr, double alpha,
                       const int *dst, int dst_buf);

/* ---- Resident AMR: levels in one graph per attempt --------------------------
 * ForestClaw keeps a forest of leaves of several levels and, with subcycling
 * off (every GeoFlood example: subcycle = F in fclaw_options.ini), advances
 * all of them with one dt per step after a ghost update across levels
 * (fc2d_geoclaw_run.c:505, fclaw2d_advance_all_levels); its regrid hooks
 * refine and coarsen leaves between steps (fc2d_cpucuda.cpp:513-694,
 * fc2d_geoclaw_run.c:569-575).  For a pool whose slots hold the leaves of
 * such a forest, the entry points below keep the level structure on the
 * device, so that each gf_pool_run attempt -- one hipGraph launch -- is the
 * whole ghost update (fine -> coarse averaging over all links, then level by
 * level from coarse to fine: the level's hydrograph sides and the
 * interpolation to the next finer level) followed by one persistent step of
 * every leaf of every level and the dt control.  The links and level ranges
 * live in device memory, so a regrid changes them without a new graph. */

/* Refinement level (0 .. 15) of slots [first, first+n); slots never set are
 * level 0.  Orders the hydrograph sides of the ghost update by level.
 * Replaces the level a ForestClaw patch carries (fclaw2d_patch_t::level). */
int gf_pool_set_levels(gf_pool *pool, int first, int n, const int *level);

/* The coarse/fine links of the forest (as for gf_pool_average_ghost: coarse
 * slot, fine slot one level finer, side 0..7, igrid), replacing any stored
 * before; n = 0 clears them.  Every gf_pool_run attempt and
 * gf_pool_ghost_update then performs the transfers of fc2d_cpucuda.cpp:
 * 698-857 (average_face/corner, interpolate_face/corner) over them. */
int gf_pool_set_cf_links(gf_pool *pool, int n, const int *coarse, const int *fine, const int *side,
                         const int *igrid);

/* The multi-level ghost update of the stored links on buffer buf at time t
 * (buf = -1: each scenario's current buffer and time), as each gf_pool_run
 * attempt begins with it; for a regrid or an output between steps.  Queued
 * on the library's stream: it returns before the device has finished, and
 * every later call that reads or writes the pool runs after it.  Replaces
 * fclaw2d_ghost_update with time_interp off for resident patches. */
int gf_pool_ghost_update(gf_pool *pool, int buf, double t);

/* Copy the state (buffer buf, ghost cells included) and aux of slot src[k]
 * to slot dst[k]: destinations distinct and no destination a source.  A
 * regrid keeps unchanged leaves this way when it lays the new forest out in
 * other slots.  Replaces ForestClaw's copy of a patch that survives a regrid
 * (fclaw2d_patch_copy via the plugin's copy hook). */
int gf_pool_move(gf_pool *pool, int n, const int *src, const int *dst, int buf);

/* Re-target scenario s to slots [first, first+n) (described; free or its
 * own), keeping its time, dt, buffer index, counters, friction and
 * hydrograph: the slots it held before return to no scenario.  The new
 * forest's state must already be in the scenario's current buffer there.
 * Replaces ForestClaw's switch to the new domain after a regrid
 * (fclaw2d_regrid_new_domain_setup) for a resident scenario. */
int gf_pool_set_scenario_slots(gf_pool *pool, int scenario, int first, int n);

/* tags[k] = 1 when an interior cell of slot patches[k] (buffer buf) is wet
 * (h > dry_tolerance) or has |grad eta| > eta_tol, eta = h + b, by centred
 * differences over the cell's four neighbours (read as the step reads its
 * halo: run gf_pool_ghost_update first on stored and hydrograph sides).  The
 * overland-flood runs' refinement criterion; it has no counterpart in the
 * reference's flag2refine (fc2d_geoflood_flag2refine.f90:7-130), whose
 * criteria gf_pool_tag_refine restates. */
int gf_pool_tag_eta_gradient(gf_pool *pool, int buf, double eta_tol, const int *patches, int n, int *tags);

/* Gauge files in the reference's format (geoclaw_create_gauge_files_default
 * and geoclaw_print_gauges_default, fc2d_geoclaw_gauges_default.c:172-200,
 * :275-305): gf_gauge_file_create writes the two header lines of
 * dir/gaugeNNNNN.txt (truncating it); gf_gauge_file_append appends n records
 * "level t h hu hv eta" (eta = h + b, 0 below 1e-99) from rec[4k..4k+3] =
 * h, hu, hv, b as gf_pool_sample_gauges returns them. */
int gf_gauge_file_create(const char *dir, int id, double xc, double yc);
int gf_gauge_file_append(const char *dir, int id, int n, const int *level, const double *t,
                         const double *rec);

#ifdef __cplusplus
}
#endif
#endif /* GEOFLOOD_AMD_H */
