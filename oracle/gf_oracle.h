/*
 * gf_oracle.h -- CPU oracle for the GeoFlood per-patch finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * GeoClaw Fortran step (the CPU route of the fc2d_cpu_cuda plugin).  It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and only to check or time the reference algorithm: the product
 * library (geoflood_amd) never links, loads or calls it.
 *
 * Array layouts are the reference's host layouts (Fortran column-major):
 *   q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc)   -> q[((j+mbc-1)*(mx+2mbc) + (i+mbc-1))*meqn + (m-1)]
 *   aux(maux, ...)                        -> same with maux
 *   1-D slices q1d(meqn, 1-mbc:n+mbc), fwave(meqn, mwaves, 1-mbc:n+mbc),
 *   s(mwaves, 1-mbc:n+mbc), gadd(meqn, 1-mbc:n+mbc, 2).
 * meqn = 3, mwaves = 3 (shallow water with topography); aux(1,.,.) is the
 * bathymetry b.
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfo_params {
    double grav;            /* geoclaw_module::grav */
    double dry_tolerance;   /* geoclaw_module::dry_tolerance */
    double manning_n;       /* manning_coefficient(1) (num_manning = 1, manning_break = huge) */
    double friction_depth;  /* geoclaw_module::friction_depth */
    int friction_forcing;   /* geoclaw_module::friction_forcing (src2 Manning on/off) */
    int src_term;           /* fc2d_cpucuda_options::src_term (call src2 at all) */
    int order;              /* method(2): 1 = Godunov, 2 = + limited corrections */
    int transverse;         /* method(3): 0 none, 1 increment waves, 2 + correction waves */
    int mthlim[3];          /* limiter per wave family (4 = MC) */
} gfo_params;

/* riemanntype (geoclaw_riemann_utils.f:566-677).  out = {hm, s1m, s2m}, rare = {rare1, rare2} */
void gfo_riemanntype(double hL, double hR, double uL, double uR, int maxiter,
                     double drytol, double g, double out[3], int rare[2]);

/* riemann_aug_JCP (geoclaw_riemann_utils.f:2-311) with meqn = mwaves = 3.
 * in = {hL,hR,huL,huR,hvL,hvR,bL,bR,uL,uR,vL,vR,phiL,phiR,pL,pR,sE1,sE2}
 * sw[3], fw[3*3] (fw(m,mw) at fw[(mw-1)*3 + m-1]). */
void gfo_riemann_aug_jcp(int maxiter, const double in[18], double drytol, double g,
                         double sw[3], double fw[9]);

/* rpn2 (rpn2_geoclaw.f:2-386).  q1d is modified in place exactly as the
 * reference modifies its ql/qr alias (negative depths zeroed). */
void gfo_rpn2(const gfo_params *p, int ixy, int mx, int mbc, int maux,
              double *q1d, const double *aux2,
              double *fwave, double *s, double *amdq, double *apdq);

/* rpt2 (rpt2_geoclaw.f:2-190). */
void gfo_rpt2(const gfo_params *p, int ixy, int imp, int mx, int mbc, int maux,
              const double *q1d, const double *aux1, const double *aux2,
              const double *aux3, const double *asdq,
              double *bmasdq, double *bpasdq);

/* philim (philim.f:65-126) and fc2d_geoclaw_limiter (geoclaw_limiter_fort.f:4-61). */
double gfo_philim(double a, double b, int meth);
void gfo_limiter(int mx, int mbc, double *wave, const double *s, const int mthlim[3]);

/* fc2d_geoclaw_flux2 (geoclaw_flux2fw_fort.f:4-269).  dtdx1d is modified in
 * place as in the reference. */
void gfo_flux2(const gfo_params *p, int ixy, int mx, int mbc, int maux,
               double *q1d, double *dtdx1d,
               const double *aux1, const double *aux2, const double *aux3,
               double *faddm, double *faddp, double *gaddm, double *gaddp,
               double *cfl1d);

/* fc2d_geoclaw_step2 (geoclaw_step2_fort.f90:1-294) followed by the update in
 * fc2d_geoclaw_step2_wrap (geoclaw_step2_wrap_fort.f:1-117).  qold is updated
 * in place (ghost cells included, as the reference does); returns cflgrid. */
double gfo_step2_wrap(const gfo_params *p, int mx, int my, int mbc, int maux,
                      double *qold, double *aux, double dx, double dy, double dt);

/* fc2d_geoclaw_b4step2 (geoclaw_b4step2_fort.f90:153-156), all cells. */
void gfo_b4step2(const gfo_params *p, int mx, int my, int mbc, double *q);

/* fc2d_geoclaw_src2 friction part (geoclaw_src2_fort.f90:41-71), interior. */
void gfo_src2(const gfo_params *p, int mx, int my, int mbc, int maux,
              double *q, const double *aux, double dt);

/* fc2d_geoclaw_bc2 (geoflood_bc2_fort.f90) for mthbc in {-1 (none), 1
 * (extrapolation), 3 (solid wall)}; sides left, right, bottom, top. */
void gfo_bc2(int mx, int my, int mbc, int maux, double *q, double *aux,
             const int mthbc[4]);

/* Hydrograph inflow (hydrograph/hydrograh_module.f90) as used by bc2's
 * mthbc = 0 branch.  Only hydrograph_type 'discharge' is restated: the 'eta'
 * type leaves q0(2), which bc2 then reads, undefined in the reference.
 * hu[k] is the discharge per unit width at time[k]; b, x0, y0 are the
 * channel width and centre; side[k] != 0 enables the inflow on bc2 side k
 * (left, right, bottom, top), i.e. boundary_location(1) == 'left',
 * (2) == 'right', (4) == 'bottom', (3) == 'top' in the reference. */
typedef struct gfo_hydrograph {
    int nrows;
    const double *time;
    const double *hu;
    double b, x0, y0;
    int side[4];
} gfo_hydrograph;

/* interpolation (hydrograh_module.f90:170-210): the discharge at time t. */
double gfo_hydro_interpolate(const gfo_hydrograph *h, double t);

/* Riemann_invariants and two_shock (hydrograh_module.f90:213-331),
 * discharge form: q0 = {h0, hu0} (h0 is solved for), q1 = {h1, hu1}. */
void gfo_riemann_invariants(double q0[2], const double q1[2], double grav);
void gfo_two_shock(double q0[2], const double q1[2], double grav);

/* fc2d_geoclaw_bc2 (geoflood_bc2_fort.f90:7-499) for mthbc in {-1, 0
 * (hydrograph inflow), 1, 3}; sides processed left, right, bottom, top.
 * Returns -1 (the reference stops) when a side has mthbc = 0 but h is NULL
 * or does not enable that side, else 0. */
int gfo_bc2_ex(const gfo_params *p, int mx, int my, int mbc, int maux, double *q,
               double *aux, const int mthbc[4], double xlower, double ylower,
               double dx, double dy, double t, const gfo_hydrograph *h);

/* Refinement criteria (geolib_source/refinement_data_module.f90, the
 * geoclaw_module sea level and the regions common block of
 * fclaw2d_source/regions.i).  region[8*m .. 8*m+7] = xlow, ylow, xhi, yhi,
 * tlow, thi, minlevel, maxlevel of region m. */
typedef struct gfo_refine {
    double wave_tolerance;
    double deep_depth;
    int max_level_deep;
    double max_velocity_depth_product;
    int num_speeds;
    const double *speed_tolerance;
    int num_flowgrades;
    const double *flowgradevalue;
    const int *iflowgradevariable; /* 1 depth, 2 momentum, 3 surface elevation */
    const int *iflowgrademinlevel;
    double sea_level;
    int mregions;
    const double *region;
} gfo_refine;

/* fc2d_geoclaw_flag2refine (fclaw2d_source/fc2d_geoflood_flag2refine.f90:
 * 7-130) at cell (i, j): 1 refine, 0 not, -1 inconclusive. */
int gfo_flag2refine(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc, int maux,
                    const double *q, const double *aux, double x1, double y1, double x2,
                    double y2, double t, int level, int maxlevel, int init_flag,
                    int is_coarsening, int i, int j);

/* fc2d_geoclaw_fort_tag4refinement (fc2d_geoclaw_tag4refinement_fort.f90:
 * 1-59): 1 when any interior cell flags the patch. */
int gfo_tag4refinement(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc,
                       int maux, double xlower, double ylower, double dx, double dy, double t,
                       const double *q, const double *aux, int level, int maxlevel,
                       int init_flag);

/* fc2d_geoclaw_fort_tag4coarsening (fc2d_geoclaw_tag4coarsening_fort.f90:
 * 1-110): 1 when no cell of the four siblings q[0..3] flags (coarsening
 * thresholds), else 0. */
int gfo_tag4coarsening(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc,
                       int maux, const double xlower[4], const double ylower[4], double dx,
                       double dy, double t, const double *const q[4],
                       const double *const aux[4], int level, int maxlevel, int init_flag);

/* geoclaw_update (fc2d_cpucuda.cpp:362-406): b4step2, step2 (+update), the
 * per-patch MAXCFL_CAP (1000 -> 0), then src2 when src_term > 0. */
double gfo_patch_update(const gfo_params *p, int mx, int my, int mbc, int maux,
                        double *q, double *aux, double dx, double dy,
                        double t, double dt);

/* ---------------------------------------------------------------------------
 * Uniform single-level brick of mi x mj patches (the forest of the uniform
 * configurations).  Patch (pi, pj) has index pj*mi + pi and its own AoS q and
 * aux arrays with ghosts.  The host-side ghost update follows ForestClaw's
 * order: same-level face copies, same-level corner copies, then bc2 on the
 * patch sides that lie on the physical boundary (mthbc = -1 elsewhere).
 * ------------------------------------------------------------------------- */
typedef double (*gfo_update_fn)(const void *ctx, int mx, int my, int mbc, int maux,
                                double *q, double *aux, double dx, double dy,
                                double t, double dt);

void gfo_brick_ghost_fill(int mi, int mj, int mx, int my, int mbc, int maux,
                          double **q, double **aux, const int mthbc[4]);

/* One step attempt over all patches (ghost fill, then every patch's update
 * through fn(ctx, ...) or the C restatement when fn is NULL).  Returns the
 * maximum of the per-patch capped CFL numbers.  nthreads > 1 uses OpenMP. */
double gfo_brick_advance(const gfo_params *p, int mi, int mj, int mx, int my,
                         int mbc, int maux, double **q, double **aux,
                         const int mthbc[4], double dx, double dy, double t,
                         double dt, gfo_update_fn fn, const void *ctx,
                         int nthreads);

/* outstyle_3 time loop (fc2d_geoclaw_run.c:417-590) over nsteps accepted
 * steps with variable dt: retake when maxcfl > max_cfl, next dt =
 * dt*desired_cfl/maxcfl.  Records per accepted step: dt used, maxcfl, the
 * interior water volume sum(h)*dx*dy after the step.  Returns the number of
 * retakes.  Arrays may be NULL. */
int gfo_brick_run(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                  int maux, double **q, double **aux, const int mthbc[4],
                  double dx, double dy, double t0, double initial_dt,
                  double desired_cfl, double max_cfl, int nsteps,
                  gfo_update_fn fn, const void *ctx, int nthreads,
                  double *dt_out, double *cfl_out, double *mass_out);

/* The same driver with the brick's lower-left corner (xlo, ylo), so that bc2
 * sees each patch's xlower/ylower, and an optional hydrograph for sides with
 * mthbc = 0.  Return -1 on a bc2 error (see gfo_bc2_ex). */
int gfo_brick_ghost_fill_ex(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                            int maux, double **q, double **aux, const int mthbc[4],
                            double xlo, double ylo, double dx, double dy, double t,
                            const gfo_hydrograph *h);

/* How the step's two powers x**(7/3) and x**(2/3) are evaluated: 0 (the
 * default) the host libm's pow, as the reference build does; 1 the
 * correctly rounded value (see gf_oracle.c).  gfo_pow(x, y) is that
 * evaluation (other exponents: pow). */
void gfo_set_pow_mode(int correctly_rounded);
double gfo_pow(double x, double y);

int gfo_brick_run_ex(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                     int maux, double **q, double **aux, const int mthbc[4],
                     double xlo, double ylo, double dx, double dy,
                     const gfo_hydrograph *h, double t0, double initial_dt,
                     double desired_cfl, double max_cfl, int nsteps,
                     gfo_update_fn fn, const void *ctx, int nthreads,
                     double *dt_out, double *cfl_out, double *mass_out);

/* gfo_brick_run_ex with the library's dt bound (gf_pool_set_scenario_dt_max;
 * no counterpart in the reference): with dt_max > 0 the next dt after an
 * accepted step is dt * desired_cfl / maxcfl capped at dt_max, and dt_max
 * itself when maxcfl = 0; dt_max = 0 is the reference's rule unchanged. */
int gfo_brick_run_dtmax(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                        int maux, double **q, double **aux, const int mthbc[4],
                        double xlo, double ylo, double dx, double dy,
                        const gfo_hydrograph *h, double t0, double initial_dt,
                        double desired_cfl, double max_cfl, double dt_max, int nsteps,
                        gfo_update_fn fn, const void *ctx, int nthreads,
                        double *dt_out, double *cfl_out, double *mass_out);

/* ---- ASCII frame output and gauges ---------------------------------------
 * fc2d_geoclaw_output_ascii_fort.f and fc2d_geoclaw_gauges_default.c.  The
 * writers append bytes to out (capacity cap) and return the number written,
 * or -1 when cap is too small. */

/* Fortran Ew.d edit descriptor (Flang): [-]0.<d digits>E+xx, or +xxx for a
 * three-digit exponent, right-justified in w characters; Inf / -Inf / NaN
 * right-justified.  Writes exactly w characters. */
void gfo_format_E(char *out, double v, int w, int d);

/* fort.tNNNN contents (fc2d_geoclaw_fort_write_header). */
int gfo_frame_header(char *out, int cap, double time, int meqn, int maux, int ngrids, int mbc);

/* One patch's block of fort.qNNNN (fc2d_geoclaw_fort_write_file); q and aux
 * are the AoS patch arrays with ghosts, and |q| < 1e-99 is set to 0 in q
 * itself, as the reference does. */
int gfo_frame_patch(char *out, int cap, int mx, int my, int meqn, int maux, int mbc,
                    double xlower, double ylower, double dx, double dy, double *q,
                    const double *aux, int patch_num, int level, int blockno, int mpirank);

/* fc2d_geoclaw_update_gauge (geoclaw_gauges_fort.f90): q and aux(1) at
 * (xc, yc) of the patch. */
void gfo_update_gauge(const gfo_params *p, int mx, int my, int mbc, int maux, double xlower,
                      double ylower, double dx, double dy, const double *q, const double *aux,
                      double xc, double yc, double qvar[3], double avar[1]);

/* Gauge file header lines and one record line
 * (geoclaw_create_gauge_files_default, geoclaw_print_gauges_default). */
int gfo_gauge_header(char *out, int cap, int id, double xc, double yc);
int gfo_gauge_line(char *out, int cap, int level, double t, const double qvar[3], const double avar[1]);

/* ---- Regrid and time interpolation between levels ------------------------
 * Patches in the reference's AoS layout with ghosts, refinement ratio 2, four
 * children numbered igrid = 0..3 (x fastest). */

/* fc2d_geoclaw_fort_average2coarse (fc2d_geoclaw_average_fort.f:482-603),
 * mcapa = 0: the quadrant igrid of qcoarse's interior from the fine patch. */
void gfo_average2coarse(const gfo_params *p, int mx, int my, int mbc, int maux, double *qcoarse,
                        const double *qfine, const double *aux_coarse, const double *aux_fine, int igrid);

/* fc2d_geoclaw_fort_interpolate2fine (fc2d_geoclaw_interpolate_fort.f:405-579)
 * with ForestClaw's fclaw2d_clawpatch_compute_slopes at mth = 5 (not in the
 * reference tree; restated, parity unpinned): the fine patch igrid's interior
 * from the coarse quadrant and its neighbours, sea_level = 0. */
void gfo_interpolate2fine(const gfo_params *p, int mx, int my, int mbc, int maux, const double *qcoarse,
                          double *qfine, const double *aux_coarse, const double *aux_fine, int igrid);

/* Coarse/fine ghost transfers between a coarse patch and a half-size
 * neighbour of the same orientation: side 0..3 = across face 0..3 of the
 * coarse patch (left, right, bottom, top), igrid = the fine patch's half
 * along that face (0: lower); side 4..7 = across corner 0..3 (lower-left,
 * lower-right, upper-left, upper-right).  The index map is restated from the
 * geometry (ForestClaw's transform_face_half / _corner_half are not in the
 * reference tree).  Return -1 for a bad side / igrid (or mx != my for the
 * interpolation).
 *
 * gfo_average_ghost: fc2d_geoclaw_fort_average_face / _average_corner
 * (fc2d_geoclaw_average_fort.f:33-480), mcapa = 0: qcoarse's ghost cells
 * from the fine interior.
 * gfo_interpolate_ghost: fc2d_geoclaw_fort_interpolate_face / _corner
 * (fc2d_geoclaw_interpolate_fort.f:24-403): qfine's ghost cells from the
 * coarse patch (its interior and stored ghost cells). */
int gfo_average_ghost(const gfo_params *p, int mx, int my, int mbc, int maux, double *qcoarse,
                      const double *qfine, const double *aux_coarse, const double *aux_fine, int side, int igrid);
int gfo_interpolate_ghost(const gfo_params *p, int mx, int my, int mbc, int maux, const double *qcoarse,
                          double *qfine, const double *aux_coarse, const double *aux_fine, int side, int igrid);

/* ForestClaw's slope for mth = 5 (the AMRClaw slope), as used above. */
double gfo_compute_slopes(double sl, double sr, int mth);

/* fc2d_geoclaw_fort_timeinterp (fc2d_geoclaw_timeinterp_fort.f:1-69):
 * qinterp = qlast + alpha (qcurr - qlast) on the two interior layers next to
 * each face.  Returns ierror (2 when the count differs from psize). */
int gfo_timeinterp(int mx, int my, int mbc, int meqn, int psize, const double *qcurr,
                   const double *qlast, double *qinterp, double alpha);

#ifdef __cplusplus
}
#endif
#endif
