// gf_kernels.h -- data structures shared by the HIP kernels and the host
// runtime of the GeoFlood patch step.
//
// HBM layout of a patch pool (all FP64):
//   q buffer  : [patch][m = h,hu,hv][row 0..my+3][col 0..mx+3]   (SoA per patch,
//               2 ghost layers; two buffers, current and next)
//   aux       : [patch][row][col]                                 (bathymetry b)
// Patch k occupies 3*(mx+4)*(my+4) doubles of a q buffer (31,104 B at 32x32).
#pragma once

#include <hip/hip_runtime.h>

#include "gf_physics.h"

namespace gf {

constexpr int MBC = 2;

// Neighbour codes of PatchDesc::nbr.  A value >= 0 is the pool index of the
// same-level neighbour patch whose interior supplies the ghost cells.
constexpr int NB_WALL = -1;    // solid wall (bc2 mthbc = 3)
constexpr int NB_EXTRAP = -2;  // zero-order extrapolation (bc2 mthbc = 1)
constexpr int NB_STORED = -3;  // ghost cells are supplied by the host
constexpr int NB_INFLOW = -4;  // hydrograph inflow (bc2 mthbc = 0); ghosts written by k_inflow

// nbr[] order: left, right, bottom, top, left-bottom, right-bottom, left-top, right-top
struct PatchDesc {
    double dx, dy;
    double xlo, ylo;  // lower-left corner of the interior (bc2's xlower, ylower)
    int scen;
    int nbr[8];
    int pad;
};

// Per-scenario time-step control record, resident on the device
// (the outstyle_3 rule of fc2d_geoclaw_run.c:417-590).
struct ScenCtrl {
    double t;            // simulated time of the accepted state
    double dt;           // dt of the next attempt
    double maxcfl;       // max CFL of the last attempt
    double desired_cfl;
    double max_cfl;
    double t_end;        // stop once t >= t_end (<= 0: no limit)
    double last_dt;      // dt of the last accepted step
    double manning_n;    // Manning coefficient of this scenario's src2
    const double* hyd;   // hydrograph table: hyd[0..rows) times, hyd[rows..2 rows) discharge per unit width
    double hyd_b, hyd_x0, hyd_y0;  // channel width and centre of the inflow strip
    double dt_max;       // upper bound of the next dt (<= 0: none)
    int hyd_rows;        // 0: no hydrograph
    int status;          // SCEN_OK, or why the scenario stopped (SCEN_NONFINITE)
    int cur;             // index of the q buffer holding the accepted state
    int nsteps;          // accepted steps
    int target_steps;    // stop after this many accepted steps (<= 0: no limit)
    int retakes;
    int active;          // 1 while the scenario is stepping
    int first_patch;
    int npatch;
    int attempts;
};

// ScenCtrl::status (the values of gf_scenario_info::status)
constexpr int SCEN_OK = 0;
constexpr int SCEN_NONFINITE = 1;  // a NaN/Inf state or CFL number: stopped, no step accepted

// dt of the attempt after an accepted step with CFL number m (the outstyle_3
// rule dt * desired_cfl / m, fc2d_geoclaw_run.c:577-580), bounded by dt_max
// when that is > 0.  A zero CFL number (every face dry on both sides, so no
// wave anywhere) keeps dt, or takes dt_max when one is set, where the
// reference's rule would give an infinite dt.
__host__ __device__ inline double next_dt(double dt, double desired_cfl, double m, double dt_max)
{
    double d = m > 0.0 ? dt * desired_cfl / m : dt;
    if (dt_max > 0.0 && (m == 0.0 || d > dt_max)) d = dt_max;
    return d;
}

// The CFL number is finite (not NaN, not +-Inf).
__host__ __device__ inline bool cfl_finite(double m) { return m - m == 0.0; }

}  // namespace gf
