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
    int hyd_rows;        // 0: no hydrograph
    int hyd_pad;
    int cur;             // index of the q buffer holding the accepted state
    int nsteps;          // accepted steps
    int target_steps;    // stop after this many accepted steps (<= 0: no limit)
    int retakes;
    int active;          // 1 while the scenario is stepping
    int first_patch;
    int npatch;
    int attempts;
};

}  // namespace gf
